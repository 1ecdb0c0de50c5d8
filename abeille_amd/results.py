"""Results and source files in the reference's layout, as .npz archives
(h5py is not available here; the member names are the reference's HDF5
paths, so a converter is a loop over keys).

Tallies::write_tallies (src/tallies.cpp:205-266):
  results/generations            recorded (active) generations
  results/kcol                   k_col per generation        k-eigenvalue only
  results/ktrk                   k_trk per generation        k-eigenvalue, surface tracking
  results/kcol-avg, kcol-std     mean, error of the mean     k-eigenvalue, generations > 0
  results/ktrk-avg, ktrk-std                                 ... and surface tracking
  results/leakage, mig-area      per generation              every mode
  results/leakage-avg, leakage-str (the reference's own spelling),
  results/mig-area-avg, mig-area-std                         generations > 0
MeshTally::write_tally (src/mesh_tally.cpp:154-207), generations > 0:
  results/<tally>/avg, std       [Ne][Nx][Ny][Nz] mean, error of the mean
  results/<tally>/{x,y,z}-bounds, energy-bounds, quantity, estimator
Simulation::write_source (src/simulation.cpp:137-175):
  source                         [N][9] x y z ux uy uz E w w2

In multigroup mode the reference writes no kabs vectors (k_abs equals
k_col there), and neither does this module.
"""
import numpy as np

from . import abi

_QUANTITY = {abi.Q_FLUX: "flux", abi.Q_ELASTIC: "elastic", abi.Q_ABSORPTION: "absorption",
             abi.Q_FISSION: "fission", abi.Q_TOTAL: "total"}
_ESTIMATOR = {abi.EST_COLLISION: "collision", abi.EST_TRACK_LENGTH: "track-length"}


def gen_values_of(problem):
    """The per-generation vectors of a problem's run, keyed by the
    reference's dataset names (kcol, ktrk, leakage, mig-area)."""
    return {"kcol": problem.gen_values(abi.GEN_KCOL),
            "ktrk": problem.gen_values(abi.GEN_KTRK),
            "leakage": problem.gen_values(abi.GEN_LEAKAGE),
            "mig-area": problem.gen_values(abi.GEN_MIG)}


def results_dict(model, stats, tallies, gen_values=None, names=None):
    """Flat {path: array} of a finished run.  `tallies` is [(avg, std)] per
    mesh tally (Problem.tally_stats), `gen_values` the per-generation
    vectors ({"kcol", "ktrk", "leakage", "mig-area"}, see gen_values_of)."""
    gen = int(stats.active_generations)
    gv = gen_values or {}
    kmode = model.settings["mode"] == abi.K_EIGENVALUE
    surface = model.settings["tracking"] == abi.SURFACE_TRACKING
    out = {"results/generations": np.int64(gen)}
    if kmode:
        if "kcol" in gv:
            out["results/kcol"] = np.asarray(gv["kcol"], np.float64)
        if surface and "ktrk" in gv:
            out["results/ktrk"] = np.asarray(gv["ktrk"], np.float64)
        if gen > 0:
            out["results/kcol-avg"] = np.float64(stats.kcol_avg)
            out["results/kcol-std"] = np.float64(stats.kcol_err)
            if surface:
                out["results/ktrk-avg"] = np.float64(stats.ktrk_avg)
                out["results/ktrk-std"] = np.float64(stats.ktrk_err)
    if "leakage" in gv:
        out["results/leakage"] = np.asarray(gv["leakage"], np.float64)
    if "mig-area" in gv:
        out["results/mig-area"] = np.asarray(gv["mig-area"], np.float64)
    if gen > 0:
        out["results/leakage-avg"] = np.float64(stats.leak_avg)
        out["results/leakage-str"] = np.float64(stats.leak_err)
        out["results/mig-area-avg"] = np.float64(stats.mig_avg)
        out["results/mig-area-std"] = np.float64(stats.mig_err)
        for t, (spec, (avg, std)) in enumerate(zip(model.tallies, tallies)):
            name = names[t] if names else f"tally{t}"
            nx, ny, nz = spec["shape"]
            eb = np.asarray(spec["ebounds"], np.float64)
            g = f"results/{name}/"
            out[g + "avg"] = np.asarray(avg).reshape(len(eb) - 1, nx, ny, nz)
            out[g + "std"] = np.asarray(std).reshape(len(eb) - 1, nx, ny, nz)
            for k, ax in enumerate("xyz"):
                n = spec["shape"][k]
                d = (spec["high"][k] - spec["low"][k]) / n
                out[g + f"{ax}-bounds"] = np.arange(n + 1) * d + spec["low"][k]
            out[g + "energy-bounds"] = eb
            out[g + "quantity"] = np.array(_QUANTITY[spec["quantity"]])
            out[g + "estimator"] = np.array(_ESTIMATOR[spec["estimator"]])
    return out


def write_results(path, model, stats, tallies, gen_values=None, names=None):
    np.savez(path, **results_dict(model, stats, tallies, gen_values, names))


def read_results(path):
    with np.load(path, allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def write_source(path, bank9):
    """The reference's source file: one [N][9] array named `source`."""
    b = np.ascontiguousarray(bank9, np.float64)
    if b.ndim != 2 or b.shape[1] != 9:
        raise ValueError("a source bank is [N][9]")
    np.savez(path, source=b)


def read_source(path):
    with np.load(path, allow_pickle=False) as f:
        b = f["source"]
    if b.ndim != 2 or b.shape[1] != 9:
        raise ValueError("Invalid source from file dimensions.")
    return b

"""Results and source files in the reference's layout, as .npz archives
(h5py is not available here; the member names are the reference's HDF5
paths, so a converter is a loop over keys).

  results/generations            active generations         (Tallies::write_tallies,
  results/kcol, kcol-avg/-std    k_col per generation, mean   src/tallies.cpp:205-266)
  results/leakage-avg/-std, mig-area-avg/-std
  results/<tally>/avg, std       [Ne][Nx][Ny][Nz] mean, error of the mean
  results/<tally>/{x,y,z}-bounds, energy-bounds, quantity, estimator
                                                            (MeshTally::write_tally,
                                                             src/mesh_tally.cpp:154-207)
  source                         [N][9] x y z ux uy uz E w w2 (src/simulation.cpp:137-175)
"""
import numpy as np

from . import abi

_QUANTITY = {abi.Q_FLUX: "flux", abi.Q_ELASTIC: "elastic", abi.Q_ABSORPTION: "absorption",
             abi.Q_FISSION: "fission", abi.Q_TOTAL: "total"}
_ESTIMATOR = {abi.EST_COLLISION: "collision", abi.EST_TRACK_LENGTH: "track-length"}


def results_dict(model, stats, tallies, kcol=None, names=None):
    """Flat {path: array} of a finished run.  `tallies` is [(avg, std)] per
    mesh tally (Problem.tally_stats), `kcol` the per-generation k_col."""
    out = {"results/generations": np.int64(stats.active_generations)}
    if kcol is not None:
        out["results/kcol"] = np.asarray(kcol, np.float64)
    if model.settings["mode"] == abi.K_EIGENVALUE and stats.active_generations > 0:
        out["results/kcol-avg"] = np.float64(stats.kcol_avg)
        out["results/kcol-std"] = np.float64(stats.kcol_err)
        if model.settings["tracking"] == abi.SURFACE_TRACKING:
            out["results/ktrk-avg"] = np.float64(stats.ktrk_avg)
            out["results/ktrk-std"] = np.float64(stats.ktrk_err)
    out["results/leakage-avg"] = np.float64(stats.leak_avg)
    out["results/leakage-std"] = np.float64(stats.leak_err)
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


def write_results(path, model, stats, tallies, kcol=None, names=None):
    np.savez(path, **results_dict(model, stats, tallies, kcol, names))


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

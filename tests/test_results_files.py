"""Results and source files (abeille_amd.results): the reference's layout
(results/<tally>/{avg,std,x-bounds,...}, results/kcol..., source [N][9])
round-trips, with bounds computed like MeshTally::write_tally."""
import numpy as np
import pytest

from abeille_amd import abi, problems
from abeille_amd.results import read_results, read_source, results_dict, write_results, \
    write_source


class _Stats:
    active_generations = 3
    kcol_avg, kcol_err = 1.01, 0.002
    ktrk_avg, ktrk_err = 1.0, 0.003
    leak_avg, leak_err = 0.1, 0.01
    mig_avg, mig_err = 50.0, 1.0


_GV = {"kcol": [1.0, 1.02, 1.01], "ktrk": [0.99, 1.0, 1.01], "leakage": [0.1, 0.11, 0.09],
       "mig-area": [49.0, 50.0, 51.0]}


def test_results_layout_round_trip(tmp_path):
    m = problems.pin_lattice_2d(100)
    nb = 7 * 51 * 51
    avg = np.arange(nb, dtype=float)
    std = np.sqrt(avg)
    d = results_dict(m, _Stats(), [(avg, std)], gen_values=_GV, names=["flux"])
    assert d["results/flux/avg"].shape == (7, 51, 51, 1)
    xb = d["results/flux/x-bounds"]
    assert len(xb) == 52 and xb[0] == pytest.approx(-10.71) and xb[-1] == pytest.approx(10.71)
    assert str(d["results/flux/estimator"]) == "track-length"
    assert "results/kcol-avg" in d and "results/ktrk-avg" not in d  # delta tracking
    p = tmp_path / "r.npz"
    write_results(p, m, _Stats(), [(avg, std)], gen_values=_GV, names=["flux"])
    back = read_results(p)
    np.testing.assert_array_equal(back["results/flux/std"].ravel(), std)
    assert int(back["results/generations"]) == 3


def test_results_names_follow_write_tallies():
    """Tallies::write_tallies (src/tallies.cpp:205-266): per-generation
    kcol (k-eigenvalue), ktrk (surface tracking), leakage and mig-area
    vectors; leakage-str (the reference's spelling); averages and mesh
    tallies only once a generation is recorded."""
    st = _Stats()
    m = problems.pin_lattice_2d(100, tracking="surface")
    tl = [(np.zeros(7 * 51 * 51), np.zeros(7 * 51 * 51))]
    d = results_dict(m, st, tl, gen_values=_GV, names=["flux"])
    for k in ("kcol", "ktrk", "leakage", "mig-area"):
        np.testing.assert_array_equal(d[f"results/{k}"], _GV[k])
    for k in ("kcol-avg", "kcol-std", "ktrk-avg", "ktrk-std", "leakage-avg", "leakage-str",
              "mig-area-avg", "mig-area-std"):
        assert f"results/{k}" in d, k
    assert "results/leakage-std" not in d
    assert d["results/leakage-str"] == st.leak_err
    # no recorded generation: vectors only
    st.active_generations = 0
    d0 = results_dict(m, st, tl, gen_values=_GV, names=["flux"])
    assert "results/kcol" in d0 and "results/leakage" in d0
    assert not any(k.endswith(("-avg", "-std", "-str")) for k in d0)
    assert "results/flux/avg" not in d0
    # fixed source: no k vectors, leakage and mig-area still written
    mf = problems.fixed_source_cube(100)
    st.active_generations = 3
    df = results_dict(mf, st, [(np.zeros(16000), np.zeros(16000))], gen_values=_GV)
    assert "results/kcol" not in df and "results/kcol-avg" not in df
    assert "results/leakage" in df and "results/leakage-str" in df


def test_source_file_round_trip(tmp_path):
    b = np.random.default_rng(1).random((100, 9))
    write_source(tmp_path / "s.npz", b)
    np.testing.assert_array_equal(read_source(tmp_path / "s.npz"), b)
    with pytest.raises(ValueError):
        write_source(tmp_path / "bad.npz", b[:, :8])


@pytest.mark.gpu
def test_bank_dump_and_resume():
    """A bank fetched after some generations seeds a fresh problem
    (PowerIterator::load_source_from_file): the loaded first generation
    transports exactly that bank and k stays consistent."""
    from abeille_amd import Problem
    m, k_ref = problems.la13511("PUa-1-0-SL", 20000)
    a = Problem(m)
    a.run(20, 10)
    bank = a.bank()
    b = Problem(m)
    b.load_bank(bank)
    back = b.bank()
    np.testing.assert_allclose(back[:, :3], bank[:, :3])
    np.testing.assert_allclose(back[:, 6:], bank[:, 6:])
    st, kh = b.run_continue(30, 0)
    assert abs(st.kcol_avg - k_ref) < 4 * st.kcol_err + 1e-3


@pytest.mark.gpu
def test_per_generation_vectors_and_statistics():
    """The per-generation vectors (abx_gen_values) hold every generation,
    inactive ones included, and the running statistics equal numpy's over the
    recorded ones: mean, and sqrt(sample variance / g) as
    Tallies::update_avg_and_var and kcol_err() give (src/tallies.cpp:268-280,
    include/simulation/tallies.hpp:103-151)."""
    from abeille_amd import Problem
    from abeille_amd.results import gen_values_of
    m = problems.pin_lattice_2d(20000, tracking="surface")
    p = Problem(m)
    st, kh = p.run(12, 4)
    gv = gen_values_of(p)
    assert len(gv["kcol"]) == 12
    np.testing.assert_array_equal(gv["kcol"], kh)
    for name, avg, err in (("kcol", st.kcol_avg, st.kcol_err), ("ktrk", st.ktrk_avg, st.ktrk_err),
                           ("leakage", st.leak_avg, st.leak_err),
                           ("mig-area", st.mig_avg, st.mig_err)):
        x = np.asarray(gv[name][4:])
        assert avg == pytest.approx(x.mean(), rel=1e-12, abs=1e-15), name
        assert err == pytest.approx(np.sqrt(x.var(ddof=1) / len(x)), rel=1e-9, abs=1e-15), name
    kabs = p.gen_values(abi.GEN_KABS)[4:]
    assert st.kabs_avg == pytest.approx(kabs.mean(), rel=1e-12)
    assert st.kabs_err == pytest.approx(np.sqrt(kabs.var(ddof=1) / len(kabs)), rel=1e-9)
    # multigroup: k_abs scores nu*fission/total per collision, as k_col's
    # nu-fission/total does, so the two agree
    np.testing.assert_allclose(p.gen_values(abi.GEN_KABS), kh, rtol=1e-12)
    d = results_dict(m, st, [p.tally_stats(0)], gen_values=gv, names=["flux"])
    assert len(d["results/ktrk"]) == 12 and "results/leakage-str" in d

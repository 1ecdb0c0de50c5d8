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


def test_results_layout_round_trip(tmp_path):
    m = problems.pin_lattice_2d(100)
    nb = 7 * 51 * 51
    avg = np.arange(nb, dtype=float)
    std = np.sqrt(avg)
    d = results_dict(m, _Stats(), [(avg, std)], kcol=[1.0, 1.02, 1.01], names=["flux"])
    assert d["results/flux/avg"].shape == (7, 51, 51, 1)
    xb = d["results/flux/x-bounds"]
    assert len(xb) == 52 and xb[0] == pytest.approx(-10.71) and xb[-1] == pytest.approx(10.71)
    assert str(d["results/flux/estimator"]) == "track-length"
    assert "results/kcol-avg" in d and "results/ktrk-avg" not in d  # delta tracking
    p = tmp_path / "r.npz"
    write_results(p, m, _Stats(), [(avg, std)], kcol=[1.0, 1.02, 1.01], names=["flux"])
    back = read_results(p)
    np.testing.assert_array_equal(back["results/flux/std"].ravel(), std)
    assert int(back["results/generations"]) == 3


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

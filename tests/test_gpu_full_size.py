"""The configurations at their full sizes, checked through properties that
do not depend on size (the oracle is too slow there):

* 20 cm cube, 2 groups, 1e6 histories as 40 batches of 25,000: leakage plus
  absorption (Σa times the volume-integrated collision-estimator flux, the
  weight implicit capture deposits) is one source particle, up to the
  Russian-roulette noise;
* 17×17 lattice, one generation of 1e6 source particles, and the 3×3 core,
  2e6: the fission bank comes back in parent order, then push order
  (src/delta_tracker.cpp:251-253; the library itself checks that its size
  is the sum of the per-parent counts), and a second transport of the same
  bank gives the same bank bit for bit and the same event counts
  (per-history RNG streams), with tallies equal up to the summation order
  of the FP64 atomics.
"""
import numpy as np
import pytest

from abeille_amd import Problem, problems

pytestmark = pytest.mark.gpu


def test_cube_full_size_balance():
    m = problems.fixed_source_cube(25_000)
    p = Problem(m)
    st, _ = p.run(40, 0)
    assert st.histories == 1_000_000
    avg, _ = p.tally_stats(0)
    flux = avg.reshape(2, -1).sum(axis=1)          # [group] volume-integrated
    absorption = 0.05 * flux[0] + 0.20 * flux[1]
    print(f"[cube] leakage {st.leak_avg:.6f} absorption {absorption:.6f}", flush=True)
    assert 0.0 < st.leak_avg < 1.0
    assert abs(st.leak_avg + absorption - 1.0) < 3e-3


def _bank_order_and_repeat(m, n):
    p = Problem(m)
    bank = p.sample_sources(0, n)
    a = p.transport(bank, kcol=1.0, converged=True, cap=2 * n)
    ta = p.tally(0)
    p.tally_clear()
    b = p.transport(bank, kcol=1.0, converged=True, cap=2 * n)
    tb = p.tally(0)
    fa, fb = a["fission"], b["fission"]
    assert len(fa) > n // 10
    key = fa["parent_history_id"].astype(np.float64) * 2.0**20 + fa["parent_daughter_id"]
    assert np.all(np.diff(fa["parent_history_id"].astype(np.int64)) >= 0)
    assert np.all(np.diff(key) > 0), "sites out of (parent, daughter) order"
    assert fa.tobytes() == fb.tobytes()
    assert a["diag"].collisions == b["diag"].collisions
    assert a["diag"].flights == b["diag"].flights
    assert a["diag"].histories == n
    np.testing.assert_allclose(tb, ta, rtol=1e-12, atol=1e-18)
    return a


def test_lattice_full_size_bank_order_and_repeat():
    a = _bank_order_and_repeat(problems.pin_lattice_2d(1_000_000), 1_000_000)
    print(f"[lattice] {len(a['fission'])} sites, {a['diag'].collisions} collisions", flush=True)


def test_core_full_size_bank_order_and_repeat():
    a = _bank_order_and_repeat(problems.core_3x3(2_000_000, tracking="surface"), 2_000_000)
    print(f"[core] {len(a['fission'])} sites, {a['diag'].collisions} collisions", flush=True)

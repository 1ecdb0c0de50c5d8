"""Split precision (FP32 flight distances, cross-section ratios, directions
and score factors; FP64 positions, surface tests, weights and tally
accumulators) against the FP64 path on the same problems and seeds.

The two paths sample different random walks after the first FP32 rounding,
so they are compared statistically: k and every tally bin within their
combined standard errors, and analytic answers where the problem has one."""
import numpy as np
import pytest

from abeille_amd import Problem, abi, problems
from helpers import assert_within_sigma

pytestmark = pytest.mark.gpu


def _pair(make, ngen, nign, engine):
    out = []
    for prec in (abi.FP64, abi.SPLIT):
        m = make()
        m.set(precision=prec)
        p = Problem(m, engine=engine)
        st, _ = p.run(ngen, nign)
        tallies = [p.tally_stats(t) for t in range(len(m.tallies))]
        out.append((st, tallies))
    return out


def _z_scores(a64, e64, a32, e32):
    sig = np.hypot(e64, e32)
    ok = sig > 0
    return (a32[ok] - a64[ok]) / sig[ok]


@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_lattice_split_matches_fp64(engine):
    """17x17 lattice, delta tracking, track-length tally 51x51x7: k agrees
    within statistics and the per-bin differences, in units of their combined
    standard error, are distributed as noise (mean ~ 0, spread ~ 1): a bias
    of the FP32 path as small as a tenth of the per-bin error would shift the
    mean of 18207 z-scores by 0.1."""
    print(f"[split] lattice engine={engine}", flush=True)
    (s64, t64), (s32, t32) = _pair(lambda: problems.pin_lattice_2d(400_000), 14, 4, engine)
    sig = np.hypot(s64.kcol_err, s32.kcol_err)
    assert abs(s64.kcol_avg - s32.kcol_avg) < 4 * sig, (s64.kcol_avg, s32.kcol_avg, sig)
    (a64, e64), (a32, e32) = t64[0], t32[0]
    assert_within_sigma(a64, e64, a32, e32, nsig=4.0, min_frac=0.99)
    z = _z_scores(a64, e64, a32, e32)
    assert len(z) > 10000
    assert abs(z.mean()) < 0.1, z.mean()
    assert 0.8 < z.std() < 1.25, z.std()


def test_cube_split_matches_fp64():
    """Fixed-source cube, collision-estimator tally 20^3 x 2."""
    (s64, t64), (s32, t32) = _pair(lambda: problems.fixed_source_cube(20000, "delta"), 10, 0,
                                   abi.ENGINE_HISTORY)
    (a64, e64), (a32, e32) = t64[0], t32[0]
    assert_within_sigma(a64, e64, a32, e32, nsig=4.0, min_frac=0.99)
    assert abs(s64.leak_avg - s32.leak_avg) < 4 * np.hypot(s64.leak_err, s32.leak_err)


@pytest.mark.parametrize("tracking", ["surface", "delta"])
def test_split_critical_slab_k_is_one(tracking):
    """LA-13511 PUa-1-0-SL is exactly critical."""
    m, k_ref = problems.la13511("PUa-1-0-SL", 20000, tracking=tracking)
    m.set(precision=abi.SPLIT)
    st, _ = Problem(m).run(40, 10)
    assert abs(st.kcol_avg - k_ref) < 4 * st.kcol_err, (st.kcol_avg, st.kcol_err)


def test_split_infinite_medium_flux():
    """Reflective 2-group box: volume-integrated collision and track-length
    fluxes per source particle are 5.0 and 3.75 cm."""
    m = problems.infinite_medium_2g(20000, tracking="delta")
    m.set(precision=abi.SPLIT)
    p = Problem(m)
    p.run(10, 0)
    for t in range(2):
        avg, err = p.tally_stats(t)
        assert abs(avg[0] - 5.0) < 4 * err[0] + 1e-9, (t, avg, err)
        assert abs(avg[1] - 3.75) < 4 * err[1] + 1e-9, (t, avg, err)


def test_split_rejects_traces():
    m = problems.fixed_source_cube(100)
    m.set(precision=abi.SPLIT)
    p = Problem(m)
    bank = p.sample_sources(0, 10)
    with pytest.raises(Exception):
        p.transport(bank, trace_cap=8)


def test_split_flight_never_zero_length():
    """A history of the 17x17 lattice (seed + 2, generation 10 of a 1e6 run)
    that once drew a flight of zero length in FP32 (1 - xi rounds to 1.0f
    for xi < 2^-25), collided on the reflective plane it had just reflected
    from, and left the geometry (ABX_E_LOST).  The FP32 flight distance now
    takes -log(1 - xi) from its series for small xi, and the history ends
    normally."""
    m = problems.pin_lattice_2d(1_000_000)
    m.set(precision=abi.SPLIT, seed=abi.DEFAULT_SEED + 2)
    p = Problem(m)
    b = np.zeros(1, abi.PARTICLE_DTYPE)
    b[0]["r"] = [-1.3622847318790436, 1.5491719473775523, -9.360116872407609]
    b[0]["u"] = [0.826570451259613, -0.5488812327384949, -0.12454143911600113]
    b[0]["E"] = 1.5
    b[0]["wgt"] = 1.0003661340050458
    b[0]["history_id"] = 8701559
    b[0]["rng_state"] = 10619179219843491746
    res = p.transport(b, kcol=0.2942699356771855, converged=True)
    assert res["diag"].histories == 1 and res["diag"].lost == 0

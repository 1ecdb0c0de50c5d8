"""GPU parity: the HIP path (through the C ABI) against the CPU restatement
on the same seeded inputs.

Bit-exact: source particles, per-history event traces (event kind, cell and
material index, group, position, direction, weight, RNG draw count), the
ordered fission bank.  Scalar scores are sums whose order differs (atomics),
so they are compared to 1e-12 relative.
"""
import numpy as np
import pytest

from abeille_amd import Problem, abi, problems
from helpers import assert_traces_equal
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

CASES = {
    "cube_dt": lambda: problems.fixed_source_cube(2000, tracking="delta"),
    "cube_st": lambda: problems.fixed_source_cube(2000, tracking="surface"),
    "lattice_dt": lambda: problems.pin_lattice_2d(2000, tracking="delta"),
    "lattice_st": lambda: problems.pin_lattice_2d(2000, tracking="surface"),
    "core_st": lambda: problems.core_3x3(2000, tracking="surface"),
    "core_dt": lambda: problems.core_3x3(2000, tracking="delta"),
    "slab_PUa_1_0_SL": lambda: problems.la13511("PUa-1-0-SL", 2000)[0],
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_trace_bit_exact(name, engine):
    print(f"[parity] case={name} engine={engine}", flush=True)
    m = CASES[name]()
    n, cap = 2000, 512
    o = Oracle(m)
    g = Problem(m, engine=engine)
    so = o.sample_sources(0, n)
    sg = g.sample_sources(0, n)
    assert so.tobytes() == sg.tobytes(), "source particles differ"
    kcol = 0.9
    ro = o.transport(so, kcol=kcol, nthreads=1, trace_cap=cap)
    rg = g.transport(so, kcol=kcol, trace_cap=cap)
    assert_traces_equal(ro, rg, cap)
    assert ro["fission"].tobytes() == rg["fission"].tobytes(), "fission banks differ"
    for f in ("k_col", "k_abs", "k_trk", "leakage", "mig"):
        a, b = getattr(ro["scores"], f), getattr(rg["scores"], f)
        assert abs(a - b) <= 1e-12 * max(1.0, abs(a)), (f, a, b)


@pytest.mark.parametrize("name", ["cube_dt", "lattice_dt", "core_st"])
@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_tally_generation_matches(name, engine):
    """One batch: the GPU tally array equals the oracle's up to summation
    order (atomics), bin by bin to 1e-9 relative."""
    m = CASES[name]()
    o = Oracle(m)
    g = Problem(m, engine=engine)
    so = o.sample_sources(0, 2000)
    o.tally_clear()
    g.tally_clear()
    o.transport(so, kcol=0.9, nthreads=1)
    g.transport(so, kcol=0.9)
    to, tg = o.tally(0), g.tally(0)
    assert to.sum() > 0
    np.testing.assert_allclose(tg, to, rtol=1e-9, atol=1e-15)


def test_cancellation_bit_exact():
    """GPU approximate mesh cancellation equals the CPU restatement bit for
    bit (bins summed in bank order on both sides)."""
    from test_cancellation import make_bank
    m = problems.fixed_source_cube(100)
    m.cancellation_mesh([-10, -10, -10], [10, 10, 10], [4, 4, 4], energy_bounds=[0.0, 1.0, 2.0])
    o = Oracle(m)
    g = Problem(m)
    b = make_bank(200000, np.random.default_rng(11))
    a_o = o.cancel(b)
    a_g = g.cancel_bank(b)
    assert a_o.tobytes() == a_g.tobytes()
    assert not np.array_equal(a_o["wgt"], b["wgt"])

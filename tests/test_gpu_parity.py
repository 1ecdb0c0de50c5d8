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
    "slab_mfs_carter": lambda: problems.subcritical_slab(2000, tracking="carter", cancel=False),
    "slab_mfs_delta": lambda: problems.subcritical_slab(2000, tracking="delta", cancel=False),
    "slab_p2_scatter": lambda: problems.anisotropic_slab(2000, tracking="surface"),
    "cube_cylinder_st": lambda: problems.cylinder_cube(2000, tracking="surface"),
    # a general Cylinder off the coordinate axes: the reference's quadric
    # alpha x^2 + beta y^2 + gamma z^2 = R^2, on the extended-geometry kernels
    "cube_cylinder_oblique_st": lambda: problems.cylinder_cube(2000, tracking="surface",
                                                               axis=(1., 1., 2.)),
    "cube_cylinder_oblique_dt": lambda: problems.cylinder_cube(2000, tracking="delta",
                                                               axis=(-2., 1., 0.5)),
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_trace_bit_exact(name, engine):
    print(f"[parity] case={name} engine={engine}", flush=True)
    _trace_bit_exact(CASES[name](), engine)


@pytest.mark.parametrize("name", ["lattice_dt", "core_st", "slab_mfs_carter", "cube_st"])
def test_trace_bit_exact_event_sorted(name):
    """The event engine with its collision queue sorted by (material, group)
    before every collision launch, and no in-lane tail: the order in which
    particles are processed does not change any history."""
    print(f"[parity] case={name} event engine, sorted queue", flush=True)
    _trace_bit_exact(CASES[name](), abi.ENGINE_EVENT,
                     dict(tail_threshold=1, sort_every=1))


def _trace_bit_exact(m, engine, tuning=None):
    n, cap = 2000, 512
    o = Oracle(m)
    g = Problem(m, engine=engine)
    if tuning:
        g.set_tuning(**tuning)
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


@pytest.mark.parametrize("case", ["one_bin", "fine", "same_sign", "tiny", "ragged"])
def test_cancellation_bit_exact_shapes(case):
    """The wave-per-bin cancellation against the CPU restatement on bins
    that span many blocks of sorted positions (one bin for the whole mesh),
    bins of one site (a fine mesh), bins of one sign (nothing changes),
    banks smaller than a wave, and a ragged size."""
    print(f"[cancel] case={case}", flush=True)
    from test_cancellation import make_bank
    m = problems.fixed_source_cube(100)
    shape = {"one_bin": [1, 1, 1], "fine": [60, 60, 60]}.get(case, [3, 5, 2])
    m.cancellation_mesh([-10, -10, -10], [10, 10, 10], shape)
    o = Oracle(m)
    g = Problem(m)
    n = {"one_bin": 150000, "fine": 120000, "tiny": 37, "ragged": 70001}.get(case, 50000)
    b = make_bank(n, np.random.default_rng(23))
    if case == "same_sign":
        b["wgt"] = np.abs(b["wgt"])
        b["wgt2"] = -np.abs(b["wgt2"])
    for _ in range(2):  # the second call reuses the buffers of the first
        a_o = o.cancel(b)
        a_g = g.cancel_bank(b)
        assert a_o.tobytes() == a_g.tobytes()
    if case == "same_sign":
        assert a_g.tobytes() == b.tobytes()
    elif case != "tiny":
        assert not np.array_equal(a_g["wgt"], b["wgt"])


RUNS = {
    "lattice_dt": (lambda: problems.pin_lattice_2d(4000, tracking="delta"), 4, 1),
    "slab_PUa_1_0_SL": (lambda: problems.la13511("PUa-1-0-SL", 4000)[0], 4, 2),
    "cube_st": (lambda: problems.fixed_source_cube(4000, tracking="surface"), 3, 0),
    "slab_mfs_carter_cancel": (lambda: problems.subcritical_slab(4000, tracking="carter"), 3, 0),
}


@pytest.mark.parametrize("name", sorted(RUNS))
@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_run_matches_oracle(name, engine):
    """Whole device-resident runs (generation loop, weight normalisation,
    re-banking, cancellation, statistics) against the CPU restatement's
    run: per-generation k_col and every tally bin's mean and standard error
    agree up to summation order."""
    print(f"[run] case={name} engine={engine}", flush=True)
    make, ngen, nign = RUNS[name]
    m = make()
    o = Oracle(m)
    g = Problem(m, engine=engine)
    so, ko = o.run(ngen, nign, nthreads=1)
    sg, kg = g.run(ngen, nign)
    np.testing.assert_allclose(kg, ko, rtol=1e-9)
    assert sg.active_generations == so.active_generations
    assert sg.diag.collisions == so.diag.collisions
    for t in range(len(m.tallies)):
        ao, eo = o.tally_stats(t)
        ag, eg = g.tally_stats(t)
        np.testing.assert_allclose(ag, ao, rtol=1e-8, atol=1e-14)
        np.testing.assert_allclose(eg, eo, rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_run_continue_keeps_the_source_bank(engine):
    """run(1) followed by run_continue(3) transports the same generations as
    run(4): the k-eigenvalue source survives between calls."""
    m = problems.pin_lattice_2d(4000, tracking="delta", tally=False)
    a = Problem(m, engine=engine)
    _, k1 = a.run(1, nignored=1)
    _, k2 = a.run_continue(3, nignored=0)
    b = Problem(m, engine=engine)
    _, k = b.run(4, nignored=1)
    np.testing.assert_allclose(np.concatenate([k1, k2]), k, rtol=1e-9)


@pytest.mark.parametrize("name", ["lattice_dt", "cube_st", "slab_mfs_carter"])
@pytest.mark.parametrize("engine", [abi.ENGINE_EVENT, abi.ENGINE_HISTORY], ids=["event", "history"])
def test_empty_bank(name, engine):
    """An empty bank in every mode: no sites, no scores, no tally, no
    launch error; the next transport on the same problem still matches the
    CPU restatement bit for bit."""
    print(f"[parity] case={name} engine={engine} empty bank", flush=True)
    m = CASES[name]()
    o = Oracle(m)
    g = Problem(m, engine=engine)
    g.tally_clear()
    r = g.transport(g.sample_sources(0, 0), kcol=0.9)
    assert len(r["fission"]) == 0
    s = r["scores"]
    assert (s.k_col, s.k_abs, s.k_trk, s.leakage, s.mig) == (0.0,) * 5
    assert r["diag"].histories == 0 and r["diag"].collisions == 0
    assert not g.tally(0).any()
    so = o.sample_sources(0, 500)
    ro = o.transport(so, kcol=0.9, nthreads=1)
    rg = g.transport(so, kcol=0.9)
    assert ro["fission"].tobytes() == rg["fission"].tobytes()
    assert ro["diag"].collisions == rg["diag"].collisions


def test_fission_capacity_error_then_recovers():
    """A fission_out buffer smaller than the generation's sites is refused
    with ABX_E_CAPACITY (the library cannot grow the caller's buffer as the
    reference's vector grows); the problem stays usable and the next
    transport matches the CPU restatement bit for bit."""
    from abeille_amd import AbxError
    m = CASES["lattice_dt"]()
    o = Oracle(m)
    g = Problem(m)
    so = o.sample_sources(0, 500)
    with pytest.raises(AbxError, match="ABX_E_CAPACITY") as e:
        g.transport(so, kcol=0.9, cap=1)
    assert e.value.code == -5   # ABX_E_CAPACITY
    ro = o.transport(so, kcol=0.9, nthreads=1)
    assert len(ro["fission"]) > 1
    # the refused call's ordered sites are still on the device
    fetched = g.fission_fetch(len(ro["fission"]))
    assert ro["fission"].tobytes() == fetched.tobytes()
    rg = g.transport(so, kcol=0.9)
    assert ro["fission"].tobytes() == rg["fission"].tobytes()

"""Exact regional cancellation: ExactMGCancelator (src/exact_mg_cancelator.cpp)
and BasicExactMGCancelator (src/basic_exact_mg_cancelator.cpp), applied to a
k-eigenvalue generation's fission sites by PowerIterator
(src/power_iterator.cpp:354-356, 751-777).

The CPU restatement is pinned by what fixes it independently of its own
formulas: the Sobol generator matrices against the reference's vendored
table (when the reference is at hand) and against the net properties every
Sobol sequence has; the parent records against the per-history event traces
and against the device code compiled for the host; weight conservation of
every cancellation pass (kept + uniform = before, for wgt and wgt2); the
uniform particles' positions, materials, energies and directions; and an
unbiased k with cancellation.  The GPU path must then reproduce the
restatement bit for bit."""
import os
import re

import numpy as np
import pytest

from abeille_amd import Problem, abi, problems
from oracle.pyoracle import Oracle, lib as olib
from pyhd import HostDevice, lib as hlib

REF_SOBOL = "/root/reference/vendor/sobol/src/sobol.cpp"
LOW, HIGH = [-4., -5., -5.], [4., 5., 5.]

# (kind, beta, sobol, use_virtual_collisions, group_bins, shape)
CONFIGS = {
    "exact": ("exact", "average-g", True, None, [[0], [1, 2]], [8, 2, 2]),
    "exact_novirt": ("exact", "average-g", True, False, [[0], [1, 2]], [8, 2, 2]),
    "exact_one_region": ("exact", "average-g", True, True, [[0, 1, 2]], [1, 1, 1]),
    "basic_avg_g": ("basic", "average-g", True, None, None, [8, 2, 2]),
    "basic_avg_g_rng": ("basic", "average-g", False, None, None, [8, 2, 2]),
    "basic_avg_f": ("basic", "average-f", True, None, None, [8, 2, 2]),
    "basic_avg_f_rng": ("basic", "average-f", False, None, None, [4, 1, 1]),
    "basic_min": ("basic", "minimum", True, None, None, [8, 2, 2]),
    "basic_zero": ("basic", "zero", True, None, None, [8, 2, 2]),
}


def exact_model(name, n=3000, tracking="carter", ratio=0.85):
    kind, beta, sobol, virt, gbins, shape = CONFIGS[name]
    m = problems.exact_slab(n, tracking=tracking, ratio=ratio)
    m.exact_cancellation(LOW, HIGH, shape, kind=kind, beta=beta, sobol=sobol,
                         group_bins=gbins, use_virtual_collisions=virt)
    return m


def generation(o, n):
    """One transport of n fresh sources: (sites, parent records)."""
    src = o.sample_sources(0, n)
    r = o.transport(src, kcol=1.0, nthreads=0, cap=40 * n + 4096)
    return r["fission"], o.parents()


# ---------------------------------------------------------------- Sobol


def _sobol(idx, dim):
    return olib().orc_sobol(idx, dim)


def test_sobol_matrices_match_reference_table():
    if not os.path.exists(REF_SOBOL):
        pytest.skip("reference sources not present")
    txt = open(REF_SOBOL).read()
    i = txt.index("matrices[Matrices::num_dimensions * Matrices::size]")
    vals = [int(v, 16) for v in re.findall(r"0x([0-9A-Fa-f]+)ULL", txt[i:i + 40000])[:208]]
    mine = [olib().orc_sobol_matrix(d, k) for d in range(4) for k in range(52)]
    assert vals == mine


def test_sobol_first_points():
    """Points worked out by hand from the direction numbers: index 1 is 1/2
    in every dimension, index 2 is (1/4, 3/4, 3/4, 3/4), index 3 the XOR of
    both, index 4 (1/8, 5/8, 3/8, 1/8): dimension 3's third direction
    number is its initial m_3 = 1 over 2^3."""
    want = {0: [0, 0, 0, 0], 1: [.5, .5, .5, .5], 2: [.25, .75, .75, .75],
            3: [.75, .25, .25, .25], 4: [.125, .625, .375, .125]}
    for idx, pt in want.items():
        assert [_sobol(idx, d) for d in range(4)] == pt


@pytest.mark.parametrize("dim", range(4))
def test_sobol_is_stratified(dim):
    """The first 2^m points of each dimension put one point in each of the
    2^m intervals of width 2^-m (a (0, m, 1)-net in base 2)."""
    x = np.array([_sobol(i, dim) for i in range(4096)])
    for m in range(1, 13):
        cells = np.floor(x[: 1 << m] * (1 << m)).astype(int)
        assert sorted(cells) == list(range(1 << m))


def test_sobol_first_two_dimensions_form_a_0_2_sequence():
    """Dimensions 0 and 1 form a (0, 2)-sequence: every elementary box of
    area 2^-m holds exactly one of the first 2^m points."""
    pts = np.array([[_sobol(i, 0), _sobol(i, 1)] for i in range(1024)])
    for m in range(1, 11):
        p = pts[: 1 << m]
        for a in range(m + 1):
            ix = np.floor(p[:, 0] * (1 << a)).astype(int)
            iy = np.floor(p[:, 1] * (1 << (m - a))).astype(int)
            assert len(set(zip(ix, iy))) == 1 << m


def test_device_sobol_matches_restatement():
    for d in range(4):
        for i in list(range(300)) + [2 ** 40 + 12345, 2 ** 51 - 1]:
            assert hlib().hd_sobol(i, d) == _sobol(i, d)


# ----------------------------------------------------------- parent records


@pytest.mark.parametrize("tracking", ["delta", "carter"])
def test_parents_follow_the_traces(tracking):
    """Each site's parent record names the collision before the fission
    collision: in a run without reflection (here one-group, vacuum on every
    face) r1 is the position of an event of the parent history traced
    before the site's position, u1 the direction of flight into that event,
    and a parent's first collision has r1 = 0, u1 = (1, 0, 0)
    (Particle's defaults, particle.hpp:231-239)."""
    m = problems.subcritical_slab(300, tracking=tracking, cancel=False)
    m.set(mode=abi.K_EIGENVALUE)
    m.exact_cancellation([-1., -5., -5.], [1., 5., 5.], [4, 1, 1], kind="exact")
    o = Oracle(m)
    src = o.sample_sources(0, 300)
    r = o.transport(src, kcol=1.0, nthreads=1, cap=100000, trace_cap=4000)
    fis, par = r["fission"], o.parents()
    assert len(par) == len(fis) > 50
    tr, nrec = r["trace"], r["n_rec"]
    hist_ids = src["history_id"]
    checked = 0
    for s, q in zip(fis, par):
        h = int(np.nonzero(hist_ids == s["parent_history_id"])[0][0])
        ev = tr[h][: nrec[h]]
        pos = ev["r"]
        # the fission event: first record at the site's position
        k = int(np.nonzero(np.all(pos == s["r"], axis=1))[0][0])
        if np.all(q["r1"] == 0) and np.all(q["u1"] == [1, 0, 0]):
            continue
        prev = np.nonzero(np.all(pos[:k] == q["r1"], axis=1))[0]
        assert len(prev), "r1 is no earlier event of the parent"
        checked += 1
    assert checked > 20


@pytest.mark.parametrize("name,tracking", [("exact", "carter"), ("exact", "delta"),
                                           ("basic_avg_g", "carter")])
def test_device_parents_on_host_match_restatement(name, tracking):
    """The device transport code compiled for the host keeps the same parent
    records (reflections, virtual collisions, splitting included)."""
    m = exact_model(name, 400, tracking=tracking)
    o = Oracle(m)
    h = HostDevice(m)
    src = o.sample_sources(0, 400)
    ro = o.transport(src, kcol=1.0, nthreads=1, cap=40000)
    rh = h.transport(src, kcol=1.0, cap=40000)
    assert ro["fission"].tobytes() == rh["fission"].tobytes()
    po, ph = o.parents(), h.parents()
    assert len(po) == len(ro["fission"])
    assert po.tobytes() == ph.tobytes()
    if tracking == "delta":
        assert po["prev_virtual"].any() and not po["prev_virtual"].all()


# ------------------------------------------------------- cancellation pass


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_cancellation_conserves_weight(name):
    """Kept weight plus uniform weight equals the weight before, region by
    region and in total, for wgt and wgt2; the uniform particles lie in the
    mesh, in a cell, at a group midpoint, with unit directions; sites outside
    the mesh or (exact, without virtual collisions) born right after a
    virtual collision keep their weight."""
    m = exact_model(name)
    o = Oracle(m)
    fis, par = generation(o, 3000)
    rng = np.random.default_rng(5)
    fis = fis.copy()
    fis["wgt2"] = rng.normal(0.0, 1.0, len(fis))  # a second, mixed-sign weight
    o.rng_reset()
    out = o.exact_cancel(fis, par)
    n = len(fis)
    kept, uni = out[:n], out[n:]
    for f in ("wgt", "wgt2"):
        before = fis[f].sum()
        after = kept[f].sum() + uni[f].sum()
        assert after == pytest.approx(before, abs=1e-10 * np.abs(fis[f]).sum())
    kind, beta = CONFIGS[name][0], CONFIGS[name][1]
    if kind == "basic" and beta == "zero":
        assert out.tobytes() == fis.tobytes()
        return
    assert len(uni) > 0
    assert not np.array_equal(kept["wgt"], fis["wgt"])
    assert np.all(uni["r"] >= LOW) and np.all(uni["r"] <= HIGH)
    assert set(np.unique(uni["E"])) <= {0.5, 1.5, 2.5}
    np.testing.assert_allclose(np.linalg.norm(uni["u"], axis=1), 1.0, rtol=1e-14)
    if name == "exact_novirt":
        v = par["prev_virtual"] != 0
        assert v.any()
        assert kept[v].tobytes() == fis[v].tobytes()


def test_cancellation_reduces_the_absolute_weight():
    """Exact cancellation with average-g (the reference's default) removes
    signed weight: sum |w| after is below sum |w| before."""
    m = exact_model("exact")
    o = Oracle(m)
    fis, par = generation(o, 3000)
    out = o.exact_cancel(fis, par)
    assert (fis["wgt"] < 0).sum() > 50
    assert np.abs(out["wgt"]).sum() < 0.995 * np.abs(fis["wgt"]).sum()


def test_k_is_unbiased_with_exact_cancellation():
    """k with exact cancellation agrees with k without it (same problem, 4
    sigma): cancellation moves weight without biasing the eigenvalue."""
    ks = []
    for cancel in (False, True):
        m = problems.exact_slab(3000, ratio=0.9)
        if cancel:
            m.exact_cancellation(LOW, HIGH, [8, 2, 2], kind="exact", group_bins=[[0], [1, 2]])
        st, kh = Oracle(m).run(40, 10, nthreads=0)
        ks.append((st.kcol_avg, st.kcol_err))
    (a, sa), (b, sb) = ks
    assert abs(a - b) < 4 * np.hypot(sa, sb)


def test_exact_configuration_errors():
    m = problems.exact_slab(100)
    m.exact_cancellation(LOW, HIGH, [2, 1, 1], kind="exact", group_bins=[[0], [0, 1]])
    with pytest.raises(Exception, match="more than once"):
        Oracle(m)
    m.exact_cancellation(LOW, HIGH, [2, 1, 1], kind="exact", group_bins=[])
    with pytest.raises(Exception, match="Chi matrix is used"):
        Oracle(m)
    m = problems.exact_slab(100, tracking="surface")
    m.exact_cancellation(LOW, HIGH, [2, 1, 1], kind="basic", beta="minimum")
    with pytest.raises(Exception, match="surface-tracking"):
        Oracle(m)


# --------------------------------------------------------------- GPU path


@pytest.mark.gpu
@pytest.mark.parametrize("tracking", ["carter", "delta"])
def test_gpu_parents_match_restatement(tracking):
    print(f"[exact] parents tracking={tracking}", flush=True)
    m = exact_model("exact", 2000, tracking=tracking)
    o = Oracle(m)
    g = Problem(m)
    src = o.sample_sources(0, 2000)
    ro = o.transport(src, kcol=1.0, nthreads=0, cap=100000)
    rg = g.transport(src, kcol=1.0, cap=100000)
    assert ro["fission"].tobytes() == rg["fission"].tobytes()
    assert o.parents().tobytes() == g.parents().tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_gpu_cancellation_matches_restatement(name):
    """The same sites and parent records through both: weights and uniform
    particles bit for bit, twice in a row (settings::rng carries over)."""
    print(f"[exact] cancel config={name}", flush=True)
    m = exact_model(name)
    o = Oracle(m)
    g = Problem(m)
    fis, par = generation(o, 3000)
    fis = fis.copy()
    fis["wgt2"] = np.random.default_rng(9).normal(0.0, 1.0, len(fis))
    for rep in range(2):
        a = o.exact_cancel(fis, par)
        b = g.exact_cancel(fis, par, cap=len(a) + 16)
        assert len(a) == len(b), f"call {rep}"
        assert a.tobytes() == b.tobytes(), f"call {rep}"
    g.rng_reset()
    o.rng_reset()
    assert o.exact_cancel(fis, par).tobytes() == g.exact_cancel(fis, par).tobytes()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["exact", "basic_avg_g_rng", "basic_min"])
def test_gpu_run_with_exact_cancellation_matches_restatement(name):
    """PowerIterator::run with exact cancellation on the device: k per
    generation and the statistics agree with the restatement's run."""
    print(f"[exact] run config={name}", flush=True)
    m = exact_model(name, 2000)
    o = Oracle(m)
    g = Problem(m)
    so, ko = o.run(6, 2, nthreads=1)
    sg, kg = g.run(6, 2)
    np.testing.assert_allclose(kg, ko, rtol=1e-9)
    assert sg.kcol_avg == pytest.approx(so.kcol_avg, rel=1e-9)


@pytest.mark.gpu
def test_gpu_exact_cancellation_refusals():
    m = exact_model("exact", 500)
    g = Problem(m, engine=abi.ENGINE_EVENT)
    with pytest.raises(Exception, match="history engine"):
        g.run(2, 1)
    with pytest.raises(Exception, match="replicas"):
        Problem(m).run_replicas(2, 2, 1)


@pytest.mark.gpu
def test_gpu_rng_state_round_trip():
    """abx_rng_get_state / abx_rng_set_state: a pass repeated from a saved
    settings::rng state repeats its uniform particles and its end state."""
    m = exact_model("basic_avg_g_rng")
    o = Oracle(m)
    g = Problem(m)
    fis, par = generation(o, 2000)
    s0 = g.rng_state()
    a = g.exact_cancel(fis, par)
    s1 = g.rng_state()
    assert s1 != s0
    g.set_rng_state(s0)
    b = g.exact_cancel(fis, par)
    assert a.tobytes() == b.tobytes() and g.rng_state() == s1


def _one_group(kind, beta):
    m, _ = problems.la13511("PUa-1-0-SL", 3000, tracking="carter")
    m.set(sample_xs_ratio=[0.85])
    m.exact_cancellation([-1.853722, -5., -5.], [1.853722, 5., 5.], [8, 1, 1], kind=kind,
                         beta=beta)
    return m


def test_one_group_uniform_particles_draw_no_energy():
    """With one group the uniform particles' energy is a discrete draw over
    one weight, which takes no number: the pass still conserves weight and
    every uniform particle sits at the group's midpoint."""
    for kind, beta in (("exact", "average-g"), ("basic", "average-g")):
        m = _one_group(kind, beta)
        o = Oracle(m)
        fis, par = generation(o, 3000)
        out = o.exact_cancel(fis, par)
        assert len(out) > len(fis)
        assert out["wgt"].sum() == pytest.approx(fis["wgt"].sum(), abs=1e-10 * np.abs(fis["wgt"]).sum())
        assert np.all(out["E"][len(fis):] == 0.5)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["exact", "basic"])
def test_gpu_one_group_cancellation_matches_restatement(kind):
    print(f"[exact] one group kind={kind}", flush=True)
    m = _one_group(kind, "average-g")
    o = Oracle(m)
    g = Problem(m)
    fis, par = generation(o, 3000)
    for rep in range(2):
        a = o.exact_cancel(fis, par)
        b = g.exact_cancel(fis, par, cap=len(a) + 16)
        assert a.tobytes() == b.tobytes(), f"call {rep}"


def _hex_exact(n):
    m, _ = problems.hex_assembly(n, tracking="carter", top="flat", tally=False)
    m.set(sample_xs_ratio=[0.9] * 7)
    m.exact_cancellation([-5.54, -5.54, -10.], [5.54, 5.54, 10.], [4, 4, 2], kind="exact",
                         group_bins=[[0, 1, 2], [3, 4, 5, 6]])
    return m


def test_device_parents_on_host_match_restatement_hex():
    """Parent records through a hexagonal lattice (carter tracking, lattice
    relocations, reflective faces) from the host build of the device code."""
    m = _hex_exact(300)
    o = Oracle(m)
    h = HostDevice(m)
    src = o.sample_sources(0, 300)
    ro = o.transport(src, kcol=1.0, nthreads=1, cap=40000)
    rh = h.transport(src, kcol=1.0, cap=40000)
    assert ro["fission"].tobytes() == rh["fission"].tobytes()
    assert o.parents().tobytes() == h.parents().tobytes()


@pytest.mark.gpu
def test_gpu_exact_cancellation_hex_lattice():
    """Exact cancellation over a hexagonal assembly (regions split by fuel,
    moderator and guide tube, located through the lattice): parent records,
    one cancellation pass and a short run against the restatement."""
    print("[exact] hex assembly", flush=True)
    m = _hex_exact(3000)
    o = Oracle(m)
    g = Problem(m)
    src = o.sample_sources(0, 3000)
    ro = o.transport(src, kcol=1.0, nthreads=0, cap=200000)
    rg = g.transport(src, kcol=1.0, cap=200000)
    assert ro["fission"].tobytes() == rg["fission"].tobytes()
    par = o.parents()
    assert par.tobytes() == g.parents().tobytes()
    a = o.exact_cancel(ro["fission"], par)
    b = g.exact_cancel(ro["fission"], par, cap=len(a) + 16)
    assert len(a) > len(ro["fission"])
    assert a.tobytes() == b.tobytes()
    so, ko = Oracle(m).run(4, 1, nthreads=1)
    sg, kg = Problem(m).run(4, 1)
    np.testing.assert_allclose(kg, ko, rtol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["basic_avg_g", "basic_min", "basic_avg_f", "exact"])
def test_gpu_cancellation_stops_match_restatement(name):
    """A zero weight ends the basic cancelator's pass through a region
    (cancel_bin returns, basic :438-468) and a non-finite P_p or P_u ends
    either cancelator's: the second weights here have zeros sprinkled among
    both signs, so several regions stop part-way."""
    print(f"[exact] stops config={name}", flush=True)
    m = exact_model(name)
    o = Oracle(m)
    g = Problem(m)
    fis, par = generation(o, 3000)
    fis = fis.copy()
    rng = np.random.default_rng(17)
    w2 = rng.normal(0.0, 1.0, len(fis))
    w2[rng.random(len(fis)) < 0.03] = 0.0
    fis["wgt2"] = w2
    a = o.exact_cancel(fis, par)
    b = g.exact_cancel(fis, par, cap=len(a) + 16)
    assert a.tobytes() == b.tobytes()
    if name.startswith("basic"):
        # some regions kept sites after a zero untouched: not every
        # mixed-sign region changed all its second weights
        n = len(fis)
        assert np.any(a["wgt2"][:n] == fis["wgt2"]) and np.any(a["wgt2"][:n] != fis["wgt2"])

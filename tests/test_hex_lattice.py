"""Hexagonal lattices (src/hex_lattice.cpp) on the CPU restatement and on the
GPU.

Location: points of the assembly are located by the restatement's tracker
and compared with a brute-force answer computed here: the tile is the one
whose centre (get_hex_center's layout, src/hex_lattice.cpp:310-324) is
nearest among every tile of the plane - found by exhaustive search, not by
the cube rounding of get_nearest_hex - the layer is counted from the
lattice's lower z, the ring decides between the tile's universe (from the
element list as it was written, keyed by tile) and the outer universe, and
the pin cylinder decides the cell.  Both orientations, and an off-centre
origin (the reference subtracts the origin once more from z when counting
layers: src/hex_lattice.cpp:86-89, 303-307).

Tile distances: in a lattice of surface-free tiles, the nearest boundary
from points inside a tile is compared with the intersection of the ray with
the six lines the reference's distance_to_line solves (src/hex_lattice.cpp:
358-452) and the two layer planes, computed here with cross products.  Its
line equation A x + B y = D with D = (x2 - x1) y1 - (y2 - y1) x1 holds the
points reflected through the tile centre, so the lines are those through
(-x1, -y1) and (-x2, -y2).  These are not the edges of the hexagon that
get_nearest_hex assigns (a closed form at the centre shows it: moving
towards -y in a pointy lattice the boundary is at p / (2 cos 30), not at the
edge p / 2); they are kept as the reference has them.  Lines and layer
planes within 1e-10 cm of the point are skipped (RectLattice's tolerance):
the reference, which has none, crosses a line it stopped on for ever -
test_reference_lines_trap_a_particle shows the point where the first
histories of the flat assembly stall.

Known answer: a lattice whose every cell holds the LA-13511 PUa material in
a reflective box has k = k_inf = 2.612903.

GPU: per-history traces and ordered fission banks bit-exact against the
restatement (pointy and flat, surface and delta tracking, off-centre
origin; history and event engines), whole k-eigenvalue runs and tallies
against it, the k_inf known answer, batched k-eigenvalue chains equal to
each chain alone, and split precision against FP64 within statistics.
(The batch engine runs a hexagonal problem in tests/test_batch.py.)
"""
import numpy as np
import pytest

from abeille_amd import abi, problems
from abeille_amd.model import hex_tiles
from helpers import assert_traces_equal, bins_close
from oracle.pyoracle import Oracle

P, PZ, RPIN = 1.26, 10.0, 0.41
C30, S30 = np.cos(np.pi / 6), np.sin(np.pi / 6)
C60, S60 = np.cos(np.pi / 3), np.sin(np.pi / 3)


def centre(top, q, r):
    """get_hex_center (src/hex_lattice.cpp:310-324)."""
    if top == "flat":
        return np.array([P * (C60 * q + r), P * S60 * q])
    return np.array([P * C30 * r, P * (q + S30 * r)])


def brute_tile(top, xy, reach):
    """(q, r) of the nearest tile centre by exhaustive search, and the gap
    between the nearest and second-nearest distances."""
    best = []
    for q in range(-reach, reach + 1):
        for r in range(-reach, reach + 1):
            best.append((np.hypot(*(xy - centre(top, q, r))), q, r))
    best.sort()
    return (best[0][1], best[0][2]), best[1][0] - best[0][0]


def cell_ids(m):
    return np.array([c["id"] for c in m.cells])


@pytest.mark.parametrize("top", ["pointy", "flat"])
@pytest.mark.parametrize("origin", [(0., 0., 0.), (0.4, -0.3, 2.5)])
def test_oracle_hex_locate_brute_force(top, origin):
    rings, nz = 4, 2
    m, layout = problems.hex_assembly(1000, top=top, rings=rings, nz=nz, origin=origin)
    o = Oracle(m)
    rng = np.random.default_rng(7)
    hx, hz = rings * P + 0.5, nz * PZ / 2
    O = np.asarray(origin)
    pts = O + rng.uniform([-hx, -hx, -hz], [hx, hx, hz], size=(3000, 3))
    u = np.tile([0.6, 0.0, 0.8], (len(pts), 1))
    cells, _ = o.locate(pts, u)
    ids = cell_ids(m)
    z_low = O[2] - 0.5 * nz * PZ
    checked = {"fuel": 0, "gt": 0, "mod": 0, "outer": 0}
    for p, ci in zip(pts, cells):
        ro = p - O
        (q, r), gap = brute_tile(top, ro[:2], rings + 2)
        if gap < 1e-9:
            continue
        iz = int(np.floor((ro[2] - z_low) / PZ))
        if abs((ro[2] - z_low) / PZ - round((ro[2] - z_low) / PZ)) < 1e-9:
            continue
        ring = max(abs(q), abs(r), abs(q + r))
        u_id = layout.get((q, r, iz), -1) if (ring < rings and 0 <= iz < nz) else -1
        if u_id == -1:
            expect, kind = 105, "outer"
        else:
            rho = np.hypot(*(ro[:2] - centre(top, q, r)))
            if abs(rho - RPIN) < 1e-9:
                continue
            inside = rho < RPIN
            if u_id == 1:
                expect, kind = (101, "fuel") if inside else (102, "mod")
            else:
                expect, kind = (103, "gt") if inside else (104, "mod")
        assert ci >= 0 and ids[ci] == expect, (p, q, r, iz, ids[ci] if ci >= 0 else ci, expect)
        checked[kind] += 1
    assert min(checked.values()) > 50, checked


def tile_model(top, rings=3, nz=3, origin=(0., 0., 0.)):
    """A hex lattice of surface-free tiles (a cell bounded only by a far
    sphere) in a far reflective box: the nearest boundary of a point inside
    the lattice is its tile's."""
    from abeille_amd.model import Model
    m = Model()
    m.material(1, total=[1.0], absorption=[0.5], scatter=[[0.5]])
    sph = m.surface("sphere", [0., 0., 0., 1000.])
    c = m.cell(f"-{sph}", material=1, id=1)
    u = m.cell_universe([c], id=1)
    s = [m.surface("xplane", [-500.], "reflective"), m.surface("xplane", [500.], "reflective"),
         m.surface("yplane", [-500.], "reflective"), m.surface("yplane", [500.], "reflective"),
         m.surface("zplane", [-500.], "reflective"), m.surface("zplane", [500.], "reflective")]
    cb = m.cell(f"+{s[0]} & -{s[1]} & +{s[2]} & -{s[3]} & +{s[4]} & -{s[5]}", material=1, id=2)
    u_out = m.cell_universe([cb], id=2)
    n = len(hex_tiles(rings)) * nz
    m.root = m.hex_lattice(rings, nz, P, PZ, list(origin), [u] * n, top=top, outer=u_out, id=3)
    m.point_source([0., 0., 0.], energy=0.5)
    m.set(mode=abi.FIXED_SOURCE, nparticles=10, energy_bounds=[0., 1.])
    return m


def reference_lines(top):
    """The six (x1, y1, x2, y2) of distance_to_tile_boundary_pointy / _flat
    (src/hex_lattice.cpp:358-417)."""
    if top == "pointy":
        x1, y1, x2, y2 = 0., P / (2 * C30), P / 2, P / (2 * S30)
        return [(x1, y1, x2, y1), (x2, y2, x2, -y2), (x2, -y2, x1, -y1),
                (x1, -y1, -x2, -y2), (-x2, -y2, -x2, y2), (-x2, y2, x1, y1)]
    x1, y1, x2, y2 = P / (2 * S30), P / 2, P / (2 * C30), 0.
    return [(x1, y1, x2, y1), (x2, y2, x1, -y1), (x1, -y1, -x1, -y1),
            (-x1, -y1, -x2, -y2), (-x2, -y2, -x1, y1), (-x1, y1, x1, y1)]


def expected_distance(top, rt, u):
    """Nearest positive intersection with the reflected lines and the layer
    planes, by cross products."""
    best = np.inf
    for (x1, y1, x2, y2) in reference_lines(top):
        a = np.array([-x1, -y1])
        d = np.array([-x2, -y2]) - a
        den = u[0] * d[1] - u[1] * d[0]
        if den == 0:
            continue
        w = a - rt[:2]
        t = (w[0] * d[1] - w[1] * d[0]) / den
        if t > 0:
            best = min(best, t)
    for zp in (-PZ / 2, PZ / 2):
        if u[2] != 0:
            t = (zp - rt[2]) / u[2]
            if t > 0:
                best = min(best, t)
    return best


@pytest.mark.parametrize("top", ["pointy", "flat"])
def test_oracle_hex_tile_distance_closed_forms(top):
    """From the centre tile's centre along the axes."""
    o = Oracle(tile_model(top))
    R = P / (2 * C30)
    r0 = np.zeros((4, 3))
    u = np.array([[1., 0., 0.], [-1., 0., 0.], [0., 1., 0.], [0., -1., 0.]])
    d, _, _ = o.nearest_boundary(r0, u)
    if top == "pointy":
        # x = +-p/2; +y meets the two upper lines where they cross at (0, R),
        # -y the line y = -R (the true edges are at y = +-p/2)
        np.testing.assert_allclose(d, [P / 2, P / 2, R, R], rtol=1e-14)
    else:
        # y = +-p/2 are lines of the set; along x the slanted lines through
        # (+-R, 0) - the reference's flat x2
        np.testing.assert_allclose(d[2:], [P / 2, P / 2], rtol=1e-14)
        np.testing.assert_allclose(d[:2], [R, R], rtol=1e-14)


@pytest.mark.parametrize("top", ["pointy", "flat"])
@pytest.mark.parametrize("origin", [(0., 0., 0.), (0.2, 0.1, 0.)])
def test_oracle_hex_tile_distance_random(top, origin):
    rings, nz = 3, 3
    o = Oracle(tile_model(top, rings, nz, origin))
    rng = np.random.default_rng(11)
    O = np.asarray(origin)
    tiles = hex_tiles(rings)
    z_low = O[2] - 0.5 * nz * PZ
    pts, dirs, rts = [], [], []
    for k in range(400):
        q, r = tiles[rng.integers(len(tiles))]
        iz = int(rng.integers(nz))
        c = np.r_[centre(top, q, r), (iz + 0.5) * PZ + z_low]
        off = np.r_[rng.uniform(-0.35, 0.35, 2) * P, rng.uniform(-0.4, 0.4) * PZ]
        u = rng.normal(size=3) if k % 2 else np.r_[rng.normal(size=2), 0.]
        u /= np.linalg.norm(u)
        p = O + c + off
        pts.append(p)
        dirs.append(u)
        # the pad keeps r (not r - origin): the reference measures from the
        # tile centre without the origin (src/hex_lattice.cpp:347-356)
        rts.append(p - c)
    d, s, b = o.nearest_boundary(np.array(pts), np.array(dirs))
    for k in range(len(pts)):
        e = expected_distance(top, rts[k], dirs[k])
        assert b[k] == abi.NORMAL and s[k] == -1, k
        assert d[k] == pytest.approx(e, rel=1e-12, abs=1e-12), (k, pts[k], dirs[k], d[k], e)


def test_reference_lines_trap_a_particle():
    """At a point the first history of the flat assembly reaches, the
    reference's arithmetic (restated here in Python) finds its line 2 a
    positive 4.3e-17 cm ahead, and moving that far leaves the position
    unchanged: its surface tracker would cross the line for ever.  The
    restatement skips the line and goes on to a real distance."""
    import math
    p = 1.26
    c6, s6 = math.cos(math.pi / 6), math.sin(math.pi / 6)
    r = np.array([0.22679306344795597, 0.7075949318163149, 0.3943155797007687])
    u = np.array([0.18897322712349152, 0.977254788104953, 0.09624031668940415])
    # tile (q, r) = (1, 0) of layer 1: centre (p cos60, p sin60, 5)
    rt = r - np.array([p * (math.cos(math.pi / 3) * 1 + 0.), p * (math.sin(math.pi / 3) * 1), 5.])
    x1, y1, x2, y2 = p / (2. * s6), p / 2., p / (2. * c6), 0.
    A, B = -y1 - y2, x2 - x1          # the line (x2, y2) -> (x1, -y1)
    D = (x1 - x2) * y2 - (-y1 - y2) * x2
    d = (D - A * rt[0] - B * rt[1]) / (A * u[0] + B * u[1])
    assert 0 < d < 1e-16
    assert np.array_equal(r + d * u, r)
    m, _ = problems.hex_assembly(100, tracking="surface", top="flat", tally=False)
    dist, _, _ = Oracle(homogeneous(m)).nearest_boundary(r[None], u[None])
    assert dist[0] > 1e-3


def homogeneous(m):
    """The hex assembly with every material the LA-13511 PUa medium in a
    fully reflective box, one group."""
    pua = dict(total=[0.32640], absorption=[0.101184], scatter=[[0.225216]],
               fission=[0.081600], nu=[3.24], chi=[[1.0]])
    m.materials = []
    for mid in (1, 2, 3):
        m.material(mid, **pua)
    m.surfaces = [(t, p, abi.REFLECTIVE if bc != abi.NORMAL else bc, sid)
                  for t, p, bc, sid in m.surfaces]
    m.tallies = []
    m.set(energy_bounds=[0., 1.])
    return m


def test_oracle_hex_k_inf():
    m, _ = problems.hex_assembly(4000, tracking="surface", top="flat", tally=False)
    st, _ = Oracle(homogeneous(m)).run(30, 10)
    assert abs(st.kcol_avg - 2.612903) < 4 * st.kcol_err, (st.kcol_avg, st.kcol_err)


def test_deck_hexagonal(tmp_path):
    """A deck's `type: hexagonal` lattice reads into the same model as the
    builder's (src/hex_lattice.cpp:454-570 keys)."""
    from abeille_amd.deck import model_from_deck
    m, layout = problems.hex_assembly(1000, top="flat", tally=False)
    lat = m.universes[-1]
    deck = f"""
materials:
  - {{id: 1, name: fuel, total: [0.5], absorption: [0.2], scatter: [[0.3]], fission: [0.1], nu: [2.5], chi: [1.0]}}
  - {{id: 2, name: mod, total: [0.5], absorption: [0.01], scatter: [[0.49]]}}
  - {{id: 3, name: gt, total: [0.5], absorption: [0.01], scatter: [[0.49]]}}
surfaces:
  - {{id: 1, type: zcylinder, x0: 0., y0: 0., r: 0.41}}
  - {{id: 11, type: xplane, x0: -5.54, boundary: vacuum}}
  - {{id: 12, type: xplane, x0: 5.54, boundary: vacuum}}
  - {{id: 13, type: yplane, y0: -5.54, boundary: reflective}}
  - {{id: 14, type: yplane, y0: 5.54, boundary: reflective}}
  - {{id: 15, type: zplane, z0: -10., boundary: reflective}}
  - {{id: 16, type: zplane, z0: 10., boundary: reflective}}
cells:
  - {{id: 101, region: "-1", material: 1}}
  - {{id: 102, region: "+1", material: 2}}
  - {{id: 103, region: "-1", material: 3}}
  - {{id: 104, region: "+1", material: 2}}
  - {{id: 105, region: "+11 & -12 & +13 & -14 & +15 & -16", material: 2}}
universes:
  - {{id: 1, cells: [101, 102]}}
  - {{id: 2, cells: [103, 104]}}
  - {{id: 3, cells: [105]}}
  - id: 10
    type: hexagonal
    shape: [4, 2]
    pitch: [1.26, 10.0]
    origin: [0., 0., 0.]
    top: flat
    outer: 3
    universes: {lat["elems"]}
root-universe: 10
sources:
  - {{spatial: {{type: point, position: [0., 0., 0.]}}, direction: {{type: isotropic}}, energy: {{type: mono-energetic, energy: 0.5}}, weight: 1.0}}
settings:
  simulation: k-eigenvalue
  nparticles: 1000
  ngenerations: 10
  nignored: 2
  ngroups: 1
  energy-bounds: [0., 1.]
"""
    f = tmp_path / "hex.yaml"
    f.write_text(deck)
    md = model_from_deck(str(f))
    u = [x for x in md.universes if x["id"] == 10][0]
    assert u["type"] == abi.HEX_LATTICE and u["top"] == "flat"
    assert u["shape"][:2] == [4, 2] and u["pitch"][:2] == [1.26, 10.0]
    assert u["elems"] == lat["elems"] and u["outer"] == 3
    with pytest.raises(ValueError, match="Improper number"):
        md.hex_lattice(3, 1, 1.26, 1.0, [0, 0, 0], [1] * 18)


# ---------------------------------------------------------------- GPU
TRACE_CASES = {
    "pointy_st": dict(top="pointy", tracking="surface"),
    "flat_st": dict(top="flat", tracking="surface"),
    "pointy_dt": dict(top="pointy", tracking="delta"),
    "flat_st_origin": dict(top="flat", tracking="surface", origin=(0.4, -0.3, 2.5)),
    "pointy_st_origin": dict(top="pointy", tracking="surface", origin=(-0.2, 0.5, 0.)),
}


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(TRACE_CASES))
def test_gpu_hex_traces_and_banks(name):
    from abeille_amd import Problem
    print(f"[hex] traces {name}", flush=True)
    m, _ = problems.hex_assembly(2000, **TRACE_CASES[name])
    n, cap = 2000, 512
    o, g = Oracle(m), Problem(m)
    so = o.sample_sources(0, n)
    assert so.tobytes() == g.sample_sources(0, n).tobytes(), "source particles differ"
    ro = o.transport(so, kcol=1.0, nthreads=1, trace_cap=cap)
    rg = g.transport(so, kcol=1.0, trace_cap=cap)
    assert_traces_equal(ro, rg, cap)
    assert len(ro["fission"]) > 0
    assert ro["fission"].tobytes() == rg["fission"].tobytes(), "fission banks differ"


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["pointy_st", "flat_st_origin", "pointy_dt"])
def test_gpu_hex_run_matches_oracle(name):
    from abeille_amd import Problem
    print(f"[hex] run {name}", flush=True)
    m, _ = problems.hex_assembly(5000, **TRACE_CASES[name])
    o, g = Oracle(m), Problem(m)
    so, ko = o.run(4, 1, nthreads=0)
    sg, kg = g.run(4, 1)
    assert np.allclose(ko, kg, rtol=1e-9, atol=0), (ko, kg)
    assert so.diag.collisions == sg.diag.collisions
    for t in range(len(m.tallies)):
        assert bins_close(o.tally_stats(t)[0], g.tally_stats(t)[0], 1e-8)


@pytest.mark.gpu
@pytest.mark.parametrize("tracking", ["surface", "delta"])
def test_gpu_hex_k_inf(tracking):
    from abeille_amd import Problem
    m, _ = problems.hex_assembly(20000, tracking=tracking, top="pointy", tally=False)
    st, _ = Problem(homogeneous(m)).run(40, 10)
    assert abs(st.kcol_avg - 2.612903) < 4 * st.kcol_err, (st.kcol_avg, st.kcol_err)


@pytest.mark.gpu
def test_gpu_hex_event_engine_traces():
    from abeille_amd import Problem
    m, _ = problems.hex_assembly(1000, tracking="surface", top="pointy")
    n, cap = 1000, 512
    o, g = Oracle(m), Problem(m, engine=abi.ENGINE_EVENT)
    so = o.sample_sources(0, n)
    ro = o.transport(so, kcol=1.0, nthreads=1, trace_cap=cap)
    rg = g.transport(so, kcol=1.0, trace_cap=cap)
    assert_traces_equal(ro, rg, cap)
    assert ro["fission"].tobytes() == rg["fission"].tobytes()


@pytest.mark.gpu
def test_gpu_hex_replicas_equal_chains_alone():
    from abeille_amd import Problem
    seeds = [abi.DEFAULT_SEED, 987654321]
    make = lambda: problems.hex_assembly(3000, tracking="surface", top="flat")[0]
    g = Problem(make())
    _, kh = g.run_replicas(2, 5, 2, seeds=seeds)
    for r, sd in enumerate(seeds):
        m = make()
        m.set(seed=sd)
        _, ka = Problem(m).run(5, 2)
        assert np.allclose(kh[r], ka, rtol=1e-9, atol=0), (r, kh[r], ka)


@pytest.mark.gpu
def test_gpu_hex_split_matches_fp64():
    from abeille_amd import Problem
    out = []
    for prec in (abi.FP64, abi.SPLIT):
        m, _ = problems.hex_assembly(40000, tracking="surface", top="pointy", tally=False)
        m.set(precision=prec)
        st, _ = Problem(m).run(24, 8)
        out.append(st)
    sig = np.hypot(out[0].kcol_err, out[1].kcol_err)
    assert abs(out[0].kcol_avg - out[1].kcol_avg) < 4 * sig, (out[0].kcol_avg, out[1].kcol_avg, sig)

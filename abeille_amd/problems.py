"""Seeded generators of the benchmark problems of the transport path.

  fixed_source_cube   20x20x20 cm vacuum cube, 2 groups, point source, delta
                      tracking, 20^3 collision flux tally (the CPU-runnable
                      fixed-source case).
  pin_lattice_2d      17x17 pin lattice, pitch 1.26 cm, pin radius 0.54 cm,
                      reflective, 7 groups, k-eigenvalue, delta tracking,
                      51x51x7 track-length flux tally.
  core_3x3            3x3 assemblies of the 17x17 lattice as nested lattices,
                      0-150 cm axially (vacuum top/bottom), surface tracking,
                      153x153x30x7 track-length tally.
  slab_problem        one random fixed-source slab problem (batched form).
  subcritical_slab    modified-fixed-source one-group slab with carter
                      tracking and approximate mesh cancellation.
  la13511             analytic one-group cases (LA-13511) with their published
                      k values, used as known answers.
  hex_assembly        hexagonal pin lattice (pointy or flat) with empty
                      corner tiles, two axial layers, in a moderator box.

Cross sections of the lattice problems are drawn from a numpy Generator with
a fixed seed: per group sigma_t ~ U[0.2, 2.0], scattering ratio ~ U[0.5, 0.99],
scatter rows downscatter-dominant with <= 10 % upscatter from the two lowest
groups only; the fuel has nu*sigma_f ~ U[0.005, 0.3] written as nu = 2.5 and
chi concentrated in the first three groups.  Sigma_f <= Sigma_a is enforced
by raising Sigma_a (src/mg_nuclide.cpp:267-272 requires it).
"""
import numpy as np

from . import abi
from .model import Model, hex_tiles

# The 25 guide-tube / instrument positions of a 17x17 PWR assembly
# (row = x index, column = y index; the layout of input_files/c5g7.yaml:257-273).
GT_POSITIONS = [(2, 5), (2, 8), (2, 11), (3, 3), (3, 13), (5, 2), (5, 5), (5, 8),
                (5, 11), (5, 14), (8, 2), (8, 5), (8, 8), (8, 11), (8, 14), (11, 2),
                (11, 5), (11, 8), (11, 11), (11, 14), (13, 3), (13, 13), (14, 5),
                (14, 8), (14, 11)]


def seven_group_xs(seed=20240607, G=7):
    """Fuel, moderator and guide-tube cross sections (dicts of arrays)."""
    rng = np.random.default_rng(seed)
    mats = {}
    for name in ("fuel", "moderator", "guide_tube"):
        st = rng.uniform(0.2, 2.0, G)
        c = rng.uniform(0.5, 0.99, G)
        sa = st * (1.0 - c)
        nusf = rng.uniform(0.005, 0.3, G) if name == "fuel" else None
        sf = None
        if nusf is not None:
            sf = nusf / 2.5
            sa = np.maximum(sa, sf)
        ss = st - sa
        scat = np.zeros((G, G))
        for g in range(G):
            w = np.zeros(G)
            w[g] = rng.uniform(0.5, 1.0)               # in-group
            for o in range(g + 1, G):                   # downscatter, decaying
                w[o] = rng.uniform(0.05, 0.4) * 0.5 ** (o - g - 1)
            w = w / w.sum()
            if g >= G - 2:                              # groups 6-7: upscatter
                up = rng.uniform(0.0, 0.10)
                w *= (1.0 - up)
                w[g - 1] += up
            scat[g] = ss[g] * w
        # make Sigma_t exactly Sigma_a + row sums so check_xs passes tightly
        st = sa + scat.sum(axis=1)
        m = dict(total=st, absorption=sa, scatter=scat)
        if sf is not None:
            chi = np.zeros(G)
            chi[:3] = rng.uniform(0.2, 1.0, 3)
            chi[3:] = rng.uniform(0.0, 0.02, G - 3)
            chi /= chi.sum()
            m.update(fission=sf, nu=np.full(G, 2.5), chi=chi)
        mats[name] = m
    return mats


def fixed_source_cube(nparticles=25000, tracking="delta"):
    """20x20x20 cm vacuum cube of one 2-group material, point source at the
    centre in group 1 (E = 0.5), collision-estimator flux on 20^3 x 2."""
    m = Model()
    s = [m.surface("xplane", [-10.], "vacuum"), m.surface("xplane", [10.], "vacuum"),
         m.surface("yplane", [-10.], "vacuum"), m.surface("yplane", [10.], "vacuum"),
         m.surface("zplane", [-10.], "vacuum"), m.surface("zplane", [10.], "vacuum")]
    m.material(1, total=[0.5, 1.2], absorption=[0.05, 0.20],
               scatter=[[0.30, 0.15], [0.00, 1.00]])
    m.cell(f"+{s[0]} & -{s[1]} & +{s[2]} & -{s[3]} & +{s[4]} & -{s[5]}", material=1, id=1)
    m.root = m.cell_universe([1], id=1)
    m.point_source([0., 0., 0.], energy=0.5)
    m.mesh_tally([-10., -10., -10.], [10., 10., 10.], [20, 20, 20], [0., 1., 2.],
                 estimator="collision")
    m.set(mode=abi.FIXED_SOURCE, nparticles=nparticles, energy_bounds=[0., 1., 2.],
          tracking=_tracking(tracking))
    return m


def cylinder_cube(nparticles=25000, tracking="surface", axis=(0., 2., 0.), radius=4.,
                  centre=(0.5, 3.0, -0.5)):
    """The 20 cm vacuum cube with a general Cylinder (src/cylinder.cpp) of
    another, more absorbing material through it; `axis` need not be
    normalised.  Point source at
    the centre, inside the cylinder; collision-estimator flux on 20^3 x 2."""
    m = Model()
    s = [m.surface("xplane", [-10.], "vacuum"), m.surface("xplane", [10.], "vacuum"),
         m.surface("yplane", [-10.], "vacuum"), m.surface("yplane", [10.], "vacuum"),
         m.surface("zplane", [-10.], "vacuum"), m.surface("zplane", [10.], "vacuum")]
    cyl = m.surface("cylinder", [*centre, *axis, radius])
    m.material(1, total=[0.5, 1.2], absorption=[0.05, 0.20],
               scatter=[[0.30, 0.15], [0.00, 1.00]])
    m.material(2, total=[0.8, 1.5], absorption=[0.30, 0.60],
               scatter=[[0.35, 0.15], [0.00, 0.90]])
    box = f"+{s[0]} & -{s[1]} & +{s[2]} & -{s[3]} & +{s[4]} & -{s[5]}"
    m.cell(f"-{cyl} & {box}", material=2, id=1)
    m.cell(f"+{cyl} & {box}", material=1, id=2)
    m.root = m.cell_universe([1, 2], id=1)
    m.point_source([0., 0., 0.], energy=0.5)
    m.mesh_tally([-10., -10., -10.], [10., 10., 10.], [20, 20, 20], [0., 1., 2.],
                 estimator="collision")
    m.set(mode=abi.FIXED_SOURCE, nparticles=nparticles, energy_bounds=[0., 1., 2.],
          tracking=_tracking(tracking))
    return m


def _tracking(name):
    return {"delta": abi.DELTA_TRACKING, "surface": abi.SURFACE_TRACKING,
            "carter": abi.CARTER_TRACKING}[name]


def _pin_universes(m, fuel_id, mod_id, gt_id, cyl):
    """Fuel-pin and guide-tube-pin cell universes inside one cylinder."""
    cf = m.cell(f"-{cyl}", material=fuel_id, id=101)
    cm = m.cell(f"+{cyl}", material=mod_id, id=102)
    cg = m.cell(f"-{cyl}", material=gt_id, id=103)
    cm2 = m.cell(f"+{cyl}", material=mod_id, id=104)
    u_fuel = m.cell_universe([cf, cm], id=1)
    u_gt = m.cell_universe([cg, cm2], id=2)
    return u_fuel, u_gt


def _assembly_elems(u_fuel, u_gt):
    elems = []
    gt = set(GT_POSITIONS)
    for nx in range(17):          # x runs down the list, y fastest
        for ny in range(17):
            elems.append(u_gt if (nx, ny) in gt else u_fuel)
    return elems


def _add_lattice_materials(m, xs):
    m.material(1, **xs["fuel"])
    m.material(2, **xs["moderator"])
    m.material(3, **xs["guide_tube"])
    return 1, 2, 3


def pin_lattice_2d(nparticles=1_000_000, tracking="delta", xs_seed=20240607,
                   tally=True):
    """17x17 reflective pin lattice, 7 groups, k-eigenvalue.  The lattice is
    the root universe; its outer universe holds the reflective box, so
    boundary conditions are found through Lattice::get_boundary_condition
    (src/lattice.cpp:77-91) as in input_files/c5g7.yaml."""
    G = 7
    m = Model()
    xs = seven_group_xs(xs_seed, G)
    fuel, mod, gt = _add_lattice_materials(m, xs)
    cyl = m.surface("zcylinder", [0., 0., 0.54], id=1)
    h = 17 * 1.26 / 2.0  # 10.71
    b = [m.surface("xplane", [-h], "reflective", id=11),
         m.surface("xplane", [h], "reflective", id=12),
         m.surface("yplane", [-h], "reflective", id=13),
         m.surface("yplane", [h], "reflective", id=14),
         m.surface("zplane", [-h], "reflective", id=15),
         m.surface("zplane", [h], "reflective", id=16)]
    u_fuel, u_gt = _pin_universes(m, fuel, mod, gt, cyl)
    cbox = m.cell(f"+{b[0]} & -{b[1]} & +{b[2]} & -{b[3]} & +{b[4]} & -{b[5]}",
                  material=mod, id=105)
    u_out = m.cell_universe([cbox], id=3)
    m.root = m.rect_lattice([17, 17, 1], [1.26, 1.26, 2 * h], [0., 0., 0.],
                            _assembly_elems(u_fuel, u_gt), outer=u_out, id=10)
    m.box_source([-h, -h, -h], [h, h, h], energy=0.5, fissile_only=True)
    if tally:
        m.mesh_tally([-h, -h, -h], [h, h, h], [51, 51, 1], np.arange(G + 1.0),
                     estimator="track-length")
    m.set(mode=abi.K_EIGENVALUE, nparticles=nparticles,
          energy_bounds=np.arange(G + 1.0), tracking=_tracking(tracking))
    return m


def core_3x3(nparticles=2_000_000, tracking="surface", xs_seed=20240607,
             precision=abi.FP64, tally=True):
    """3x3 core of the 17x17 assembly as nested lattices, z in [0, 150] cm
    with vacuum top and bottom and reflective radial faces."""
    G = 7
    m = Model()
    xs = seven_group_xs(xs_seed, G)
    fuel, mod, gt = _add_lattice_materials(m, xs)
    cyl = m.surface("zcylinder", [0., 0., 0.54], id=1)
    H = 3 * 17 * 1.26 / 2.0  # 32.13
    b = [m.surface("xplane", [-H], "reflective", id=11),
         m.surface("xplane", [H], "reflective", id=12),
         m.surface("yplane", [-H], "reflective", id=13),
         m.surface("yplane", [H], "reflective", id=14),
         m.surface("zplane", [0.], "vacuum", id=15),
         m.surface("zplane", [150.], "vacuum", id=16)]
    u_fuel, u_gt = _pin_universes(m, fuel, mod, gt, cyl)
    cbox = m.cell(f"+{b[0]} & -{b[1]} & +{b[2]} & -{b[3]} & +{b[4]} & -{b[5]}",
                  material=mod, id=105)
    u_out = m.cell_universe([cbox], id=3)
    assy = m.rect_lattice([17, 17, 1], [1.26, 1.26, 150.], [0., 0., 0.],
                          _assembly_elems(u_fuel, u_gt), outer=None, id=10)
    m.root = m.rect_lattice([3, 3, 1], [21.42, 21.42, 150.], [0., 0., 75.],
                            [assy] * 9, outer=u_out, id=20)
    m.box_source([-H, -H, 0.], [H, H, 150.], energy=0.5, fissile_only=True)
    if tally:
        m.mesh_tally([-H, -H, 0.], [H, H, 150.], [153, 153, 30], np.arange(G + 1.0),
                     estimator="track-length")
    m.set(mode=abi.K_EIGENVALUE, nparticles=nparticles,
          energy_bounds=np.arange(G + 1.0), tracking=_tracking(tracking),
          precision=precision)
    return m


def hex_assembly(nparticles=20000, tracking="surface", top="pointy", rings=4, nz=2,
                 origin=(0., 0., 0.), xs_seed=20240607, tally=True, mode=abi.K_EIGENVALUE):
    """Hexagonal pin assembly: `rings` rings of pins (pitch 1.26 cm, pin
    radius 0.41 cm) in `nz` axial layers of 10 cm, inside a moderator box
    (vacuum x faces, reflective y and z faces) entered as the lattice's
    outer universe.  Layer 0 has guide tubes in the centre and on every
    third tile of the inner rings and leaves the outer ring's six corners
    empty (-1: the outer universe); layer 1 swaps fuel and guide tubes.
    Returns (model, {(q, r, layer): universe id or -1})."""
    G = 7
    m = Model()
    xs = seven_group_xs(xs_seed, G)
    fuel, mod, gt = _add_lattice_materials(m, xs)
    cyl = m.surface("zcylinder", [0., 0., 0.41], id=1)
    u_fuel, u_gt = _pin_universes(m, fuel, mod, gt, cyl)
    p, pz = 1.26, 10.0
    hx, hz = rings * p + 0.5, nz * pz / 2.0
    # the lattice's layers sit at 2 * origin[2] +- hz (HexLattice::get_tile
    # counts them from the origin-relative point, src/hex_lattice.cpp:
    # 303-307); the box spans them and the nominal ones
    zlo, zhi = min(origin[2], 2 * origin[2]) - hz, max(origin[2], 2 * origin[2]) + hz
    b = [m.surface("xplane", [origin[0] - hx], "vacuum", id=11),
         m.surface("xplane", [origin[0] + hx], "vacuum", id=12),
         m.surface("yplane", [origin[1] - hx], "reflective", id=13),
         m.surface("yplane", [origin[1] + hx], "reflective", id=14),
         m.surface("zplane", [zlo], "reflective", id=15),
         m.surface("zplane", [zhi], "reflective", id=16)]
    cbox = m.cell(f"+{b[0]} & -{b[1]} & +{b[2]} & -{b[3]} & +{b[4]} & -{b[5]}",
                  material=mod, id=105)
    u_out = m.cell_universe([cbox], id=3)
    layout, elems = {}, []
    last = rings - 1
    corners = {(last, 0), (-last, 0), (0, last), (0, -last), (last, -last), (-last, last)}
    for iz in range(nz):
        for (q, r) in hex_tiles(rings):
            ring = max(abs(q), abs(r), abs(q + r))
            u = u_gt if (ring < last and (q - r) % 3 == 0) else u_fuel
            if iz % 2 == 1:
                u = u_fuel if u == u_gt else u_gt
            if iz % 2 == 0 and rings > 1 and (q, r) in corners:
                u = -1
            layout[(q, r, iz)] = u
            elems.append(u)
    m.root = m.hex_lattice(rings, nz, p, pz, list(origin), elems, top=top, outer=u_out, id=10)
    m.box_source([origin[0] - hx, origin[1] - hx, zlo], [origin[0] + hx, origin[1] + hx, zhi],
                 energy=0.5, fissile_only=True)
    if tally:
        m.mesh_tally([origin[0] - hx, origin[1] - hx, zlo], [origin[0] + hx, origin[1] + hx, zhi],
                     [24, 24, 2], np.arange(G + 1.0), estimator="track-length")
    m.set(mode=mode, nparticles=nparticles, energy_bounds=np.arange(G + 1.0),
          tracking=_tracking(tracking))
    return m, layout


def la13511(case, nparticles=100_000, tracking="surface"):
    """One-group LA-13511 cases with analytic k: 'PUa-1-0-IN' (2.612903),
    'PUb-1-0-IN' (2.290323), 'PUa-1-0-SL' (1.0), 'PUb-1-0-SL' (1.0),
    'Ua-1-1-IN' (2.25).  Infinite media are reflective 10 cm boxes; slabs are
    vacuum in x, reflective in y and z."""
    data = {
        "PUa-1-0-IN": (0.32640, 0.101184, 0.081600, 3.24, 0.225216, None, 2.612903),
        "PUb-1-0-IN": (0.32640, 0.101184, 0.081600, 2.84, 0.225216, None, 2.290323),
        "PUa-1-0-SL": (0.32640, 0.101184, 0.081600, 3.24, 0.225216, 1.853722, 1.0),
        "PUb-1-0-SL": (0.32640, 0.101184, 0.081600, 2.84, 0.225216, 2.256751, 1.0),
        "Ua-1-1-IN": (0.32640, 0.078336, 0.065280, 2.7, 0.248064, None, 2.25),
    }
    st, sa, sf, nu, ss, half, k = data[case]
    m = Model()
    m.material(1, total=[st], absorption=[sa], scatter=[[ss]], fission=[sf], nu=[nu],
               chi=[[1.0]])
    xb = "vacuum" if half else "reflective"
    hx = half if half else 5.0
    s = [m.surface("xplane", [-hx], xb), m.surface("xplane", [hx], xb),
         m.surface("yplane", [-5.], "reflective"), m.surface("yplane", [5.], "reflective"),
         m.surface("zplane", [-5.], "reflective"), m.surface("zplane", [5.], "reflective")]
    m.cell(f"+{s[0]} & -{s[1]} & +{s[2]} & -{s[3]} & +{s[4]} & -{s[5]}", material=1, id=1)
    m.root = m.cell_universe([1], id=1)
    m.point_source([0., 0., 0.], energy=0.5)
    m.set(mode=abi.K_EIGENVALUE, nparticles=nparticles, energy_bounds=[0., 1.],
          tracking=_tracking(tracking))
    return m, k


def infinite_medium_2g(nparticles=20000, tracking="delta", mode=abi.FIXED_SOURCE):
    """The cube's 2-group material in a reflective 20 cm box: the volume-
    integrated collision-estimator flux per source particle is
    phi_1 = 1/(St1 - Ss11) = 5.0 cm and phi_2 = Ss12 phi_1/(St2 - Ss22) = 3.75 cm."""
    m = fixed_source_cube(nparticles, tracking)
    m.surfaces = [(t, p, abi.REFLECTIVE, sid) for (t, p, _, sid) in m.surfaces]
    m.tallies = []
    m.mesh_tally([-10., -10., -10.], [10., 10., 10.], [1, 1, 1], [0., 1., 2.],
                 estimator="collision")
    m.mesh_tally([-10., -10., -10.], [10., 10., 10.], [1, 1, 1], [0., 1., 2.],
                 estimator="track-length")
    m.set(mode=mode)
    return m


def slab_problem(rng, G=None, tracking=None, nhist=None, fission_free=None,
                 nbatch_hist=None):
    """One batched-form problem: a box of 1-8 x-stacked slab regions with a
    vacuum outer boundary, G in {2,4,7,16,33}, per-group XS from the lattice
    distributions, a point source in the first slab, and a 16^3 collision
    flux tally.  Fission-free by default; `fission_free=False` gives every
    region a subcritical fuel (nu Sigma_f = 0.3 Sigma_a, chi into the first
    group), so a fixed-source run carries fission secondaries.
    Returns (model, histories)."""
    if G is None:
        G = int(rng.choice([2, 4, 7, 16, 33]))
    nreg = int(rng.integers(1, 9))
    if tracking is None:
        tracking = "delta" if rng.random() < 0.5 else "surface"
    if nhist is None:
        nhist = int(np.exp(rng.uniform(np.log(1e3), np.log(1e6))))
    if fission_free is None:
        fission_free = True
    m = Model()
    width = rng.uniform(2.0, 10.0, nreg)
    hy = float(rng.uniform(5.0, 20.0))
    _slab_geometry(m, rng, G, width, hy, fission_free)
    m.set(mode=abi.FIXED_SOURCE, nparticles=nbatch_hist or nhist,
          energy_bounds=np.arange(G + 1.0), tracking=_tracking(tracking))
    return m, nhist


def _slab_geometry(m, rng, G, width, hy, fission_free, nu_fission_ratio=0.3,
                   c_range=(0.5, 0.99)):
    nreg = len(width)
    x = np.concatenate([[0.0], np.cumsum(width)])
    x -= x[-1] / 2.0
    sx = [m.surface("xplane", [float(v)], "vacuum" if i in (0, nreg) else "normal")
          for i, v in enumerate(x)]
    sy = [m.surface("yplane", [-hy], "vacuum"), m.surface("yplane", [hy], "vacuum")]
    sz = [m.surface("zplane", [-hy], "vacuum"), m.surface("zplane", [hy], "vacuum")]
    cells = []
    for r in range(nreg):
        st = rng.uniform(0.2, 2.0, G)
        c = rng.uniform(c_range[0], c_range[1], G)
        sa = st * (1 - c)
        scat = np.zeros((G, G))
        for g in range(G):
            w = np.zeros(G)
            w[g] = rng.uniform(0.5, 1.0)
            for o in range(g + 1, min(G, g + 4)):
                w[o] = rng.uniform(0.05, 0.4)
            w /= w.sum()
            scat[g] = (st[g] - sa[g]) * w
        st = sa + scat.sum(axis=1)
        kw = {}
        if not fission_free:
            chi = np.zeros(G)
            chi[0] = 1.0
            kw = dict(fission=nu_fission_ratio * sa / 2.5, nu=np.full(G, 2.5), chi=chi)
        m.material(r + 1, total=st, absorption=sa, scatter=scat, **kw)
        cells.append(m.cell(f"+{sx[r]} & -{sx[r + 1]} & +{sy[0]} & -{sy[1]} & "
                            f"+{sz[0]} & -{sz[1]}", material=r + 1, id=r + 1))
    m.root = m.cell_universe(cells, id=1)
    m.point_source([float(0.5 * (x[0] + x[1])), 0., 0.], energy=0.5)
    m.mesh_tally([float(x[0]), -hy, -hy], [float(x[-1]), hy, hy], [16, 16, 16],
                 np.arange(G + 1.0), estimator="collision")
    return x


def served_request(rng, nbatches=2):
    """One request of the served stream: a slab problem from the batched-form
    generator with 1e4 - 5e6 histories (log-uniform) split over `nbatches`
    batches.  One request in five is a modified-fixed-source chain in a thin
    fissile box (1-3 slabs of 1-3 cm, |y|, |z| < 2-6 cm) with carter
    tracking on a sampling cross section of 0.8 x the majorant and
    approximate mesh cancellation on 10 x 10 x 10; the rest are fission-free
    fixed-source problems with delta or surface tracking.  The fissile boxes
    have nu Sigma_f = 0.05 Sigma_a and scattering ratios in [0.3, 0.6]: in
    the negative-weight branch a flight multiplies |w| by 1.5 and a real
    collision by the scattering ratio, so with ratios near one weights (and,
    through splitting, particle counts) grow without bound, and the chain's
    particle count multiplies by roughly the |w|-weighted k per generation.
    Returns (model, total histories, kind)."""
    nhist = int(np.exp(rng.uniform(np.log(1e4), np.log(5e6))))
    per_batch = max(1, nhist // nbatches)
    if rng.random() < 0.2:
        G = int(rng.choice([2, 4, 7, 16, 33]))
        m = Model()
        width = rng.uniform(1.0, 3.0, int(rng.integers(1, 4)))
        hy = float(rng.uniform(2.0, 6.0))
        x = _slab_geometry(m, rng, G, width, hy, fission_free=False, nu_fission_ratio=0.05,
                           c_range=(0.3, 0.6))
        m.set(mode=abi.MODIFIED_FIXED_SOURCE, nparticles=per_batch,
              energy_bounds=np.arange(G + 1.0), tracking=abi.CARTER_TRACKING,
              sample_xs_ratio=[0.8] * G)
        m.cancellation_mesh([float(x[0]), -hy, -hy], [float(x[-1]), hy, hy], [10, 10, 10])
        return m, per_batch * nbatches, "carter"
    m, _ = slab_problem(rng, nhist=nhist, nbatch_hist=per_batch)
    return m, per_batch * nbatches, "fixed"


def subcritical_slab(nparticles=20000, tracking="carter", cancel=True, half=1.0):
    """Modified-fixed-source run in a subcritical one-group box of the
    PUa-1-0-SL material (2 half x 10 x 10 cm, vacuum on every face): a point
    source at the centre, fission generations followed until the chain dies.
    Carter tracking samples flights with 0.8 x the majorant, so weights go
    negative and split; `cancel` adds approximate mesh cancellation on a
    10 x 10 x 10 mesh.  A collision flux tally on 20 x 1 x 1."""
    m, _ = la13511("PUa-1-0-SL", nparticles, tracking="surface")
    m.surfaces = [(t, [(-half if p[0] < 0 else half)] if t == abi.XPLANE else p, abi.VACUUM, sid)
                  for (t, p, bc, sid) in m.surfaces]
    m.mesh_tally([-half, -5., -5.], [half, 5., 5.], [20, 1, 1], [0., 1.],
                 estimator="collision")
    m.set(mode=abi.MODIFIED_FIXED_SOURCE, tracking=_tracking(tracking))
    if tracking == "carter":
        m.set(sample_xs_ratio=[0.8])
    if cancel:
        m.cancellation_mesh([-half, -5., -5.], [half, 5., 5.], [10, 10, 10])
    return m


def anisotropic_slab(nparticles=20000, tracking="surface", moments=(0.3, 0.1)):
    """The PUa-1-0-SL critical slab with P1/P2 Legendre scattering moments,
    linearised like LegendreDistribution::linearize (the scattering tables
    of an anisotropic multigroup material)."""
    from .deck import linearize_legendre
    m, _ = la13511("PUa-1-0-SL", nparticles, tracking=tracking)
    m.materials[0]["angles"] = [linearize_legendre(list(moments))]
    return m


def exact_slab(nparticles=4000, tracking="carter", ratio=0.85, moments=((0.25, 0.05), (0.1, 0.0))):
    """A three-group k-eigenvalue slab for exact regional cancellation:
    fuel | moderator | fuel across x (2.5, 3 and 2.5 cm, vacuum at x = +-4),
    reflective at y, z = +-5.  The fuel's fission spectrum depends on the
    incident group (a chi matrix); the moderator scatters anisotropically,
    with linearised Legendre tables per (g_in, g_out) pair (P1, P2 moments
    `moments[0]` in group, `moments[1]` out of group).  Carter tracking on
    `ratio` x the majorant gives negative weights."""
    from .deck import linearize_legendre
    G = 3
    m = Model()
    sa_f = np.array([0.08, 0.12, 0.22])
    sc_f = np.array([[0.15, 0.07, 0.01], [0.0, 0.20, 0.06], [0.0, 0.01, 0.30]])
    m.material(1, total=sa_f + sc_f.sum(axis=1), absorption=sa_f, scatter=sc_f,
               fission=np.array([0.07, 0.11, 0.20]), nu=np.full(G, 2.5),
               chi=np.array([[0.8, 0.2, 0.0], [0.7, 0.25, 0.05], [0.6, 0.3, 0.1]]))
    sa_m = np.array([0.10, 0.15, 0.25])
    sc_m = np.array([[0.20, 0.12, 0.02], [0.0, 0.30, 0.15], [0.0, 0.05, 0.55]])
    m.material(2, total=sa_m + sc_m.sum(axis=1), absorption=sa_m, scatter=sc_m)
    m.materials[1]["angles"] = [linearize_legendre(list(moments[0] if gi == go else moments[1]))
                                for gi in range(G) for go in range(G)]
    xs = [-4.0, -1.5, 1.5, 4.0]
    sx = [m.surface("xplane", [x], "vacuum" if i in (0, 3) else "normal")
          for i, x in enumerate(xs)]
    sy = [m.surface("yplane", [-5.], "reflective"), m.surface("yplane", [5.], "reflective")]
    sz = [m.surface("zplane", [-5.], "reflective"), m.surface("zplane", [5.], "reflective")]
    box = f"+{sy[0]} & -{sy[1]} & +{sz[0]} & -{sz[1]}"
    cells = [m.cell(f"+{sx[i]} & -{sx[i + 1]} & {box}", material=(2 if i == 1 else 1), id=i + 1)
             for i in range(3)]
    m.root = m.cell_universe(cells, id=1)
    m.point_source([-2.5, 0., 0.], energy=2.5)
    m.set(mode=abi.K_EIGENVALUE, nparticles=nparticles, energy_bounds=[0., 1., 2., 3.],
          tracking=_tracking(tracking))
    if tracking == "carter":
        m.set(sample_xs_ratio=[ratio] * G)
    return m

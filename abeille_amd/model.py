"""Problem builder: the multigroup subset of an Abeille input deck as Python
objects, flattened into the C description of include/abx.h.

The object model follows the deck (src/parser.cpp:77-1006): surfaces with
ids and boundary conditions, cells with region strings, cell universes,
rectilinear and hexagonal lattices referring to each other by id,
multigroup materials, sources, mesh tallies and settings.  `Model.desc()` resolves ids to indices
and returns an `abx_problem_desc` (ctypes) that stays valid while the model
object is alive.
"""
import ctypes as C

import numpy as np

from . import abi


def hex_tiles(rings):
    """Axial (q, r) of one layer of a hexagonal lattice of `rings` rings in
    the order its universes are listed: rows of r ascending, q ascending
    within a row, tiles outside the rings skipped (src/hex_lattice.cpp:
    219-243)."""
    mid = rings - 1
    return [(q, r) for r in range(-mid, mid + 1) for q in range(-mid, mid + 1)
            if max(abs(q), abs(r), abs(q + r)) < rings]


class Model:
    def __init__(self):
        self.surfaces = []      # (type, params, boundary, id)
        self.cells = []         # dict(id, region, material|universe)
        self.universes = []     # dict
        self.materials = []     # dict
        self.sources = []       # dict
        self.tallies = []       # dict
        self.root = None
        self.cancel = None
        self.exact = None
        self.settings = dict(mode=abi.K_EIGENVALUE, tracking=abi.SURFACE_TRACKING,
                             seed=abi.DEFAULT_SEED, stride=abi.DEFAULT_STRIDE,
                             wgt_cutoff=0.25, wgt_survival=1.0, wgt_split=2.0,
                             nparticles=100000, energy_bounds=None,
                             precision=abi.FP64, sample_xs_ratio=None)
        self._keep = []

    # ---- deck-like API -------------------------------------------------
    def surface(self, kind, params, boundary="normal", id=None):
        sid = id if id is not None else len(self.surfaces) + 1
        bc = {"vacuum": abi.VACUUM, "reflective": abi.REFLECTIVE,
              "normal": abi.NORMAL}[boundary]
        self.surfaces.append((abi.SURFACE_TYPES[kind], list(params), bc, sid))
        return sid

    def cell(self, region, material=None, universe=None, id=None):
        cid = id if id is not None else len(self.cells) + 1
        if (material is None) == (universe is None):
            raise ValueError("cell needs exactly one of material / universe")
        self.cells.append(dict(id=cid, region=region, material=material,
                               universe=universe))
        return cid

    def cell_universe(self, cells, id=None):
        uid = id if id is not None else len(self.universes) + 1
        self.universes.append(dict(id=uid, type=abi.CELL_UNIVERSE, cells=list(cells)))
        return uid

    def rect_lattice(self, shape, pitch, origin, universes, outer=None, id=None):
        """`universes` lists universe ids (or -1) in the deck's order:
        index = nz*Nx*Ny + nx*Ny + ny (src/rect_lattice.cpp:300)."""
        uid = id if id is not None else len(self.universes) + 1
        if len(universes) != shape[0] * shape[1] * shape[2]:
            raise ValueError("Lattice instance has improper number of universes.")
        self.universes.append(dict(id=uid, type=abi.RECT_LATTICE, shape=list(shape),
                                   pitch=list(pitch), origin=list(origin),
                                   elems=list(universes), outer=outer))
        return uid

    def hex_lattice(self, rings, nz, pitch, pitch_z, origin, universes, top="pointy",
                    outer=None, id=None):
        """Hexagonal lattice of `rings` rings and `nz` axial layers
        (src/hex_lattice.cpp:454-570, deck `type: hexagonal`).  `universes`
        lists universe ids (or -1) in the deck's order: layer by layer, each
        layer's tiles in rows of axial r ascending, q ascending within a row,
        tiles outside the rings skipped (src/hex_lattice.cpp:219-243)."""
        uid = id if id is not None else len(self.universes) + 1
        if rings < 1 or nz < 1:
            raise ValueError("All elements of HexLattice shape must be >= 1.")
        if pitch <= 0. or pitch_z <= 0.:
            raise ValueError("All elements of HexLattice pitch must > 0.")
        if top not in ("pointy", "flat"):
            raise ValueError(f"Uknown top for hexagonal lattice {top}.")
        if len(universes) != (3 * rings * (rings - 1) + 1) * nz:
            raise ValueError("Improper number of universes for HexLattice.")
        self.universes.append(dict(id=uid, type=abi.HEX_LATTICE, shape=[rings, nz, 1],
                                   pitch=[pitch, pitch_z, 0.], origin=list(origin),
                                   top=top, elems=list(universes), outer=outer))
        return uid

    def material(self, id, total, absorption, scatter, fission=None, nu=None,
                 nu_delayed=None, chi=None, delayed_probabilities=None,
                 delayed_constants=None):
        G = len(total)
        chi_arr = None
        if chi is not None:
            chi_arr = np.asarray(chi, dtype=np.float64)
            if chi_arr.ndim == 1:
                chi_arr = chi_arr[None, :]
            if chi_arr.shape[0] == 1:  # single row replicated (src/mg_nuclide.cpp:842-853)
                chi_arr = np.repeat(chi_arr, G, axis=0)
        self.materials.append(dict(
            id=id, total=np.asarray(total, np.float64),
            absorption=np.asarray(absorption, np.float64),
            scatter=np.ascontiguousarray(np.asarray(scatter, np.float64).reshape(G, G)),
            fission=None if fission is None else np.asarray(fission, np.float64),
            nu=None if nu is None else np.asarray(nu, np.float64),
            nu_delayed=None if nu_delayed is None else np.asarray(nu_delayed, np.float64),
            chi=None if chi_arr is None else np.ascontiguousarray(chi_arr),
            dp=None if delayed_probabilities is None else np.asarray(delayed_probabilities, np.float64),
            dc=None if delayed_constants is None else np.asarray(delayed_constants, np.float64)))
        return id

    def point_source(self, position, energy, weight=1.0):
        self.sources.append(dict(spatial=abi.SRC_POINT, low=list(position),
                                 high=[0, 0, 0], energy=energy, weight=weight,
                                 fissile_only=False, direction=abi.DIR_ISOTROPIC,
                                 dir=[1, 0, 0]))

    def box_source(self, low, high, energy, weight=1.0, fissile_only=False):
        self.sources.append(dict(spatial=abi.SRC_BOX, low=list(low), high=list(high),
                                 energy=energy, weight=weight,
                                 fissile_only=fissile_only,
                                 direction=abi.DIR_ISOTROPIC, dir=[1, 0, 0]))

    def mesh_tally(self, low, high, shape, energy_bounds, estimator="collision",
                   quantity="flux"):
        est = {"collision": abi.EST_COLLISION, "track-length": abi.EST_TRACK_LENGTH}[estimator]
        q = {"flux": abi.Q_FLUX, "elastic": abi.Q_ELASTIC, "absorption": abi.Q_ABSORPTION,
             "fission": abi.Q_FISSION, "total": abi.Q_TOTAL}[quantity]
        self.tallies.append(dict(estimator=est, quantity=q, low=list(low),
                                 high=list(high), shape=list(shape),
                                 ebounds=np.asarray(energy_bounds, np.float64)))
        return len(self.tallies) - 1

    def cancellation_mesh(self, low, high, shape, energy_bounds=None):
        """Approximate mesh cancellation of banked particles
        (src/approximate_mesh_cancelator.cpp)."""
        self.cancel = dict(low=list(low), high=list(high), shape=list(shape),
                           ebounds=None if energy_bounds is None
                           else np.asarray(energy_bounds, np.float64))

    def exact_cancellation(self, low, high, shape, kind="exact", n_samples=10, group_bins=None,
                           beta="average-g", sobol=True, chi_matrix=None,
                           use_virtual_collisions=None):
        """ExactMGCancelator ("exact", src/exact_mg_cancelator.cpp) or
        BasicExactMGCancelator ("basic", src/basic_exact_mg_cancelator.cpp)
        regional cancellation of k-eigenvalue generations (delta or carter
        tracking).  `chi_matrix` and `use_virtual_collisions` are the
        reference's settings::chi_matrix / use_virtual_collisions; by default
        they follow from the materials: a chi whose rows differ is a matrix,
        and a zero in-group scattering entry forbids virtual collisions
        (src/mg_nuclide.cpp:639, 713, 808-810).  The deck reader passes the
        deck's own answers."""
        if kind not in ("exact", "basic"):
            raise ValueError(f"Unkown cancelator type \"{kind}\".")
        betas = {"zero": abi.BETA_ZERO, "minimum": abi.BETA_MINIMUM,
                 "average-f": abi.BETA_AVERAGE_F, "average-g": abi.BETA_AVERAGE_G}
        if beta not in betas:
            raise ValueError(f"Unkown beta entry \"{beta}\" for basic exact MG cancelator.")
        if chi_matrix is None:
            chi_matrix = any(m["chi"] is not None and not np.all(m["chi"] == m["chi"][0])
                             for m in self.materials)
        if use_virtual_collisions is None:
            use_virtual_collisions = all(np.all(np.diag(m["scatter"]) != 0.)
                                         for m in self.materials)
        self.exact = dict(kind=kind, low=list(low), high=list(high), shape=list(shape),
                          n_samples=int(n_samples),
                          group_bins=[list(map(int, g)) for g in (group_bins or [])],
                          beta=beta, sobol=bool(sobol), chi_matrix=bool(chi_matrix),
                          use_virtual_collisions=bool(use_virtual_collisions))

    def set(self, **kw):
        for k, v in kw.items():
            if k not in self.settings:
                raise KeyError(k)
            self.settings[k] = v

    @property
    def ngroups(self):
        return len(self.settings["energy_bounds"]) - 1

    # ---- flat JSON form ------------------------------------------------
    def to_dict(self):
        """Plain lists and numbers (JSON-ready): the model as built, ids and
        region strings included, so that `Model.from_dict` rebuilds the same
        description without the deck it was read from."""
        def arr(v):
            return None if v is None else np.asarray(v, np.float64).tolist()

        mats = []
        for m in self.materials:
            d = {k: (arr(v) if k not in ("id", "angles") else v) for k, v in m.items()}
            if m.get("angles"):
                d["angles"] = [[arr(a) for a in t] for t in m["angles"]]
            mats.append(d)
        tallies = [dict(t, ebounds=arr(t["ebounds"])) for t in self.tallies]
        st = dict(self.settings)
        st["energy_bounds"] = arr(st["energy_bounds"])
        st["sample_xs_ratio"] = arr(st["sample_xs_ratio"])
        cancel = None
        if self.cancel is not None:
            cancel = dict(self.cancel, ebounds=arr(self.cancel["ebounds"]))
        d = dict(surfaces=[list(s) for s in self.surfaces], cells=self.cells,
                 universes=self.universes, materials=mats, sources=self.sources,
                 tallies=tallies, root=self.root, cancel=cancel, settings=st)
        if self.exact is not None:
            d["exact_cancel"] = dict(self.exact)
        return d

    @classmethod
    def from_dict(cls, d):
        m = cls()
        m.surfaces = [(int(t), list(p), int(bc), sid) for t, p, bc, sid in d["surfaces"]]
        m.cells = [dict(c) for c in d["cells"]]
        m.universes = [dict(u) for u in d["universes"]]
        for md in d["materials"]:
            mm = {k: (None if v is None else np.asarray(v, np.float64))
                  for k, v in md.items() if k not in ("id", "angles")}
            mm["id"] = md["id"]
            if mm.get("scatter") is not None:
                G = len(mm["total"])
                mm["scatter"] = np.ascontiguousarray(mm["scatter"].reshape(G, G))
            if mm.get("chi") is not None:
                mm["chi"] = np.ascontiguousarray(mm["chi"].reshape(len(mm["total"]), -1))
            if md.get("angles"):
                mm["angles"] = [tuple(np.asarray(a, np.float64) for a in t) for t in md["angles"]]
            m.materials.append(mm)
        m.sources = [dict(s) for s in d["sources"]]
        m.tallies = [dict(t, ebounds=np.asarray(t["ebounds"], np.float64)) for t in d["tallies"]]
        m.root = d["root"]
        if d.get("cancel"):
            c = dict(d["cancel"])
            c["ebounds"] = None if c["ebounds"] is None else np.asarray(c["ebounds"], np.float64)
            m.cancel = c
        m.settings.update(d["settings"])
        if d.get("exact_cancel"):
            m.exact = dict(d["exact_cancel"])
        return m

    # ---- flattening ----------------------------------------------------
    def desc(self):
        keep = []
        sidx = {sid: i for i, (_, _, _, sid) in enumerate(self.surfaces)}
        midx = {m["id"]: i for i, m in enumerate(self.materials)}
        uidx = {u["id"]: i for i, u in enumerate(self.universes)}
        cidx = {c["id"]: i for i, c in enumerate(self.cells)}

        surfs = (abi.Surface * max(1, len(self.surfaces)))()
        for i, (t, p, bc, sid) in enumerate(self.surfaces):
            surfs[i].type = t
            surfs[i].boundary = bc
            surfs[i].id = sid
            for k, v in enumerate(p):
                surfs[i].p[k] = v

        rpn_all = []
        cells = (abi.Cell * max(1, len(self.cells)))()
        for i, c in enumerate(self.cells):
            rpn = abi.infix_to_rpn(abi.tokenize_region(c["region"], sidx))
            cells[i].id = c["id"]
            if c["material"] is not None:
                cells[i].fill_type = abi.FILL_MATERIAL
                cells[i].fill = midx[c["material"]]
            else:
                cells[i].fill_type = abi.FILL_UNIVERSE
                cells[i].fill = uidx[c["universe"]]
            cells[i].rpn_offset = len(rpn_all)
            cells[i].rpn_len = len(rpn)
            rpn_all += rpn
        rpn_arr = np.asarray(rpn_all if rpn_all else [0], np.int32)

        ucells, elems = [], []
        unis = (abi.Universe * max(1, len(self.universes)))()
        for i, u in enumerate(self.universes):
            unis[i].id = u["id"]
            unis[i].type = u["type"]
            unis[i].outer = -1
            unis[i].elems_offset = -1
            if u["type"] == abi.CELL_UNIVERSE:
                unis[i].cells_offset = len(ucells)
                unis[i].n_cells = len(u["cells"])
                ucells += [cidx[c] for c in u["cells"]]
            else:
                for k in range(3):
                    unis[i].shape[k] = u["shape"][k]
                    unis[i].pitch[k] = u["pitch"][k]
                    unis[i].origin[k] = u["origin"][k]
                if u["type"] == abi.HEX_LATTICE:
                    unis[i].hex_top = abi.HEX_FLAT if u["top"] == "flat" else abi.HEX_POINTY
                unis[i].elems_offset = len(elems)
                elems += [(-1 if e == -1 else uidx[e]) for e in u["elems"]]
                unis[i].outer = -1 if u["outer"] is None else uidx[u["outer"]]
        ucell_arr = np.asarray(ucells if ucells else [0], np.int32)
        elem_arr = np.asarray(elems if elems else [0], np.int32)

        mats = (abi.Material * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i].id = m["id"]
            mats[i].total = abi.ptr(m["total"])
            mats[i].absorption = abi.ptr(m["absorption"])
            mats[i].scatter = abi.ptr(m["scatter"])
            mats[i].fission = abi.ptr(m["fission"])
            mats[i].nu_prompt = abi.ptr(m["nu"])
            mats[i].nu_delayed = abi.ptr(m["nu_delayed"])
            mats[i].chi = abi.ptr(m["chi"])
            n_d = 0 if m["dp"] is None else len(m["dp"])
            mats[i].n_delayed = n_d
            mats[i].delayed_probabilities = abi.ptr(m["dp"])
            mats[i].delayed_constants = abi.ptr(m["dc"])
            if m.get("angles"):
                # piecewise-linear cosine tables per (g_in, g_out)
                offs = [0]
                for (mu, _, _) in m["angles"]:
                    offs.append(offs[-1] + len(mu))
                offs = np.asarray(offs, np.int32)
                mu_a = np.concatenate([a[0] for a in m["angles"]])
                pdf_a = np.concatenate([a[1] for a in m["angles"]])
                cdf_a = np.concatenate([a[2] for a in m["angles"]])
                keep += [offs, mu_a, pdf_a, cdf_a]
                mats[i].angle_offsets = abi.ptr(offs, C.c_int32)
                mats[i].angle_mu = abi.ptr(mu_a)
                mats[i].angle_pdf = abi.ptr(pdf_a)
                mats[i].angle_cdf = abi.ptr(cdf_a)
            else:
                mats[i].angle_offsets = abi.ptr(None, C.c_int32)
                mats[i].angle_mu = abi.ptr(None)
                mats[i].angle_pdf = abi.ptr(None)
                mats[i].angle_cdf = abi.ptr(None)

        srcs = (abi.Source * max(1, len(self.sources)))()
        for i, s in enumerate(self.sources):
            srcs[i].spatial = s["spatial"]
            srcs[i].direction = s["direction"]
            srcs[i].fissile_only = 1 if s["fissile_only"] else 0
            for k in range(3):
                srcs[i].low[k] = s["low"][k]
                srcs[i].high[k] = s["high"][k]
                srcs[i].dir[k] = s["dir"][k]
            srcs[i].energy = s["energy"]
            srcs[i].weight = s["weight"]

        tals = (abi.MeshTally * max(1, len(self.tallies)))()
        for i, t in enumerate(self.tallies):
            tals[i].estimator = t["estimator"]
            tals[i].quantity = t["quantity"]
            for k in range(3):
                tals[i].shape[k] = t["shape"][k]
                tals[i].low[k] = t["low"][k]
                tals[i].high[k] = t["high"][k]
            tals[i].n_ebounds = len(t["ebounds"])
            tals[i].ebounds = abi.ptr(t["ebounds"])

        st = self.settings
        eb = np.asarray(st["energy_bounds"], np.float64)
        ratio = None if st["sample_xs_ratio"] is None else np.asarray(st["sample_xs_ratio"], np.float64)
        d = abi.ProblemDesc()
        d.n_surfaces = len(self.surfaces)
        d.surfaces = surfs
        d.n_cells = len(self.cells)
        d.cells = cells
        d.n_rpn = len(rpn_all)
        d.rpn = abi.ptr(rpn_arr, C.c_int32)
        d.n_universes = len(self.universes)
        d.universes = unis
        d.n_universe_cells = len(ucells)
        d.universe_cells = abi.ptr(ucell_arr, C.c_int32)
        d.n_lattice_elems = len(elems)
        d.lattice_elems = abi.ptr(elem_arr, C.c_int32)
        d.root_universe = uidx[self.root]
        d.n_materials = len(self.materials)
        d.materials = mats
        d.n_sources = len(self.sources)
        d.sources = srcs
        d.n_tallies = len(self.tallies)
        d.tallies = tals
        s = d.settings
        s.mode = st["mode"]
        s.tracking = st["tracking"]
        s.seed = st["seed"]
        s.stride = st["stride"]
        s.wgt_cutoff = st["wgt_cutoff"]
        s.wgt_survival = st["wgt_survival"]
        s.wgt_split = st["wgt_split"]
        s.nparticles = st["nparticles"]
        s.ngroups = len(eb) - 1
        s.precision = st["precision"]
        s.energy_bounds = abi.ptr(eb)
        s.sample_xs_ratio = abi.ptr(ratio)
        cm = None
        if self.cancel is not None:
            cm = abi.CancelMesh()
            for k in range(3):
                cm.low[k] = self.cancel["low"][k]
                cm.high[k] = self.cancel["high"][k]
                cm.shape[k] = self.cancel["shape"][k]
            ceb = self.cancel["ebounds"]
            cm.n_ebounds = 0 if ceb is None else len(ceb)
            cm.ebounds = abi.ptr(ceb)
            d.cancel = C.pointer(cm)
        ex = None
        if self.exact is not None:
            e = self.exact
            ex = abi.ExactCancel()
            ex.kind = abi.EXACT_MG if e["kind"] == "exact" else abi.BASIC_EXACT_MG
            for k in range(3):
                ex.shape[k] = e["shape"][k]
                ex.low[k] = e["low"][k]
                ex.high[k] = e["high"][k]
            ex.n_samples = e["n_samples"]
            ex.chi_matrix = 1 if e["chi_matrix"] else 0
            ex.use_virtual_collisions = 1 if e["use_virtual_collisions"] else 0
            ex.beta = {"zero": abi.BETA_ZERO, "minimum": abi.BETA_MINIMUM,
                       "average-f": abi.BETA_AVERAGE_F, "average-g": abi.BETA_AVERAGE_G}[e["beta"]]
            ex.sobol = 1 if e["sobol"] else 0
            offs = np.cumsum([0] + [len(g) for g in e["group_bins"]]).astype(np.int32)
            grp = np.asarray([x for g in e["group_bins"] for x in g] or [0], np.int32)
            ex.n_group_bins = len(e["group_bins"])
            ex.group_bin_offsets = abi.ptr(offs, C.c_int32)
            ex.group_bins = abi.ptr(grp, C.c_int32)
            d.exact_cancel = C.pointer(ex)
            keep += [offs, grp]
        keep += [surfs, cells, rpn_arr, unis, ucell_arr, elem_arr, mats, srcs, tals, eb,
                 ratio, d, cm, ex]
        self._keep = keep
        return d

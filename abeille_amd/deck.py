"""Reads the multigroup subset of an Abeille YAML input deck into a Model.

Follows the reference's parser (src/parser.cpp:77-1006): `materials` with
multigroup data (src/mg_nuclide.cpp:579-922, including P1..P5 Legendre
moments linearised like src/legendre_distribution.cpp:105-155), `surfaces`,
`cells` (region strings), `universes` (cell lists, `rectlinear` or
`hexagonal` lattices), `root-universe`, `sources` (point / box, isotropic /
mono-directional, mono-energetic), `tallies` (mesh tallies with
collision / track-length estimators), `settings`, `sampling-xs-ratio` and a
`cancelator` (approximate, exact or basic-exact, src/cancelator.cpp:32-78).
Keys outside that subset (continuous energy, noise, ...) raise ValueError.
"""
import math

import numpy as np
import yaml

from . import abi
from .model import Model

TOLERANCE = 0.0001  # include/utils/constants.hpp


def _legendre(n, x):
    if n == 0:
        return 1.0
    if n == 1:
        return x
    if n == 2:
        return 0.5 * (3. * x * x - 1.)
    if n == 3:
        return 0.5 * (5. * x * x * x - 3. * x)
    if n == 4:
        return 0.125 * (35. * x ** 4 - 30. * x * x + 3.)
    if n == 5:
        return 0.125 * (63. * x ** 5 - 70. * x ** 3 + 15. * x)
    raise ValueError("Legendre moments above 5 are not read")


def linearize_legendre(moments):
    """Piecewise-linear table (mu, pdf, cdf) of the distribution with
    Legendre moments a_1.. (a_0 = 1), bisecting until linear interpolation
    is within the 1e-4 relative tolerance (LegendreDistribution::linearize)."""
    a = [0.5] + [c * (2. * l + 1.) / 2. for l, c in enumerate(moments, start=1)]

    def pdf(mu):
        return sum(a[l] * _legendre(l, mu) for l in range(len(a)))

    mu = [-1., 1.]
    p = [pdf(-1.), pdf(1.)]
    i = 0
    while i < len(mu) - 1:
        mid = 0.5 * (mu[i] + mu[i + 1])
        p_interp = 0.5 * (p[i] + p[i + 1])
        p_real = pdf(mid)
        if abs(p_interp - p_real) / p_real > TOLERANCE:
            mu.insert(i + 1, mid)
            p.insert(i + 1, p_real)
        else:
            i += 1
    cdf = [0.] * len(mu)
    for i in range(len(mu) - 1):
        cdf[i + 1] = ((mu[i + 1] - mu[i]) * 0.5 * (p[i + 1] + p[i])) + cdf[i]
    norm = cdf[-1]
    return (np.array(mu), np.array(p) / norm, np.array(cdf) / norm)


def _vec(node, key, G, name):
    v = node.get(key)
    if v is None or len(v) != G:
        raise ValueError(f"Invalid {key} entry in material {name}.")
    return [float(x) for x in v]


def _material(m, node, G):
    """Adds the material; returns (chi given as a matrix, virtual collisions
    allowed), the flags MGNuclide's reader sets for exact cancellation
    (src/mg_nuclide.cpp:639, 713, 808-810)."""
    mid = node["id"]
    total = _vec(node, "total", G, mid)
    absorption = _vec(node, "absorption", G, mid)
    scatter = node.get("scatter")
    if scatter is None or len(scatter) != G or any(len(r) != G for r in scatter):
        raise ValueError(f"Invalid scatter matrix entry in material {mid}.")
    kw = {}
    if "fission" in node:
        kw["fission"] = _vec(node, "fission", G, mid)
        if "nu" in node:
            kw["nu"] = _vec(node, "nu", G, mid)
        elif "nu_prompt" in node and "nu_delayed" in node:
            kw["nu"] = _vec(node, "nu_prompt", G, mid)
            kw["nu_delayed"] = _vec(node, "nu_delayed", G, mid)
        if "chi" in node:
            kw["chi"] = node["chi"]
    if "delayed_groups" in node:
        dg = node["delayed_groups"]
        kw["delayed_probabilities"] = dg["probabilities"]
        kw["delayed_constants"] = dg["constants"]
    m.material(mid, total=total, absorption=absorption, scatter=scatter, **kw)
    fissile = "fission" in node and any(float(v) > 0. for v in kw["fission"])
    chi_matrix = fissile and "chi" in node and len(node["chi"]) == G
    virtual_ok = all(float(scatter[g][g]) != 0. for g in range(G))
    moments = {}
    for l in range(1, 6):
        key = f"P{l}"
        if key in node:
            moments[l] = np.asarray(node[key], float)
    if moments:
        angles = []
        for gi in range(G):
            for go in range(G):
                coeffs = [moments[l][gi][go] if l in moments else 0.0
                          for l in range(1, max(moments) + 1)]
                angles.append(linearize_legendre(coeffs))
        m.materials[-1]["angles"] = angles
        for g in range(G):  # in-group pdf at mu = 1
            mu, pdf, _ = angles[g * G + g]
            if float(np.interp(1.0, mu, pdf)) == 0.:
                virtual_ok = False
    return chi_matrix, virtual_ok


def _triple(c, key, what):
    v = c.get(key)
    if not isinstance(v, (list, tuple)) or len(v) != 3:
        raise ValueError(f"No valid {key} entry for {what}.")
    return [float(x) for x in v]


def _cancelator(m, c, tracking, chi_matrix, virtual_ok):
    """make_cancelator (src/cancelator.cpp:32-78) and the exact cancelators'
    readers (src/exact_mg_cancelator.cpp:595-680,
    src/basic_exact_mg_cancelator.cpp:625-710)."""
    if not c or not isinstance(c.get("type"), str):
        raise ValueError("Invalid type entry for cancelator.")
    t = c["type"]
    if t == "approximate":
        m.cancellation_mesh(c["low"], c["hi"], c["shape"], c.get("energy-bounds"))
        return
    if t not in ("exact", "basic-exact"):
        raise ValueError(f"Unkown cancelator type \"{t}\".")
    if tracking == abi.SURFACE_TRACKING:
        raise ValueError(("basic-exect" if t == "basic-exact" else "exect") +
                         " cancelators may not be used with surface-tracking.")
    what = "basic exact MG cancelator"
    low, hi = _triple(c, "low", what), _triple(c, "hi", what)
    shape = c.get("shape")
    if not isinstance(shape, (list, tuple)) or len(shape) != 3:
        raise ValueError(f"No valid shape entry for {what}.")
    n_samples = c.get("n-samples", 10)
    if not isinstance(n_samples, int):
        raise ValueError("Invalid n-samples entry for cancelator.")
    if n_samples == 0:
        raise ValueError("n-samples must be greater than zero.")
    if t == "exact":
        gb = c.get("group-bins")
        gb = [[int(g) for g in b] for b in gb] if isinstance(gb, list) else None
        m.exact_cancellation(low, hi, [int(x) for x in shape], kind="exact",
                             n_samples=n_samples, group_bins=gb, chi_matrix=chi_matrix,
                             use_virtual_collisions=virtual_ok)
        return
    beta = c.get("beta")
    if not isinstance(beta, str):
        raise ValueError(f"No valid beta entry for {what}.")
    sobol = c.get("sobol", True)
    if not isinstance(sobol, bool):
        raise ValueError("Invalid sobol entry for cancelator.")
    m.exact_cancellation(low, hi, [int(x) for x in shape], kind="basic", n_samples=n_samples,
                         beta=beta, sobol=sobol, chi_matrix=chi_matrix,
                         use_virtual_collisions=virtual_ok)


SURFACE_PARAMS = {
    "xplane": ["x0"], "yplane": ["y0"], "zplane": ["z0"],
    "plane": ["A", "B", "C", "D"],
    "xcylinder": ["y0", "z0", "r"], "ycylinder": ["x0", "z0", "r"],
    "zcylinder": ["x0", "y0", "r"], "sphere": ["x0", "y0", "z0", "r"],
    "cylinder": ["x0", "y0", "z0", "u0", "v0", "w0", "r"],
}


def model_from_deck(deck):
    """Build a Model from a parsed deck (dict) or a YAML file path."""
    if isinstance(deck, str):
        with open(deck) as f:
            deck = yaml.safe_load(f)
    st = deck.get("settings", {})
    if st.get("energy-mode", "multi-group") not in ("multi-group",):
        raise ValueError("only multi-group decks are supported")
    G = int(st["ngroups"])
    ebounds = [float(x) for x in st["energy-bounds"]]
    m = Model()
    chi_matrix, virtual_ok = False, True
    for node in deck["materials"]:
        cm, vo = _material(m, node, G)
        chi_matrix = chi_matrix or cm
        virtual_ok = virtual_ok and vo
    for node in deck["surfaces"]:
        t = node["type"]
        if t not in SURFACE_PARAMS:
            raise ValueError(f"Surface type \"{t}\" is not supported.")
        m.surface(t, [float(node[k]) for k in SURFACE_PARAMS[t]],
                  node.get("boundary", "normal"), id=int(node["id"]))
    for node in deck["cells"]:
        if "material" in node:
            m.cell(node["region"], material=int(node["material"]), id=int(node["id"]))
        else:
            m.cell(node["region"], universe=int(node["universe"]), id=int(node["id"]))
    for node in deck["universes"]:
        if "cells" in node:
            m.cell_universe([int(c) for c in node["cells"]], id=int(node["id"]))
        else:
            outer = node.get("outer", -1)
            outer = None if outer == -1 else int(outer)
            unis = [int(u) for u in node["universes"]]
            if node.get("type") == "rectlinear":
                m.rect_lattice(node["shape"], node["pitch"], node["origin"], unis,
                               outer=outer, id=int(node["id"]))
            elif node.get("type") == "hexagonal":
                # src/hex_lattice.cpp:454-570
                shape, pitch = node["shape"], node["pitch"]
                if len(shape) != 2:
                    raise ValueError("Lattice must have a valid shape.")
                if len(pitch) != 2:
                    raise ValueError("Lattice must have a valid pitch.")
                m.hex_lattice(int(shape[0]), int(shape[1]), float(pitch[0]), float(pitch[1]),
                              [float(x) for x in node["origin"]], unis,
                              top=str(node.get("top", "pointy")), outer=outer,
                              id=int(node["id"]))
            else:
                raise ValueError(f"Unknown lattice type {node.get('type')}.")
    m.root = int(deck["root-universe"])
    for node in deck.get("sources", []):
        sp, di, en = node["spatial"], node["direction"], node["energy"]
        if en["type"] != "mono-energetic":
            raise ValueError("only mono-energetic sources are supported")
        if sp["type"] == "point":
            m.point_source(sp["position"], float(en["energy"]), float(node["weight"]))
        elif sp["type"] == "box":
            m.box_source(sp["low"], sp["hi"], float(en["energy"]), float(node["weight"]),
                         fissile_only=bool(node.get("fissile-only", sp.get("fissile-only", False))))
        else:
            raise ValueError(f"spatial distribution {sp['type']} is not supported")
        if di["type"] == "mono-directional":
            m.sources[-1]["direction"] = abi.DIR_MONO
            m.sources[-1]["dir"] = [float(x) for x in di["direction"]]
        elif di["type"] != "isotropic":
            raise ValueError(f"direction distribution {di['type']} is not supported")
    for node in deck.get("tallies", []):
        m.mesh_tally(node["low"], node["hi"], node["shape"],
                     node.get("energy-bounds", [ebounds[0], ebounds[-1]]),
                     estimator=node.get("estimator", "collision"),
                     quantity=node.get("quantity", "flux"))
    mode = {"k-eigenvalue": abi.K_EIGENVALUE, "fixed-source": abi.FIXED_SOURCE,
            "modified-fixed-source": abi.MODIFIED_FIXED_SOURCE}[st.get("simulation", "k-eigenvalue")]
    tracking = {"surface-tracking": abi.SURFACE_TRACKING, "delta-tracking": abi.DELTA_TRACKING,
                "carter-tracking": abi.CARTER_TRACKING}[st.get("transport", "surface-tracking")]
    m.set(mode=mode, tracking=tracking, nparticles=int(st.get("nparticles", 100000)),
          energy_bounds=ebounds, seed=int(st.get("seed", abi.DEFAULT_SEED)),
          stride=int(st.get("stride", abi.DEFAULT_STRIDE)),
          wgt_cutoff=float(st.get("wgt-cutoff", 0.25)),
          wgt_survival=float(st.get("wgt-survival", 1.0)),
          wgt_split=float(st.get("wgt-split", 2.0)))
    if tracking == abi.CARTER_TRACKING:
        ratio = deck.get("sampling-xs-ratio")
        if ratio is None or any(v <= 0. for v in ratio):
            raise ValueError("Must provide positive \"sampling-xs-ratio\" for carter tracking.")
        m.set(sample_xs_ratio=[float(v) for v in ratio])
    if st.get("cancellation", False):
        _cancelator(m, deck.get("cancelator"), tracking, chi_matrix, virtual_ok)
    m.generations = int(st.get("ngenerations", 1))
    m.ignored = int(st.get("nignored", 0))
    return m

"""ctypes mirror of include/abx.h (the C ABI of the transport path).

Only plain data lives here: structure layouts, enum values and a builder
that turns Python-side geometry/material lists into an `abx_problem_desc`
whose arrays stay alive as long as the builder does.
"""
import ctypes as C

import numpy as np

ABX_OK = 0
ERRORS = {
    -1: "ABX_E_INVALID",
    -2: "ABX_E_HIP",
    -3: "ABX_E_LOST",
    -4: "ABX_E_MAJORANT",
    -5: "ABX_E_CAPACITY",
    -6: "ABX_E_NO_DEVICE",
    -7: "ABX_E_BUDGET",
}
ERRORS_BY_NAME = {v: k for k, v in ERRORS.items()}

# surfaces
XPLANE, YPLANE, ZPLANE, PLANE, XCYL, YCYL, ZCYL, SPHERE, CYL = range(9)
SURFACE_TYPES = {
    "xplane": XPLANE, "yplane": YPLANE, "zplane": ZPLANE, "plane": PLANE,
    "xcylinder": XCYL, "ycylinder": YCYL, "zcylinder": ZCYL, "sphere": SPHERE,
    "cylinder": CYL,
}
VACUUM, REFLECTIVE, NORMAL = 0, 1, 2
OP_UNIN = 2147483647 - 4
OP_INTR = 2147483647 - 3
OP_COMP = 2147483647 - 2
OP_R_PAR = 2147483647 - 1
OP_L_PAR = 2147483647
FILL_MATERIAL, FILL_UNIVERSE = 0, 1
CELL_UNIVERSE, RECT_LATTICE, HEX_LATTICE = 0, 1, 2
HEX_POINTY, HEX_FLAT = 0, 1
SRC_POINT, SRC_BOX = 0, 1
DIR_ISOTROPIC, DIR_MONO = 0, 1
EST_COLLISION, EST_TRACK_LENGTH = 0, 1
Q_FLUX, Q_ELASTIC, Q_ABSORPTION, Q_FISSION, Q_TOTAL = 0, 1, 2, 3, 4
K_EIGENVALUE, FIXED_SOURCE, MODIFIED_FIXED_SOURCE = 0, 1, 2
SURFACE_TRACKING, DELTA_TRACKING, CARTER_TRACKING = 0, 1, 2
FP64, SPLIT = 0, 1
ENGINE_EVENT, ENGINE_HISTORY = 0, 1
EV_VIRTUAL, EV_COLLISION, EV_CROSS, EV_REFLECT, EV_LEAK, EV_DEATH, EV_RESURRECT = range(7)

DEFAULT_SEED = 19073486328125
DEFAULT_STRIDE = 152917

PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)


class Surface(C.Structure):
    _fields_ = [("type", C.c_int32), ("boundary", C.c_int32), ("id", C.c_uint32),
                ("pad_", C.c_int32), ("p", C.c_double * 8)]


class Cell(C.Structure):
    _fields_ = [("id", C.c_uint32), ("fill_type", C.c_int32), ("fill", C.c_int32),
                ("rpn_offset", C.c_int32), ("rpn_len", C.c_int32)]


class Universe(C.Structure):
    _fields_ = [("id", C.c_uint32), ("type", C.c_int32), ("cells_offset", C.c_int32),
                ("n_cells", C.c_int32), ("shape", C.c_int32 * 3),
                ("elems_offset", C.c_int32), ("outer", C.c_int32), ("hex_top", C.c_int32),
                ("pitch", C.c_double * 3), ("origin", C.c_double * 3)]


class Material(C.Structure):
    _fields_ = [("id", C.c_uint32), ("n_delayed", C.c_int32),
                ("total", PD), ("absorption", PD), ("scatter", PD), ("fission", PD),
                ("nu_prompt", PD), ("nu_delayed", PD), ("chi", PD),
                ("delayed_probabilities", PD), ("delayed_constants", PD),
                ("angle_offsets", PI32), ("angle_mu", PD), ("angle_pdf", PD),
                ("angle_cdf", PD)]


class Source(C.Structure):
    _fields_ = [("spatial", C.c_int32), ("direction", C.c_int32),
                ("fissile_only", C.c_int32), ("pad_", C.c_int32),
                ("low", C.c_double * 3), ("high", C.c_double * 3),
                ("dir", C.c_double * 3), ("energy", C.c_double), ("weight", C.c_double)]


class MeshTally(C.Structure):
    _fields_ = [("estimator", C.c_int32), ("quantity", C.c_int32),
                ("shape", C.c_int32 * 3), ("n_ebounds", C.c_int32),
                ("low", C.c_double * 3), ("high", C.c_double * 3), ("ebounds", PD)]


class Settings(C.Structure):
    _fields_ = [("mode", C.c_int32), ("tracking", C.c_int32), ("seed", C.c_uint64),
                ("stride", C.c_uint64), ("wgt_cutoff", C.c_double),
                ("wgt_survival", C.c_double), ("wgt_split", C.c_double),
                ("nparticles", C.c_int64), ("ngroups", C.c_int32),
                ("precision", C.c_int32), ("energy_bounds", PD),
                ("sample_xs_ratio", PD)]


class CancelMesh(C.Structure):
    _fields_ = [("low", C.c_double * 3), ("high", C.c_double * 3), ("shape", C.c_int32 * 3),
                ("n_ebounds", C.c_int32), ("ebounds", PD)]


EXACT_MG, BASIC_EXACT_MG = 1, 2
BETA_ZERO, BETA_MINIMUM, BETA_AVERAGE_F, BETA_AVERAGE_G = 0, 1, 2, 3


class ExactCancel(C.Structure):
    _fields_ = [("kind", C.c_int32), ("shape", C.c_int32 * 3),
                ("low", C.c_double * 3), ("high", C.c_double * 3),
                ("n_samples", C.c_int32), ("chi_matrix", C.c_int32),
                ("use_virtual_collisions", C.c_int32), ("beta", C.c_int32),
                ("sobol", C.c_int32), ("n_group_bins", C.c_int32),
                ("group_bin_offsets", PI32), ("group_bins", PI32)]


class ProblemDesc(C.Structure):
    _fields_ = [("n_surfaces", C.c_int32), ("surfaces", C.POINTER(Surface)),
                ("n_cells", C.c_int32), ("cells", C.POINTER(Cell)),
                ("n_rpn", C.c_int32), ("rpn", PI32),
                ("n_universes", C.c_int32), ("universes", C.POINTER(Universe)),
                ("n_universe_cells", C.c_int32), ("universe_cells", PI32),
                ("n_lattice_elems", C.c_int32), ("lattice_elems", PI32),
                ("root_universe", C.c_int32),
                ("n_materials", C.c_int32), ("materials", C.POINTER(Material)),
                ("n_sources", C.c_int32), ("sources", C.POINTER(Source)),
                ("n_tallies", C.c_int32), ("tallies", C.POINTER(MeshTally)),
                ("settings", Settings), ("cancel", C.POINTER(CancelMesh)),
                ("exact_cancel", C.POINTER(ExactCancel))]


class Particle(C.Structure):
    _fields_ = [("r", C.c_double * 3), ("u", C.c_double * 3), ("E", C.c_double),
                ("wgt", C.c_double), ("wgt2", C.c_double), ("history_id", C.c_uint64),
                ("family_id", C.c_uint64), ("rng_state", C.c_uint64)]


class Banked(C.Structure):
    _fields_ = [("r", C.c_double * 3), ("u", C.c_double * 3), ("E", C.c_double),
                ("wgt", C.c_double), ("wgt2", C.c_double),
                ("parent_history_id", C.c_uint64), ("parent_daughter_id", C.c_uint64),
                ("family_id", C.c_uint64)]


class Scores(C.Structure):
    _fields_ = [("k_col", C.c_double), ("k_abs", C.c_double), ("k_trk", C.c_double),
                ("k_tot", C.c_double), ("leakage", C.c_double), ("mig", C.c_double)]


class Diag(C.Structure):
    _fields_ = [("histories", C.c_uint64), ("collisions", C.c_uint64),
                ("virtual_collisions", C.c_uint64), ("flights", C.c_uint64),
                ("crossings", C.c_uint64), ("tally_updates", C.c_uint64),
                ("lost", C.c_uint64), ("secondaries", C.c_uint64),
                ("error", C.c_int32), ("pad_", C.c_int32), ("error_history", C.c_uint64)]


class RunStats(C.Structure):
    _fields_ = [("kcol_avg", C.c_double), ("kcol_err", C.c_double),
                ("kabs_avg", C.c_double), ("kabs_err", C.c_double),
                ("ktrk_avg", C.c_double), ("ktrk_err", C.c_double),
                ("ktot_avg", C.c_double), ("ktot_err", C.c_double),
                ("leak_avg", C.c_double), ("leak_err", C.c_double),
                ("mig_avg", C.c_double), ("mig_err", C.c_double),
                ("active_generations", C.c_int32), ("generations_run", C.c_int32),
                ("transport_seconds", C.c_double), ("kernel_seconds", C.c_double),
                ("kernel_launches", C.c_int64), ("kernel_items", C.c_int64),
                ("histories", C.c_int64), ("diag", Diag)]


class TraceRec(C.Structure):
    _fields_ = [("r", C.c_double * 3), ("u", C.c_double * 3), ("wgt", C.c_double),
                ("history_id", C.c_uint64), ("secondary_id", C.c_int32),
                ("event", C.c_int32), ("cell", C.c_int32), ("material", C.c_int32),
                ("group", C.c_int32), ("rng_calls", C.c_int32)]


# numpy views of the record types (same layout as the C structs)
PARTICLE_DTYPE = np.dtype([("r", "<f8", 3), ("u", "<f8", 3), ("E", "<f8"), ("wgt", "<f8"),
                           ("wgt2", "<f8"), ("history_id", "<u8"), ("family_id", "<u8"),
                           ("rng_state", "<u8")])
BANKED_DTYPE = np.dtype([("r", "<f8", 3), ("u", "<f8", 3), ("E", "<f8"), ("wgt", "<f8"),
                         ("wgt2", "<f8"), ("parent_history_id", "<u8"),
                         ("parent_daughter_id", "<u8"), ("family_id", "<u8")])
PARENT_DTYPE = np.dtype([("r1", "<f8", 3), ("u1", "<f8", 3), ("E1", "<f8"), ("E3", "<f8"),
                         ("Esmp", "<f8"), ("prev_virtual", "<i4"), ("pad_", "<i4")])
TRACE_DTYPE = np.dtype([("r", "<f8", 3), ("u", "<f8", 3), ("wgt", "<f8"),
                        ("history_id", "<u8"), ("secondary_id", "<i4"), ("event", "<i4"),
                        ("cell", "<i4"), ("material", "<i4"), ("group", "<i4"),
                        ("rng_calls", "<i4")])
assert PARTICLE_DTYPE.itemsize == C.sizeof(Particle)
assert BANKED_DTYPE.itemsize == C.sizeof(Banked)
assert TRACE_DTYPE.itemsize == C.sizeof(TraceRec)


def ptr(arr, ctype=C.c_double):
    """Pointer to a contiguous numpy array (or NULL for None)."""
    if arr is None:
        return C.cast(None, C.POINTER(ctype))
    return arr.ctypes.data_as(C.POINTER(ctype))


def infix_to_rpn(infix):
    """Shunting-yard conversion of a tokenised cell region, following
    infix_to_rpn (src/cell.cpp:234-276): operators are compared by their
    numeric codes (COMP binds tighter than INTR, INTR than UNIN)."""
    rpn, stack = [], []
    for tok in infix:
        if tok < OP_UNIN:
            rpn.append(tok)
        elif tok < OP_R_PAR:
            while stack:
                op = stack[-1]
                if op < OP_R_PAR and ((tok == OP_COMP and tok < op) or
                                      (tok != OP_COMP and tok <= op)):
                    rpn.append(op)
                    stack.pop()
                else:
                    break
            stack.append(tok)
        elif tok == OP_L_PAR:
            stack.append(tok)
        else:
            while stack and stack[-1] != OP_L_PAR:
                rpn.append(stack.pop())
            if not stack:
                raise ValueError("Mismatched parentheses in cell region definition.")
            stack.pop()
    while stack:
        op = stack.pop()
        if op >= OP_R_PAR:
            raise ValueError("Mismatched parentheses in cell region definition.")
        rpn.append(op)
    return rpn


def tokenize_region(region, surface_index_of_id):
    """Tokenise a region string such as "+1 & -2 & ~(3 U -4)" the way
    make_cell does (src/cell.cpp:296-338): signed surface ids become
    +-(index + 1)."""
    out, tmp = [], ""

    def flush():
        nonlocal tmp
        if tmp:
            sid = int(tmp)
            idx = surface_index_of_id[abs(sid)] + 1
            out.append(-idx if sid < 0 else idx)
            tmp = ""

    for ch in region:
        if ch in "&()U~":
            flush()
            out.append({"&": OP_INTR, "(": OP_L_PAR, ")": OP_R_PAR,
                        "U": OP_UNIN, "~": OP_COMP}[ch])
        elif ch in "+-0123456789":
            tmp += ch
        elif ch != " ":
            raise ValueError("Invalid character in cell region definition.")
    flush()
    return out


class Tuning(C.Structure):
    _fields_ = [("block_size", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("stage_tables", C.c_int32), ("tail_threshold", C.c_int32),
                ("sort_every", C.c_int32)]

# abx_gen_quantity: per-generation vectors (abx_gen_values)
GEN_KCOL, GEN_KABS, GEN_KTRK, GEN_KTOT, GEN_LEAKAGE, GEN_MIG = range(6)
GEN_NAMES = ("kcol", "kabs", "ktrk", "ktot", "leakage", "mig-area")

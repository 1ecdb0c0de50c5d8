"""abeille_amd - MI355X-native multigroup transport hot path.

Python handle on the C ABI (include/abx.h) exported by the in-tree gfx950
library `abeille_amd/libabx.so`.  There is no CPU fallback: if the library
is missing or no GPU is visible, constructing a `Problem` raises.
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .model import Model  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libabx.so")
_lib = None


class AbxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.ERRORS.get(code, code)}: {msg}")
        self.code = code


def lib():
    """Load libabx.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                               "(python -m abeille_amd.build) first")
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.abx_last_error.restype = C.c_char_p
        L.abx_version.restype = C.c_char_p
        L.abx_synchronize.argtypes = []
        L.abx_create.argtypes = [C.POINTER(abi.ProblemDesc), C.POINTER(P)]
        L.abx_destroy.argtypes = [P]
        L.abx_destroy.restype = None
        L.abx_set_engine.argtypes = [P, C.c_int]
        L.abx_set_tuning.argtypes = [P, C.POINTER(abi.Tuning)]
        L.abx_get_tuning.argtypes = [P, C.POINTER(abi.Tuning)]
        L.abx_time_transport.argtypes = [P, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.abx_sample_sources.argtypes = [P, C.c_uint64, C.c_size_t, C.c_void_p]
        L.abx_transport.argtypes = [P, C.c_void_p, C.c_size_t, C.c_double, C.c_int, C.c_void_p,
                                    C.c_size_t, C.POINTER(C.c_size_t), C.POINTER(abi.Scores),
                                    C.POINTER(abi.Diag)]
        L.abx_transport_trace.argtypes = [P, C.c_void_p, C.c_size_t, C.c_double, C.c_int,
                                          C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                          C.POINTER(abi.Scores), C.c_void_p, C.c_int,
                                          C.c_void_p]
        L.abx_tally_size.argtypes = [P, C.c_int, C.POINTER(C.c_size_t)]
        L.abx_tally_fetch.argtypes = [P, C.c_int, C.c_void_p]
        L.abx_tally_clear.argtypes = [P]
        L.abx_tally_record.argtypes = [P, C.c_double]
        L.abx_tally_stats.argtypes = [P, C.c_int, C.c_void_p, C.c_void_p]
        L.abx_run.argtypes = [P, C.c_int, C.c_int, C.c_void_p, C.POINTER(abi.RunStats)]
        L.abx_run_continue.argtypes = [P, C.c_int, C.c_int, C.c_void_p,
                                       C.POINTER(abi.RunStats)]
        L.abx_request_stop.argtypes = [P]
        L.abx_set_max_time.argtypes = [P, C.c_double]
        L.abx_cancel_bank.argtypes = [P, C.c_void_p, C.c_size_t]
        L.abx_parent_fetch.argtypes = [P, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.abx_exact_cancel_bank.argtypes = [P, C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p,
                                            C.c_size_t, C.POINTER(C.c_size_t)]
        L.abx_rng_reset.argtypes = [P]
        L.abx_rng_get_state.argtypes = [P, C.POINTER(C.c_uint64)]
        L.abx_rng_set_state.argtypes = [P, C.c_uint64]
        L.abx_bank_size.argtypes = [P, C.POINTER(C.c_size_t)]
        L.abx_bank_fetch.argtypes = [P, C.c_void_p]
        L.abx_bank_load.argtypes = [P, C.c_void_p, C.c_size_t]
        L.abx_batch_create.argtypes = [C.c_size_t, C.POINTER(P)]
        L.abx_batch_destroy.argtypes = [P]
        L.abx_batch_destroy.restype = None
        L.abx_problem_footprint.argtypes = [P, C.POINTER(C.c_size_t)]
        L.abx_batch_add.argtypes = [P, P, C.c_int]
        L.abx_batch_step.argtypes = [P, C.POINTER(P), C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)]
        L.abx_batch_kernel_seconds.argtypes = [P, C.POINTER(C.c_double)]
        L.abx_batch_set_step_histories.argtypes = [P, C.c_int64]
        L.abx_batch_set_params.argtypes = [P, C.c_int64, C.c_int64]
        L.abx_batch_set_block_size.argtypes = [P, C.c_int]
        L.abx_batch_set_graphs.argtypes = [P, C.c_int]
        L.abx_batch_graph_stats.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.abx_batch_need.argtypes = [P, P, C.c_int, C.POINTER(C.c_size_t)]
        L.abx_problem_stats.argtypes = [P, C.POINTER(abi.RunStats)]
        L.abx_fission_fetch.argtypes = [P, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.abx_gen_values.argtypes = [P, C.c_int, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.abx_run_replicas.argtypes = [P, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                       C.c_void_p, C.POINTER(abi.RunStats)]
        L.abx_run_replicas_continue.argtypes = [P, C.c_int, C.c_int, C.c_void_p,
                                                C.POINTER(abi.RunStats)]
        L.abx_transport_noalloc.argtypes = [P, C.c_void_p, C.c_size_t, C.c_double, C.c_int,
                                            C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                            C.POINTER(abi.Scores), C.POINTER(abi.Diag)]
        _lib = L
    return _lib


def synchronize():
    """Wait for all work on the device (abx_synchronize)."""
    _check(lib().abx_synchronize())


def _check(rc):
    if rc != abi.ABX_OK:
        raise AbxError(rc, lib().abx_last_error().decode())


class Problem:
    """One problem staged on the GPU (abx_create)."""

    def __init__(self, model, engine=abi.ENGINE_HISTORY):
        self.model = model
        self.desc = model.desc()
        h = C.c_void_p()
        _check(lib().abx_create(C.byref(self.desc), C.byref(h)))
        self.h = h
        self.set_engine(engine)

    def close(self):
        if getattr(self, "h", None):
            lib().abx_destroy(self.h)
            self.h = None

    __del__ = close

    def set_engine(self, engine):
        _check(lib().abx_set_engine(self.h, engine))

    def sample_sources(self, first, n):
        out = np.zeros(n, abi.PARTICLE_DTYPE)
        _check(lib().abx_sample_sources(self.h, first, n, out.ctypes.data))
        return out

    def transport(self, bank, kcol=1.0, converged=True, cap=None, trace_cap=0):
        n = len(bank)
        bank = np.ascontiguousarray(bank, dtype=abi.PARTICLE_DTYPE)
        cap = cap if cap is not None else max(16, 4 * n)
        fis = np.zeros(cap, abi.BANKED_DTYPE)
        n_out = C.c_size_t()
        sc = abi.Scores()
        res = {}
        if trace_cap:
            tr = np.zeros(n * trace_cap, abi.TRACE_DTYPE)
            nrec = np.zeros(n, np.int32)
            _check(lib().abx_transport_trace(self.h, bank.ctypes.data, n, kcol,
                                             1 if converged else 0, fis.ctypes.data, cap,
                                             C.byref(n_out), C.byref(sc), tr.ctypes.data,
                                             trace_cap, nrec.ctypes.data))
            res["trace"] = tr.reshape(n, trace_cap)
            res["n_rec"] = nrec
        else:
            dg = abi.Diag()
            _check(lib().abx_transport(self.h, bank.ctypes.data, n, kcol,
                                       1 if converged else 0, fis.ctypes.data, cap,
                                       C.byref(n_out), C.byref(sc), C.byref(dg)))
            res["diag"] = dg
        res["fission"] = fis[:n_out.value].copy()
        res["scores"] = sc
        return res

    def transport_noalloc(self, d_bank, n, d_fission_out, cap, kcol=1.0, converged=True):
        """abx_transport_noalloc: d_bank (n abx_particle records) and
        d_fission_out (cap abx_banked records) are device addresses (ints).
        Returns (sites written, scores, diag)."""
        n_out = C.c_size_t()
        sc = abi.Scores()
        dg = abi.Diag()
        _check(lib().abx_transport_noalloc(self.h, d_bank, n, kcol, 1 if converged else 0,
                                           d_fission_out, cap, C.byref(n_out), C.byref(sc),
                                           C.byref(dg)))
        return n_out.value, sc, dg

    def tally(self, i):
        nb = C.c_size_t()
        _check(lib().abx_tally_size(self.h, i, C.byref(nb)))
        out = np.zeros(nb.value)
        _check(lib().abx_tally_fetch(self.h, i, out.ctypes.data))
        return out

    def tally_record(self, multiplier=1.0):
        """abx_tally_record: the generation arrays join the running mean and
        variance (MeshTally::record_generation)."""
        _check(lib().abx_tally_record(self.h, multiplier))

    def tally_clear(self):
        _check(lib().abx_tally_clear(self.h))

    def tally_stats(self, i):
        nb = C.c_size_t()
        _check(lib().abx_tally_size(self.h, i, C.byref(nb)))
        avg = np.zeros(nb.value)
        err = np.zeros(nb.value)
        _check(lib().abx_tally_stats(self.h, i, avg.ctypes.data, err.ctypes.data))
        return avg, err

    def run(self, ngenerations, nignored=0):
        """abx_run; k_col per generation that ran (fewer than asked after a
        stop request or the time budget)."""
        st = abi.RunStats()
        kh = np.zeros(ngenerations)
        _check(lib().abx_run(self.h, ngenerations, nignored, kh.ctypes.data, C.byref(st)))
        return st, kh[:st.generations_run]

    def run_continue(self, ngenerations, nignored=0):
        st = abi.RunStats()
        kh = np.zeros(ngenerations)
        _check(lib().abx_run_continue(self.h, ngenerations, nignored, kh.ctypes.data,
                                      C.byref(st)))
        return st, kh[:st.generations_run]

    def request_stop(self):
        """abx_request_stop: the run in progress (or the next) stops at its
        next generation boundary; callable from another thread."""
        _check(lib().abx_request_stop(self.h))

    def set_max_time(self, seconds):
        """abx_set_max_time: wall-clock budget of each run (<= 0: none)."""
        _check(lib().abx_set_max_time(self.h, float(seconds)))

    def run_replicas(self, replicas, ngenerations, nignored=0, seeds=None, from_bank=False):
        """abx_run_replicas: `replicas` chains of this k-eigenvalue problem
        (seeds[r], default settings.seed + r) run together, each from its own
        sampled source or (from_bank) from the bank of the problem's last
        run.  Returns the statistics of the chains' mean and k_col per chain
        and generation ([replicas][ngenerations])."""
        st = abi.RunStats()
        kh = np.zeros((replicas, ngenerations))
        sd = None
        if seeds is not None:
            sd = np.ascontiguousarray(seeds, dtype=np.uint64)
            assert len(sd) == replicas
        _check(lib().abx_run_replicas(self.h, replicas, None if sd is None else sd.ctypes.data,
                                      1 if from_bank else 0, ngenerations, nignored,
                                      kh.ctypes.data, C.byref(st)))
        self._replicas = replicas
        return st, kh[:, :st.generations_run]

    def run_replicas_continue(self, ngenerations, nignored=0):
        """abx_run_replicas_continue: more generations of the last replica run."""
        st = abi.RunStats()
        kh = np.zeros((getattr(self, "_replicas", 1), ngenerations))
        _check(lib().abx_run_replicas_continue(self.h, ngenerations, nignored, kh.ctypes.data,
                                               C.byref(st)))
        return st, kh[:, :st.generations_run]

    def fission_fetch(self, cap):
        """The ordered fission sites of the last transport (abx_fission_fetch),
        e.g. after a transport whose output buffer was too small."""
        out = np.zeros(max(1, cap), abi.BANKED_DTYPE)
        n = C.c_size_t()
        _check(lib().abx_fission_fetch(self.h, out.ctypes.data, cap, C.byref(n)))
        return out[:n.value].copy()

    def gen_values(self, quantity):
        """Per-generation values since the run started (abx_gen_values; abi.GEN_*)."""
        n = C.c_size_t()
        _check(lib().abx_gen_values(self.h, quantity, None, 0, C.byref(n)))
        out = np.zeros(max(1, n.value))
        _check(lib().abx_gen_values(self.h, quantity, out.ctypes.data, len(out), C.byref(n)))
        return out[:n.value].copy()

    def cancel_bank(self, bank):
        b = np.ascontiguousarray(bank, dtype=abi.BANKED_DTYPE).copy()
        _check(lib().abx_cancel_bank(self.h, b.ctypes.data, len(b)))
        return b

    def parents(self, cap=None):
        """abx_parent_fetch: the parent records of the last transport's
        fission sites, in their order."""
        n = C.c_size_t(0)
        if cap is None:  # ask for the count first (ABX_E_CAPACITY = -5 then)
            rc = lib().abx_parent_fetch(self.h, None, 0, C.byref(n))
            if rc not in (0, -5):
                _check(rc)
            cap = n.value
        out = np.zeros(max(1, cap), abi.PARENT_DTYPE)
        _check(lib().abx_parent_fetch(self.h, out.ctypes.data, cap, C.byref(n)))
        return out[: n.value]

    def exact_cancel(self, bank, parents, cap=None):
        """abx_exact_cancel_bank: the bank with its cancelled weights, then
        the uniform particles."""
        b = np.ascontiguousarray(bank, dtype=abi.BANKED_DTYPE)
        q = np.ascontiguousarray(parents, dtype=abi.PARENT_DTYPE)
        if len(q) != len(b):
            raise ValueError("one parent record per site")
        cap = cap if cap is not None else 4 * len(b) + 1024
        out = np.zeros(max(1, cap), abi.BANKED_DTYPE)
        n = C.c_size_t(0)
        _check(lib().abx_exact_cancel_bank(self.h, b.ctypes.data, q.ctypes.data, len(b),
                                           out.ctypes.data, cap, C.byref(n)))
        return out[: n.value]

    def rng_reset(self):
        """abx_rng_reset: settings::rng back to the seed."""
        _check(lib().abx_rng_reset(self.h))

    def rng_state(self):
        """settings::rng's pcg32 state (abx_rng_get_state)."""
        x = C.c_uint64()
        _check(lib().abx_rng_get_state(self.h, C.byref(x)))
        return x.value

    def set_rng_state(self, state):
        _check(lib().abx_rng_set_state(self.h, C.c_uint64(int(state))))

    def bank(self):
        n = C.c_size_t()
        _check(lib().abx_bank_size(self.h, C.byref(n)))
        out = np.zeros((n.value, 9))
        _check(lib().abx_bank_fetch(self.h, out.ctypes.data))
        return out

    def load_bank(self, bank9):
        """Seed the next generation with a [n][9] source (abx_bank_load)."""
        b = np.ascontiguousarray(bank9, dtype=np.float64)
        assert b.ndim == 2 and b.shape[1] == 9
        _check(lib().abx_bank_load(self.h, b.ctypes.data, len(b)))

    def set_tuning(self, block_size=256, blocks_per_cu=0, stage_tables=1, tail_threshold=0,
                   sort_every=0):
        t = abi.Tuning(block_size, blocks_per_cu, stage_tables, tail_threshold, sort_every)
        _check(lib().abx_set_tuning(self.h, C.byref(t)))

    def tuning(self):
        t = abi.Tuning()
        _check(lib().abx_get_tuning(self.h, C.byref(t)))
        return dict(block_size=t.block_size, blocks_per_cu=t.blocks_per_cu,
                    stage_tables=t.stage_tables, tail_threshold=t.tail_threshold,
                    sort_every=t.sort_every)

    def time_transport(self, n, converged=True):
        """Device seconds of one transport of n fresh histories (abx_time_transport)."""
        x = C.c_double()
        _check(lib().abx_time_transport(self.h, int(n), 1 if converged else 0, C.byref(x)))
        return x.value

    def footprint(self):
        """Device bytes the problem holds while it runs (abx_problem_footprint)."""
        n = C.c_size_t()
        _check(lib().abx_problem_footprint(self.h, C.byref(n)))
        return n.value

    def stats(self):
        """Statistics of the problem's last batch-engine run (abx_problem_stats)."""
        st = abi.RunStats()
        _check(lib().abx_problem_stats(self.h, C.byref(st)))
        return st


class Batch:
    """Many fixed-source / modified-fixed-source problems advanced together
    (abx_batch_*).  `budget` caps the device bytes of attached problems."""

    def __init__(self, budget=0, step_histories=0):
        h = C.c_void_p()
        _check(lib().abx_batch_create(int(budget), C.byref(h)))
        self.h = h
        self.attached = {}
        if step_histories:
            self.set_step_histories(step_histories)

    def set_params(self, chunk=0, multi_hist=0, block=256):
        """Batch-kernel chunk length (histories; 0 = about 8 chunks per CU,
        -f = about f per CU), the cap on a problem's side-by-side batches
        (0 = default) and the kernel's lanes per block (64, 128, 256)."""
        _check(lib().abx_batch_set_params(self.h, C.c_int64(int(chunk)),
                                          C.c_int64(int(multi_hist))))
        _check(lib().abx_batch_set_block_size(self.h, int(block)))

    def set_graphs(self, on):
        """Replay the transport of steps that repeat the previous one as a
        captured graph (default on)."""
        _check(lib().abx_batch_set_graphs(self.h, 1 if on else 0))

    def graph_stats(self):
        """(graphs captured, graph replays) so far."""
        c, r = C.c_int64(), C.c_int64()
        _check(lib().abx_batch_graph_stats(self.h, C.byref(c), C.byref(r)))
        return c.value, r.value

    def set_step_histories(self, n):
        """Cap the histories of one step over all problems (0 = no cap); a
        fixed-source batch above its share is split over several steps."""
        _check(lib().abx_batch_set_step_histories(self.h, C.c_int64(int(n))))

    def __del__(self):
        if getattr(self, "h", None):
            lib().abx_batch_destroy(self.h)
            self.h = None

    def add(self, problem, nbatches):
        """Attach; returns False (and attaches nothing) if it does not fit the
        budget."""
        rc = lib().abx_batch_add(self.h, problem.h, int(nbatches))
        if rc == abi.ERRORS_BY_NAME["ABX_E_BUDGET"]:
            return False
        _check(rc)
        self.attached[problem.h.value] = problem
        return True

    def need(self, problem, nbatches):
        """Device bytes `add` would count for `problem` run `nbatches` batches."""
        x = C.c_size_t()
        _check(lib().abx_batch_need(self.h, problem.h, int(nbatches), C.byref(x)))
        return x.value

    def step(self):
        """One step; returns the problems that finished in it."""
        cap = max(1, len(self.attached))
        out = (C.c_void_p * cap)()
        nf = C.c_int()
        na = C.c_int()
        _check(lib().abx_batch_step(self.h, out, cap, C.byref(nf), C.byref(na)))
        done = [self.attached.pop(out[k]) for k in range(nf.value)]
        return done

    def __len__(self):
        return len(self.attached)

    def kernel_seconds(self):
        x = C.c_double()
        _check(lib().abx_batch_kernel_seconds(self.h, C.byref(x)))
        return x.value


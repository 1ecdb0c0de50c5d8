"""ctypes wrapper of tests/native/libhdhost.so (device transport code built
for the host; tests only)."""
import ctypes as C
import fcntl
import os
import subprocess

import numpy as np

from abeille_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libhdhost.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        # one build at a time when test workers run in parallel
        os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
        with open(os.path.join(HERE, "_build", ".lock"), "w") as lk:
            fcntl.flock(lk, fcntl.LOCK_EX)
            subprocess.run(["make", "-s", "-C", HERE], check=True)
            fcntl.flock(lk, fcntl.LOCK_UN)
        L = C.CDLL(LIB)
        P = C.c_void_p
        L.hd_last_error.restype = C.c_char_p
        L.hd_create.argtypes = [C.POINTER(abi.ProblemDesc), C.POINTER(P)]
        L.hd_destroy.argtypes = [P]
        L.hd_destroy.restype = None
        L.hd_sample_sources.argtypes = [P, C.c_uint64, C.c_size_t, C.c_void_p]
        L.hd_transport.argtypes = [P, C.c_void_p, C.c_size_t, C.c_double, C.c_int,
                                   C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t),
                                   C.POINTER(abi.Scores), C.POINTER(abi.Diag),
                                   C.c_void_p, C.c_int, C.c_void_p]
        L.hd_tally_fetch.argtypes = [P, C.c_void_p]
        L.hd_parents_fetch.argtypes = [P, C.c_void_p, C.c_size_t, C.POINTER(C.c_size_t)]
        L.hd_sobol.argtypes = [C.c_uint64, C.c_int]
        L.hd_sobol.restype = C.c_double
        _lib = L
    return _lib


class HostDevice:
    def __init__(self, model):
        self.model = model
        self.desc = model.desc()
        h = C.c_void_p()
        rc = lib().hd_create(C.byref(self.desc), C.byref(h))
        if rc:
            raise RuntimeError(lib().hd_last_error().decode())
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().hd_destroy(self.h)

    def parents(self):
        """Parent records of the last transport's fission sites."""
        n = C.c_size_t()
        lib().hd_parents_fetch(self.h, None, 0, C.byref(n))
        out = np.zeros(max(1, n.value), abi.PARENT_DTYPE)
        rc = lib().hd_parents_fetch(self.h, out.ctypes.data, len(out), C.byref(n))
        if rc:
            raise RuntimeError(f"hd_parents_fetch rc={rc}")
        return out[: n.value].copy()

    def sample_sources(self, first, n):
        out = np.zeros(n, abi.PARTICLE_DTYPE)
        rc = lib().hd_sample_sources(self.h, first, n, out.ctypes.data)
        if rc:
            raise RuntimeError(f"sample_sources rc={rc}")
        return out

    def transport(self, bank, kcol=1.0, converged=True, cap=None, trace_cap=0):
        n = len(bank)
        bank = np.ascontiguousarray(bank)
        cap = cap if cap is not None else max(16, 4 * n)
        fis = np.zeros(cap, abi.BANKED_DTYPE)
        n_out = C.c_size_t()
        sc, dg = abi.Scores(), abi.Diag()
        tr = nrec = None
        if trace_cap:
            tr = np.zeros(n * trace_cap, abi.TRACE_DTYPE)
            nrec = np.zeros(n, np.int32)
        rc = lib().hd_transport(self.h, bank.ctypes.data, n, kcol, 1 if converged else 0,
                                fis.ctypes.data, cap, C.byref(n_out), C.byref(sc),
                                C.byref(dg), None if tr is None else tr.ctypes.data,
                                trace_cap, None if nrec is None else nrec.ctypes.data)
        if rc:
            raise RuntimeError(f"hd_transport rc={rc}: {lib().hd_last_error().decode()}")
        res = dict(fission=fis[:n_out.value].copy(), scores=sc, diag=dg)
        if trace_cap:
            res["trace"] = tr.reshape(n, trace_cap)
            res["n_rec"] = nrec
        return res

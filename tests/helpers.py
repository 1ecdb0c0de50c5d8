"""Shared comparison helpers for the parity tests."""
import ctypes as C

import numpy as np


def first_trace_diff(a, b, na, nb, cap):
    """Index of the first history whose trace differs, with a description."""
    if not np.array_equal(na, nb):
        i = int(np.nonzero(na != nb)[0][0])
        return i, f"history {i}: {na[i]} vs {nb[i]} records"
    for i in range(len(na)):
        k = min(int(na[i]), cap)
        if a[i][:k].tobytes() != b[i][:k].tobytes():
            j = int(np.nonzero([x.tobytes() != y.tobytes() for x, y in zip(a[i][:k], b[i][:k])])[0][0])
            return i, f"history {i} record {j}:\n {a[i][j]}\n {b[i][j]}"
    return None, ""


def assert_traces_equal(ra, rb, cap):
    i, msg = first_trace_diff(ra["trace"], rb["trace"], ra["n_rec"], rb["n_rec"], cap)
    assert i is None, msg


def assert_within_sigma(avg_a, err_a, avg_b, err_b, nsig=3.0, min_frac=0.99):
    """Fraction of bins whose difference is within nsig combined standard
    errors must be at least min_frac (bins with zero error on both sides
    must agree exactly)."""
    comb = np.sqrt(err_a ** 2 + err_b ** 2)
    d = np.abs(avg_a - avg_b)
    ok = (d <= nsig * comb) | ((comb == 0) & (d == 0))
    frac = ok.mean()
    assert frac >= min_frac, f"only {frac:.4f} of bins within {nsig} sigma"
    return frac


def bins_close(a, b, rtol=1e-9):
    """Tally bins of two runs of the same histories: FP64 atomics add in
    another order, and the GPU's track-length walk accumulates the distance
    to the next mesh plane per axis while the reference recomputes it from
    an advanced position, so a segment's length differs by ~1e-15 cm per
    crossing.  Each bin must agree to `rtol` relative, or to 1e-9 of the
    largest bin (a bin holding only a sub-micron corner clip)."""
    scale = max(float(np.abs(a).max()), 1e-300)
    return bool(np.all(np.abs(a - b) <= rtol * np.abs(a) + 1e-9 * scale))


_hip = None


def hip():
    """The HIP runtime libabx.so itself uses (same shared object), for
    device buffers in tests of the entry points that take device pointers."""
    global _hip
    if _hip is None:
        import abeille_amd
        abeille_amd.lib()  # loads libamdhip64 as a dependency of libabx.so
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceBuffer:
    """hipMalloc'd bytes; upload / download of numpy arrays."""

    def __init__(self, nbytes):
        self.nbytes = max(1, int(nbytes))
        self.ptr = C.c_void_p()
        rc = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
        assert rc == 0, f"hipMalloc failed ({rc})"

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, arr):
        a = np.ascontiguousarray(arr)
        assert a.nbytes <= self.nbytes
        assert hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device

    def download(self, nbytes, dtype):
        out = np.zeros(max(1, nbytes), np.uint8)
        if nbytes:
            assert hip().hipMemcpy(out.ctypes.data, self.ptr, nbytes, 2) == 0  # device to host
        return out[:nbytes].view(dtype)

    def __del__(self):
        if getattr(self, "ptr", None) and self.ptr.value:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

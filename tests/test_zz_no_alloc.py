"""abx_transport_noalloc (no device allocation, scratch in a static device
array) against abx_transport on the same banks: ordered fission banks bit
for bit, event counters exact, scalar scores to 1e-12 and tally bins to
1e-9 relative (FP64 atomics add in another order).  Two calls in a row per
problem, the second smaller, so the second call runs over the first call's
leftover scratch contents."""
import numpy as np
import pytest

from abeille_amd import AbxError, Problem, abi, problems
from helpers import DeviceBuffer

pytestmark = pytest.mark.gpu

CASES = {
    "cube_dt": lambda: problems.fixed_source_cube(2000, tracking="delta"),
    "lattice_dt": lambda: problems.pin_lattice_2d(2000, tracking="delta"),
    "core_st": lambda: problems.core_3x3(2000, tracking="surface"),
    "slab_mfs_carter": lambda: problems.subcritical_slab(2000, tracking="carter", cancel=False),
}

REC_P = abi.PARTICLE_DTYPE.itemsize
REC_B = abi.BANKED_DTYPE.itemsize


def _to_device(arr):
    b = DeviceBuffer(arr.nbytes)
    b.upload(arr)
    return b


def _close(a, b, tol):
    return abs(a - b) <= tol * max(1.0, abs(a), abs(b))


@pytest.mark.parametrize("name", sorted(CASES))
def test_noalloc_matches_transport(name):
    print(f"[noalloc] case={name}", flush=True)
    m = CASES[name]()
    ref = Problem(m)
    dut = Problem(m)
    first = 0
    for n in (20000, 3000):
        bank = ref.sample_sources(first, n)
        first += n
        ref.tally_clear()
        dut.tally_clear()
        r = ref.transport(bank, kcol=0.9)
        cap = 4 * n + 4096
        d_bank = _to_device(bank)
        d_out = DeviceBuffer(cap * REC_B)
        n_out, sc, dg = dut.transport_noalloc(d_bank.addr, n, d_out.addr, cap, kcol=0.9)
        fis = d_out.download(n_out * REC_B, abi.BANKED_DTYPE)
        assert n_out == len(r["fission"]), (n, n_out, len(r["fission"]))
        assert fis.tobytes() == r["fission"].tobytes(), f"n={n}: fission banks differ"
        for f in ("histories", "collisions", "virtual_collisions", "flights", "crossings",
                  "tally_updates", "lost", "secondaries"):
            assert getattr(dg, f) == getattr(r["diag"], f), (n, f)
        for f in ("k_col", "k_abs", "k_trk", "k_tot", "leakage", "mig"):
            a, b = getattr(r["scores"], f), getattr(sc, f)
            assert _close(a, b, 1e-12), (n, f, a, b)
        for t in range(len(m.tallies)):
            ta, tb = ref.tally(t), dut.tally(t)
            assert np.all(np.abs(ta - tb) <= 1e-9 * np.maximum(np.abs(ta), 1e-300)), (n, t)
        del d_bank, d_out


def test_noalloc_slices_equal_one_call():
    """A bank split into consecutive slices gives the concatenated fission
    bank and the tallies of one abx_transport call."""
    m = problems.pin_lattice_2d(2000, tracking="delta")
    ref = Problem(m)
    dut = Problem(m)
    n = 30000
    bank = ref.sample_sources(0, n)
    r = ref.transport(bank, kcol=0.9)
    parts = []
    for a, b in ((0, 11111), (11111, n)):
        k = b - a
        cap = 4 * k + 4096
        d_bank = _to_device(bank[a:b])
        d_out = DeviceBuffer(cap * REC_B)
        n_out, _, _ = dut.transport_noalloc(d_bank.addr, k, d_out.addr, cap, kcol=0.9)
        parts.append(d_out.download(n_out * REC_B, abi.BANKED_DTYPE).copy())
    assert np.concatenate(parts).tobytes() == r["fission"].tobytes()
    ta, tb = ref.tally(0), dut.tally(0)
    assert np.all(np.abs(ta - tb) <= 1e-9 * np.maximum(np.abs(ta), 1e-300))


def test_noalloc_rejects_too_many_histories():
    """Checked on the host before anything is launched."""
    m = problems.fixed_source_cube(2000, tracking="delta")
    dut = Problem(m)
    with pytest.raises(AbxError, match="at most 32768"):
        dut.transport_noalloc(1, 32769, 0, 0)

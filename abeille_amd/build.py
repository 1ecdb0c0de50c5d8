"""Builds the gfx950 shared library libabx.so in-tree with hipcc.

abx.hip (host side of the C ABI and the small kernels) and the twelve parts of
instances.hip (the transport kernel variants) compile in parallel to objects
under _build/, which are linked into abeille_amd/libabx.so.

-ffp-contract=off keeps multiply-adds unfused so that device results match
the host restatement bit for bit; packed FP32 ops are disabled (they have
given wrong results on these machines).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libabx.so")
N_PARTS = 12
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wall", "-Wno-unused-function", "-Wno-unused-variable"]


def _units():
    units = [(os.path.join(CSRC, "abx.hip"), os.path.join(OBJ, "abx.o"), [])]
    for k in range(N_PARTS):
        units.append((os.path.join(CSRC, "instances.hip"),
                      os.path.join(OBJ, f"instances_{k}.o"), [f"-DABX_INST_PART={k}"]))
    return units


def _compile(unit, extra, verbose):
    src, obj, defs = unit
    cmd = ["hipcc"] + FLAGS + list(extra) + defs + ["-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, capture_output=True, text=True), cmd


def build(verbose=False, extra_flags=(), jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    units = _units()
    jobs = jobs or min(len(units), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda u: _compile(u, extra_flags, verbose), units))
    failed = [(r, c) for r, c in results if r.returncode != 0]
    for r, c in failed:
        sys.stderr.write(" ".join(c) + "\n" + r.stdout + r.stderr)
    if failed:
        raise RuntimeError("hipcc failed")
    cmd = ["hipcc", "-shared", "-fPIC", "--offload-arch=gfx950"] + [u[1] for u in units] + \
          ["-o", LIB + ".tmp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc link failed")
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))

"""The C++ host mirror of the reference interface (include/abeille_gpu.hpp):
a PowerIterator::run loop written against Transporter / Tallies / Particle
drives the GPU through GpuTransporter and reproduces the LA-13511
PUa-1-0-SL critical slab (analytic k = 1)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "power_iteration.cpp")


def build(tmp_path):
    exe = str(tmp_path / "power_iteration")
    lib = os.path.join(ROOT, "abeille_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", SRC, "-I", os.path.join(ROOT, "include"),
                    "-L", lib, "-labx", f"-Wl,-rpath,{lib}", "-o", exe], check=True)
    return exe


def test_cpp_driver_compiles(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_cpp_power_iteration_critical_slab(tmp_path):
    exe = build(tmp_path)
    out = subprocess.run([exe, "20000", "70", "10"], capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stderr
    _, k, err, gens = out.stdout.split()
    k, err = float(k), float(err)
    assert int(gens) == 60
    assert abs(k - 1.0) < 4 * err, (k, err)

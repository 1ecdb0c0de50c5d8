"""The device transport code (abeille_amd/csrc/*.h) compiled for the host
CPU reproduces the CPU restatement's per-history traces, fission banks and
scores bit for bit.  This pins the physics before any GPU is involved; the
GPU tests then only have to show the gfx950 build computes the same bits."""
import numpy as np
import pytest

from abeille_amd import problems
from helpers import assert_traces_equal
from oracle.pyoracle import Oracle
from pyhd import HostDevice

CASES = {
    "cube_dt": lambda: problems.fixed_source_cube(600, tracking="delta"),
    "cube_st": lambda: problems.fixed_source_cube(600, tracking="surface"),
    "lattice_dt": lambda: problems.pin_lattice_2d(600, tracking="delta"),
    "lattice_st": lambda: problems.pin_lattice_2d(600, tracking="surface"),
    "core_st": lambda: problems.core_3x3(600, tracking="surface"),
    "core_dt": lambda: problems.core_3x3(600, tracking="delta"),
    "la_slab": lambda: problems.la13511("PUb-1-0-SL", 600)[0],
    "la_inf": lambda: problems.la13511("Ua-1-1-IN", 600)[0],
    "mfs_carter": lambda: problems.subcritical_slab(600, tracking="carter", cancel=False),
    "mfs_delta": lambda: problems.subcritical_slab(600, tracking="delta", cancel=False),
    "aniso_st": lambda: problems.anisotropic_slab(600, tracking="surface"),
    "aniso_dt": lambda: problems.anisotropic_slab(600, tracking="delta"),
    "cyl_y_st": lambda: problems.cylinder_cube(600, tracking="surface"),
    "cyl_y_dt": lambda: problems.cylinder_cube(600, tracking="delta"),
    "cyl_x_st": lambda: problems.cylinder_cube(600, tracking="surface", axis=(3., 0., 0.),
                                               centre=(0., 0.5, -0.5)),
    "cyl_z_st": lambda: problems.cylinder_cube(600, tracking="surface", axis=(0., 0., -1.),
                                               centre=(0.5, -0.5, 2.)),
    "cyl_oblique_st": lambda: problems.cylinder_cube(600, tracking="surface", axis=(1., 1., 2.)),
    "cyl_oblique_dt": lambda: problems.cylinder_cube(600, tracking="delta", axis=(-2., 1., 0.5)),
    "hex_flat_st": lambda: problems.hex_assembly(600, tracking="surface", top="flat")[0],
    "hex_pointy_dt": lambda: problems.hex_assembly(600, tracking="delta", top="pointy",
                                                   origin=(0.3, -0.2, 1.0))[0],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_host_build_matches_oracle(name):
    m = CASES[name]()
    n, cap = 600, 512
    o, h = Oracle(m), HostDevice(m)
    so = o.sample_sources(0, n)
    sh = h.sample_sources(0, n)
    assert so.tobytes() == sh.tobytes()
    ro = o.transport(so, kcol=0.93, nthreads=1, trace_cap=cap)
    rh = h.transport(so, kcol=0.93, trace_cap=cap)
    assert_traces_equal(ro, rh, cap)
    assert ro["fission"].tobytes() == rh["fission"].tobytes()
    for f in ("k_col", "k_abs", "k_trk", "leakage", "mig"):
        assert getattr(ro["scores"], f) == getattr(rh["scores"], f)

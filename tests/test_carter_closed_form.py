"""Carter tracking (src/carter_tracker.cpp:92-294) against closed forms.

A one-material 2-group infinite medium (reflective 20 cm box) with sigma_t
{1.0, 1.5}, scatter [[0.3, 0.2], [0, 0.6]]: the volume-integrated flux per
source particle is phi_1 = 1/(St1 - Ss11) = 1/0.7 and
phi_2 = Ss12 phi_1/(St2 - Ss22) = 0.2/(0.7 x 0.9), for the collision and
the track-length estimators.  With the sampling cross section at 0.8 of the
majorant every flight takes the negative-weight branch (D = Et/(2Et - Esmp),
F = Et/(D Esmp), w <- +-wF, splitting at |w| >= 2); at 1.25 it takes the
delta-tracking branch with an over-estimated majorant.  A wrong D or F, a
dropped sign or a split that does not conserve weight biases both fluxes by
far more than their errors.  (The scattering ratios are below 2 x 0.8 - 1, so
the mean |w| shrinks per flight and histories end; at the lattice's ratios
carter tracking at 0.8 grows weights without bound.)
"""
import numpy as np
import pytest

from abeille_amd import abi
from abeille_amd.model import Model
from oracle.pyoracle import Oracle

PHI = np.array([1.0 / 0.7, 0.2 / (0.7 * 0.9)])


def carter_medium(nparticles, ratio):
    m = Model()
    s = [m.surface("xplane", [-10.], "reflective"), m.surface("xplane", [10.], "reflective"),
         m.surface("yplane", [-10.], "reflective"), m.surface("yplane", [10.], "reflective"),
         m.surface("zplane", [-10.], "reflective"), m.surface("zplane", [10.], "reflective")]
    m.material(1, total=[1.0, 1.5], absorption=[0.5, 0.9], scatter=[[0.3, 0.2], [0.0, 0.6]])
    m.cell(f"+{s[0]} & -{s[1]} & +{s[2]} & -{s[3]} & +{s[4]} & -{s[5]}", material=1, id=1)
    m.root = m.cell_universe([1], id=1)
    m.point_source([0., 0., 0.], energy=0.5)
    for est in ("collision", "track-length"):
        m.mesh_tally([-10.] * 3, [10.] * 3, [1, 1, 1], [0., 1., 2.], estimator=est)
    m.set(mode=abi.FIXED_SOURCE, nparticles=nparticles, energy_bounds=[0., 1., 2.],
          tracking=abi.CARTER_TRACKING, sample_xs_ratio=[ratio, ratio])
    return m


def _check_flux(solver):
    for t in range(2):
        avg, err = solver.tally_stats(t)
        z = (avg - PHI) / err
        assert np.all(np.abs(z) < 4.0), (t, avg, err, z)


@pytest.mark.parametrize("ratio", [0.8, 1.25])
def test_oracle_carter_infinite_medium(ratio):
    o = Oracle(carter_medium(20000, ratio))
    st, _ = o.run(16, 0)
    _check_flux(o)
    if ratio < 1:
        assert st.diag.secondaries > 0       # splitting happened
    else:
        assert st.diag.secondaries == 0


@pytest.mark.gpu
@pytest.mark.parametrize("ratio", [0.8, 1.25])
def test_gpu_carter_infinite_medium(ratio):
    """The same closed forms on the GPU, and every bin's mean equal to the
    CPU restatement's on the same histories (up to atomic summation order)."""
    from abeille_amd import Problem
    print(f"[carter-closed-form] ratio={ratio}", flush=True)
    m = carter_medium(20000, ratio)
    g = Problem(m)
    st, _ = g.run(16, 0)
    _check_flux(g)
    o = Oracle(m)
    so, _ = o.run(16, 0)
    assert st.diag.collisions == so.diag.collisions
    assert st.diag.secondaries == so.diag.secondaries
    for t in range(2):
        a, _ = g.tally_stats(t)
        b, _ = o.tally_stats(t)
        assert np.allclose(a, b, rtol=1e-9, atol=0), (t, a, b)

"""Drained triaxial compression of the oracle's non-associative Drucker-Prager (perfect
plasticity), through the oracle NIST solver: a confining pressure pc on the lateral faces
(surface tractions on x = lx and z = lz, symmetry planes x = 0 and z = 0) and a prescribed
axial shortening along y, both ramped over the step.  The stress field is homogeneous
(a linear displacement field is exact in every isoparametric hexahedron, jittered or not),
and once it reaches the cone it stays at the limit that the mathematics fixes:

* the outer-edge Drucker-Prager cone touches Mohr-Coulomb on the compression meridian, so
  the limit deviator is the Mohr-Coulomb one, q = 2 (c cos(phi) + pc sin(phi)) / (1 - sin(phi))
  (de Souza Neto, Peric & Owen 2008, sections 6.4.3 and 8.3.3), with sigma_r = -pc;
* the plastic strain follows the potential, so its volumetric part over the norm of its
  deviatoric part times sqrt(2) is etabar = 6 sin(psi) / (sqrt(3) (3 - sin(psi))) - the only
  place the dilation angle enters this test (the limit stress does not depend on it).
"""

import numpy as np
import pytest

from helpers import oracle_model
from oracle import nist as onist
from oracle.materials import elasticity_matrix

E, NU, PHI, PSI, C, PC = 30000.0, 0.2, 30.0, 10.0, 10.0, 50.0
L = (1.0, 2.0, 1.0)
STRAIN = 0.02


def mohr_coulomb_limit(phi, c, pc):
    sp, cp = np.sin(np.radians(phi)), np.cos(np.radians(phi))
    return 2 * (c * cp + pc * sp) / (1 - sp)


def etabar(psi):
    s = np.sin(np.radians(psi))
    return 6 * s / (np.sqrt(3) * (3 - s))


def triaxial_step(strain=STRAIN, **kw):
    return onist.Step(dirichlets=[dict(nset="gen_top", values={1: -strain * L[1]}), dict(nset="gen_bottom", values={1: 0.0}),
                                  dict(nset="gen_left", values={0: 0.0}), dict(nset="gen_back", values={2: 0.0})],
                      distributedloads=[dict(surface={5: "gen_right"}, vector=(-PC, 0.0, 0.0)),
                                        dict(surface={4: "gen_front"}, vector=(0.0, 0.0, -PC))], **kw)


def plastic_strain(state):
    st = state.reshape(-1, state.shape[-1])
    D = elasticity_matrix(np.array([E]), np.array([NU]))[0]
    return st[:, 6:12] - np.linalg.solve(D, st[:, 0:6].T).T


@pytest.mark.parametrize("el_type,n,jitter", [("C3D8", (1, 1, 1), 0.0), ("C3D20R", (1, 1, 1), 0.0),
                                              ("C3D8", (2, 3, 2), 0.15)])
@pytest.mark.parametrize("phi,psi", [(PHI, PSI), (PHI, PHI), (20.0, 0.0)])
def test_triaxial_limit_stress_and_dilatancy(el_type, n, jitter, phi, psi):
    om = oracle_model(n, L, el_type, "druckerprager", [E, NU, phi, psi, C, 0.0], jitter=jitter, seed=4)
    so = onist.NIST(convention="colmajor")
    so.solve_step(om, triaxial_step(startInc=0.1, maxInc=0.1))
    assert so.history[-1][1] == pytest.approx(1.0)
    s = om.state[..., 0:6].reshape(-1, 6)
    q = mohr_coulomb_limit(phi, C, PC)
    np.testing.assert_allclose(s[:, [0, 2]], -PC, rtol=1e-10)
    np.testing.assert_allclose(s[:, 1], -PC - q, rtol=1e-10)
    np.testing.assert_allclose(s[:, 3:], 0.0, atol=1e-10 * q)
    assert (om.state[..., 12] > 0).all()
    ep = plastic_strain(om.state)
    ev = ep[:, :3].sum(axis=1)
    dev = np.sqrt(((ep[:, :3] - ev[:, None] / 3) ** 2).sum(axis=1))
    np.testing.assert_allclose(ev / (np.sqrt(2) * dev), etabar(psi), rtol=1e-9, atol=1e-12)


def test_triaxial_limit_is_reached_only_after_yield():
    """At a tenth of the axial shortening the state is still elastic and its axial stress well
    above the limit: the check above is not satisfied trivially by an elastic state."""
    om = oracle_model((1, 1, 1), L, "C3D8", "druckerprager", [E, NU, PHI, PSI, C, 0.0])
    so = onist.NIST(convention="colmajor")
    so.solve_step(om, triaxial_step(strain=STRAIN / 10, startInc=0.5, maxInc=0.5))
    s = om.state[..., 0:6].reshape(-1, 6)
    assert (om.state[..., 12] == 0).all()
    assert (s[:, 1] > -PC - 0.5 * mohr_coulomb_limit(PHI, C, PC)).all()

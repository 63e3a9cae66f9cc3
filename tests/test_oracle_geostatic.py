"""Oracle geostatic initial stress (Marmot's rule behind stepactions/geostatic.py:94-101).

The stress rule itself has no reference implementation here (Marmot is an external
dependency), so it is pinned by what the mechanics fixes:
* QP coordinates on an undistorted box equal the affine map of the exchanged evaluation
  point (eta_i, xi_i, zeta_i) (closed form);
* a linear geostatic stress sigma_yy(y) with lateral ratio k, together with the body force
  b_y = -d sigma_yy / dy, is in equilibrium: the residual on the free DOFs vanishes and the
  geostatic step leaves U = 0, for hexa8 and for hexa20 with reduced integration (Gauss
  quadrature is exact for these polynomial integrands on an affine mesh);
* the same step with the body force of the wrong sign leaves a clearly nonzero U, so the
  check can fail.
"""

import numpy as np
import pytest

from helpers import oracle_model, oracle_step
from oracle.element import qp_table
from oracle.nist import NIST

LE = [30000.0, 0.2]
GEO = dict(p1=0.0, h1=1.0, p2=-10.0, h2=0.0, xLateral=0.5, zLateral=0.7)
ROLLERS = [("bottom", "gen_bottom", {1: 0.0}), ("left", "gen_left", {0: 0.0}), ("right", "gen_right", {0: 0.0}),
           ("back", "gen_back", {2: 0.0}), ("front", "gen_front", {2: 0.0})]


def _geostatic_step(body_y):
    st = oracle_step(ROLLERS, bodyforces=[("all", (0.0, 0.0, 0.0))], maxInc=1.0)
    st.bodyforces[0]["start"] = (0.0, body_y, 0.0)   # constant over the step
    st.geostatic = [dict(GEO)]
    return st


@pytest.mark.parametrize("el_type,n", [("C3D8", (3, 4, 2)), ("C3D20R", (2, 3, 2))])
def test_qp_coordinates_follow_the_exchanged_evaluation_point(el_type, n):
    om = oracle_model(n, (1.0, 1.0, 1.0), el_type, "linearelastic", LE)
    xyz = om.elements.qp_coordinates()
    xi, eta, zeta, _ = qp_table(el_type)
    X0 = om.coords[om.elements.conn].min(axis=1)  # lower corner of each element
    h = np.array([1.0 / n[0], 1.0 / n[1], 1.0 / n[2]])
    want = X0[:, None, :] + (1.0 + np.stack([eta, xi, zeta], axis=1))[None, :, :] * h / 2
    np.testing.assert_allclose(xyz, want, rtol=0, atol=1e-14)


@pytest.mark.parametrize("el_type,n", [("C3D8", (3, 4, 2)), ("C3D20R", (2, 3, 2))])
def test_geostatic_stress_with_matching_gravity_is_in_equilibrium(el_type, n):
    slope = (GEO["p2"] - GEO["p1"]) / (GEO["h2"] - GEO["h1"])
    om = oracle_model(n, (1.0, 1.0, 1.0), el_type, "linearelastic", LE)
    so = NIST()
    so.solve_step(om, _geostatic_step(-slope))
    assert [h[2] for h in so.history] == [1, 1]
    assert np.abs(om.U).max() == 0.0          # reset at the end of the geostatic step
    assert np.abs(so.dU).max() < 1e-12        # the step itself moved nothing
    y = om.elements.qp_coordinates()[..., 1]
    sy = GEO["p1"] + slope * (y - GEO["h1"])
    np.testing.assert_allclose(om.state[..., 1], sy, rtol=0, atol=1e-9)
    np.testing.assert_allclose(om.state[..., 0], GEO["xLateral"] * sy, rtol=0, atol=1e-9)
    np.testing.assert_allclose(om.state[..., 2], GEO["zLateral"] * sy, rtol=0, atol=1e-9)
    np.testing.assert_allclose(om.state[..., 3:6], 0.0, atol=1e-9)

    bad = oracle_model(n, (1.0, 1.0, 1.0), el_type, "linearelastic", LE)
    sb = NIST()
    sb.solve_step(bad, _geostatic_step(+slope))
    assert np.abs(sb.dU).max() > 1e-5

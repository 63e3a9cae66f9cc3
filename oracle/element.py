"""Hexahedral small-strain displacement element (oracle side).

Restates edelweissfe/elements/displacementelement/element.py for C3D8, C3D20
and C3D20R, vectorised over elements:

* quadrature tables: edelweissfe/elements/library.py:33-46,212-227,260-291;
* Jacobian and B operator: _elementcomputationmatrices.py:306-366 (_J3D8),
  :369-540 (_J3D20), :679-817 (_B3D8), :820-1060 (_B3D20).  The reference fits
  the isoparametric map with a polynomial and evaluates its derivative with
  the first two natural coordinates exchanged, i.e. quadrature point i is
  evaluated at the natural point (eta_i, xi_i, zeta_i) while its weight stays
  w_i.  Here the same quantities are formed from analytic shape-function
  derivatives at that exchanged point;
* hexa20 Jacobian, third row: the reference's hand-expanded derivative of the fitted
  map with respect to the third natural coordinate multiplies the coefficient a5 of the
  monomial (second coordinate x third coordinate) by z_i where the derivative gives the
  second coordinate of the evaluation point, xi_i (_elementcomputationmatrices.py:502,
  :514, :526; the first two rows and hexa8 are exact).  Since a5 = d2x/(d eta d zeta) at
  the element centre = sum_a h_a x_a with h_a = d2N_a/(d eta d zeta)(0) (+-1/4 on the four
  mid-side nodes (0, +-1, +-1), 0 elsewhere), the reference's J is the exact J at the
  exchanged point with (zeta_i - xi_i) sum_a h_a x_a added to its third row.  B uses that
  J with the exact natural derivatives (:820-1060), and detJ of that J enters K, P and the
  body force.  On undistorted elements a5 = 0 and nothing changes; on distorted ones J
  differs by up to ~10 %.  Pinned by tests/golden/element_fixtures.npz, the reference
  element's own J, B, Ke, Pe, state and body force on distorted elements;
* computeYourself (element.py:290-346): dEps_i = B_i dU, material update per
  QP on a copy of the reference state, K += B^T C B detJ w, P -= B^T sigma detJ w,
  strain += dEps;
* computeBodyForce (element.py:348-371): P += N_i (x) f detJ_i w_i, with N at the
  table point (xi_i, eta_i, zeta_i) (computeNOperator, :106-212) and detJ at the
  exchanged point, exactly as the reference combines them.

State layout per QP follows element.py:222-236: [stress(6), strain(6), material state].
"""

import numpy as np

_S3 = 1 / np.sqrt(3)
_S06 = np.sqrt(0.6)

# natural coordinates of the element nodes (xi, eta, zeta)
NAT8 = np.array([(-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1),
                 (-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)], dtype=float)
NAT20 = np.vstack([NAT8, np.array([
    (-1, -1, 0), (0, -1, 1), (1, -1, 0), (0, -1, -1),
    (-1, 1, 0), (0, 1, 1), (1, 1, 0), (0, 1, -1),
    (-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)], dtype=float)])


def qp_table(el_type):
    """(xi, eta, zeta, w) as the reference library defines them."""
    el_type = el_type.upper()
    if el_type in ("C3D8", "C3D20R"):
        t8 = _S3 * np.array([-1, -1, 1, 1])
        s8 = _S3 * np.array([-1, 1, 1, -1])
        xi = _S3 * np.hstack([-np.ones(4), np.ones(4)])
        return xi, np.hstack([t8, t8]), np.hstack([s8, s8]), np.ones(8)
    if el_type in ("C3D20", "C3D8E"):
        s20 = np.array([-1, 0, 1])
        t20 = np.array([-1, -1, -1, 0, 0, 0, 1, 1, 1])
        w1H = (5 / 9) ** 3
        w2H = (5 / 9) ** 2 * (8 / 9)
        w3H = (5 / 9) * (8 / 9) ** 2
        w4H = (8 / 9) ** 3
        wI = np.array([w1H, w2H, w1H])
        wII = np.array([w2H, w3H, w2H])
        xi = _S06 * np.hstack([-np.ones(9), np.zeros(9), np.ones(9)])
        eta = _S06 * np.hstack([t20, t20, t20])
        zeta = _S06 * np.hstack([s20 for _ in range(9)])
        w = np.hstack([wI, wII, wI, wII, np.array([w3H, w4H, w3H]), wII, wI, wII, wI])
        return xi, eta, zeta, w
    raise ValueError(el_type)


def nodes_per_element(el_type):
    return 8 if el_type.upper() in ("C3D8", "C3D8E") else 20


def shape_functions(nn, p):
    """N (nQP, nn) at natural points p (nQP, 3)."""
    nat = NAT8 if nn == 8 else NAT20
    x, y, z = p[:, 0:1], p[:, 1:2], p[:, 2:3]
    a, b, c = nat[:, 0], nat[:, 1], nat[:, 2]
    if nn == 8:
        return (1 + a * x) * (1 + b * y) * (1 + c * z) / 8
    N = np.empty((p.shape[0], 20))
    cor = (1 + a[:8] * x) * (1 + b[:8] * y) * (1 + c[:8] * z) * (a[:8] * x + b[:8] * y + c[:8] * z - 2) / 8
    N[:, :8] = cor
    for k in range(8, 20):
        ak, bk, ck = a[k], b[k], c[k]
        if ak == 0:
            N[:, k] = ((1 - x * x) * (1 + bk * y) * (1 + ck * z) / 4)[:, 0]
        elif bk == 0:
            N[:, k] = ((1 + ak * x) * (1 - y * y) * (1 + ck * z) / 4)[:, 0]
        else:
            N[:, k] = ((1 + ak * x) * (1 + bk * y) * (1 - z * z) / 4)[:, 0]
    return N


def shape_derivatives(nn, p):
    """dN/dxi (nQP, nn, 3) at natural points p (nQP, 3)."""
    nat = NAT8 if nn == 8 else NAT20
    q = p.shape[0]
    d = np.empty((q, nn, 3))
    x, y, z = p[:, 0][:, None], p[:, 1][:, None], p[:, 2][:, None]
    a, b, c = nat[None, :, 0], nat[None, :, 1], nat[None, :, 2]
    if nn == 8:
        d[:, :, 0] = a * (1 + b * y) * (1 + c * z) / 8
        d[:, :, 1] = b * (1 + a * x) * (1 + c * z) / 8
        d[:, :, 2] = c * (1 + a * x) * (1 + b * y) / 8
        return d
    ac, bc, cc = a[:, :8], b[:, :8], c[:, :8]
    d[:, :8, 0] = ac * (1 + bc * y) * (1 + cc * z) * (2 * ac * x + bc * y + cc * z - 1) / 8
    d[:, :8, 1] = bc * (1 + ac * x) * (1 + cc * z) * (ac * x + 2 * bc * y + cc * z - 1) / 8
    d[:, :8, 2] = cc * (1 + ac * x) * (1 + bc * y) * (ac * x + bc * y + 2 * cc * z - 1) / 8
    x, y, z = x[:, 0], y[:, 0], z[:, 0]
    for k in range(8, 20):
        ak, bk, ck = nat[k]
        if ak == 0:
            d[:, k, 0] = -2 * x * (1 + bk * y) * (1 + ck * z) / 4
            d[:, k, 1] = bk * (1 - x * x) * (1 + ck * z) / 4
            d[:, k, 2] = ck * (1 - x * x) * (1 + bk * y) / 4
        elif bk == 0:
            d[:, k, 0] = ak * (1 - y * y) * (1 + ck * z) / 4
            d[:, k, 1] = -2 * y * (1 + ak * x) * (1 + ck * z) / 4
            d[:, k, 2] = ck * (1 + ak * x) * (1 - y * y) / 4
        else:
            d[:, k, 0] = ak * (1 + bk * y) * (1 - z * z) / 4
            d[:, k, 1] = bk * (1 + ak * x) * (1 - z * z) / 4
            d[:, k, 2] = -2 * z * (1 + ak * x) * (1 + bk * y) / 4
    return d


def evaluation_points(el_type):
    """Natural points where the reference evaluates J and B: (eta_i, xi_i, zeta_i)."""
    xi, eta, zeta, _ = qp_table(el_type)
    return np.stack([eta, xi, zeta], axis=1)


def jacobian_table(el_type):
    """dN/dxi (Q, nn, 3) at the evaluation points, as the reference forms J from them: for
    hexa20 the third component carries (zeta_q - xi_q) h_a in addition (module docstring)."""
    nn = nodes_per_element(el_type)
    xi, eta, zeta, _ = qp_table(el_type)
    dJ = shape_derivatives(nn, evaluation_points(el_type)).copy()
    if nn == 20:
        nat = NAT20
        h = np.where(nat[:, 0] == 0, nat[:, 1] * nat[:, 2] / 4.0, 0.0)  # d2N/(d eta d zeta) at 0
        dJ[:, :, 2] += (zeta - xi)[:, None] * h[None, :]
    return dJ


def geometry(el_type, X):
    """Per element and QP: J (E,Q,3,3) with J[r,c] = d x_c / d xi_r (the reference's J,
    jacobian_table), detJ (E,Q), dN/dx (E,Q,nn,3) = J^-1 dN/dxi."""
    nn = nodes_per_element(el_type)
    dN = shape_derivatives(nn, evaluation_points(el_type))  # (Q, nn, 3)
    J = np.einsum("qar,eac->eqrc", jacobian_table(el_type), X)
    detJ = np.linalg.det(J)
    invJ = np.linalg.inv(J)
    dNdx = np.einsum("eqcr,qar->eqac", invJ, dN)
    return J, detJ, dNdx


def b_operator(dNdx):
    """B (E,Q,6,3nn) from dN/dx (E,Q,nn,3); Voigt 11,22,33,12,13,23 with engineering shear."""
    E, Q, nn, _ = dNdx.shape
    B = np.zeros((E, Q, 6, 3 * nn))
    gx, gy, gz = dNdx[..., 0], dNdx[..., 1], dNdx[..., 2]
    B[:, :, 0, 0::3] = gx
    B[:, :, 1, 1::3] = gy
    B[:, :, 2, 2::3] = gz
    B[:, :, 3, 0::3] = gy
    B[:, :, 3, 1::3] = gx
    B[:, :, 4, 0::3] = gz
    B[:, :, 4, 2::3] = gx
    B[:, :, 5, 1::3] = gz
    B[:, :, 5, 2::3] = gy
    return B


class HexElements:
    """A homogeneous set of hexahedral DisplacementElements."""

    def __init__(self, el_type, coords, conn):
        self.el_type = el_type.upper()
        self.nn = nodes_per_element(el_type)
        self.nDof = 3 * self.nn
        self.conn = np.asarray(conn)
        self.coords = np.asarray(coords, dtype=float)
        X = coords[self.conn]  # (E, nn, 3)
        self.J, self.detJ, self.dNdx = geometry(self.el_type, X)
        self.B = b_operator(self.dNdx)
        self.xi, self.eta, self.zeta, self.w = qp_table(self.el_type)
        self.nQP = self.w.size

    @property
    def dof_map(self):
        """Element DOF indices, node-major interleaved (dofmanager.py:455-471)."""
        return (3 * self.conn[:, :, None] + np.arange(3)[None, None, :]).reshape(self.conn.shape[0], -1)

    def compute(self, dU_e, state_ref, material, idx=None):
        """computeYourself for all elements: returns Ke (E,nDof,nDof), Pe (E,nDof), new state.

        material.compute(stress, mstate, dstrain) -> (stress, mstate, C) on flattened QPs,
        raising CutbackRequest like vonmises.py:230-231.
        """
        B = self.B if idx is None else self.B[idx]
        detJ = self.detJ if idx is None else self.detJ[idx]
        E, Q = B.shape[0], B.shape[1]
        dEps = np.einsum("eqij,ej->eqi", B, dU_e)
        st = state_ref.copy()
        stress, C, mstate = material.compute(st[:, :, 0:6].reshape(-1, 6), st[:, :, 12:].reshape(E * Q, -1),
                                             dEps.reshape(-1, 6))
        st[:, :, 0:6] = stress.reshape(E, Q, 6)
        st[:, :, 12:] = mstate.reshape(E, Q, -1)
        C = C.reshape(E, Q, 6, 6)
        f = detJ * self.w[None, :]
        Ke = np.einsum("eqki,eqkl,eqlj,eq->eij", B, C, B, f, optimize=True)
        Pe = -np.einsum("eqki,eqk,eq->ei", B, st[:, :, 0:6], f)
        st[:, :, 6:12] += dEps
        return Ke, Pe, st

    def qp_coordinates(self):
        """Physical coordinates (E, Q, 3) of the QPs at the points where J and B are evaluated."""
        N = shape_functions(self.nn, evaluation_points(self.el_type))  # (Q, nn)
        return np.einsum("qa,eac->eqc", N, self.coords[self.conn])

    def body_force(self, load):
        """computeBodyForce for all elements (element.py:348-371)."""
        N = shape_functions(self.nn, np.stack([self.xi, self.eta, self.zeta], axis=1))  # (Q, nn)
        f = self.detJ * self.w[None, :]  # (E, Q)
        Pe = np.einsum("qa,eq,c->eac", N, f, np.asarray(load, dtype=float))
        return Pe.reshape(Pe.shape[0], -1)

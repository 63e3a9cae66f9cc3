"""Plane-strain CPE8 restatement (TEST INFRASTRUCTURE ONLY).

Used to pin the oracle's VonMises material (nonlinear isotropic hardening) against the
reference's only golden vector for that material, testfiles/CantileverBeamVonMisesPE, a
plane-strain cantilever of 8-node quads.  The GPU path itself is hexahedral.

* plane_rect_quad follows edelweissfe/generators/planerectquad.py:62-140: grid nodes
  x-major then y, 8-node connectivity [c0, c1, c2, c3, m01, m12, m23, m30], node sets;
  only nodes used by an element carry DOFs (the node field order of
  models/femodel.py _populateNodeFieldVariablesFromElements), so the unused quad centre
  nodes are dropped.
* QuadElements8PE follows DisplacementElement for CPE8 (elements/library.py:100-114,
  displacementelement/element.py:290-346, _elementcomputationmatrices.py:252-303,603-676):
  9 Gauss points; J and B at the exchanged point (eta_i, xi_i), as in 3D; strain rows
  [e11, e22, g12] into Voigt slots [0, 1, 3] of the 3D material (plane strain); the
  thickness multiplies K and P; body force with N at the table point (xi_i, eta_i).
"""

import numpy as np

# natural coordinates of the Q8 nodes (_elementcomputationmatrices.py:140-152 shape functions)
NATQ8 = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0]], dtype=float)
VOIGT2D = np.array([0, 1, 3])


def plane_rect_quad(nX, nY, l, h, x0=0.0, y0=0.0, name="gen"):
    nx, ny = 2 * nX + 1, 2 * nY + 1
    xs = np.linspace(x0, x0 + l, nx)
    ys = np.linspace(y0, y0 + h, ny)
    grid = np.arange(nx * ny).reshape(nx, ny)  # grid label, x-major then y
    gconn = []
    for x in range(nX):
        for y in range(nY):
            gconn.append([grid[2 * x, 2 * y], grid[2 * x + 2, 2 * y], grid[2 * x + 2, 2 * y + 2], grid[2 * x, 2 * y + 2],
                          grid[2 * x + 1, 2 * y], grid[2 * x + 2, 2 * y + 1], grid[2 * x + 1, 2 * y + 2],
                          grid[2 * x, 2 * y + 1]])
    gconn = np.array(gconn)
    used = np.zeros(nx * ny, dtype=bool)
    used[gconn.ravel()] = True
    pos = -np.ones(nx * ny, dtype=np.int64)
    pos[used] = np.arange(np.count_nonzero(used))
    gx, gy = np.meshgrid(xs, ys, indexing="ij")
    coords = np.column_stack([gx.ravel(), gy.ravel()])[used]

    def ns(g):
        p = pos[np.ravel(g)]
        return p[p >= 0]

    nodesets = {name + "_left": ns(grid[0, :]), name + "_right": ns(grid[-1, :]), name + "_top": ns(grid[:, -1]),
                name + "_bottom": ns(grid[:, 0]), name + "_leftBottom": ns(grid[0, 0]),
                name + "_leftTop": ns(grid[0, -1]), name + "_rightBottom": ns(grid[-1, 0]),
                name + "_rightTop": ns(grid[-1, -1]), name + "_all": np.arange(coords.shape[0]),
                "all": np.arange(coords.shape[0])}
    elsets = {"all": np.arange(gconn.shape[0])}
    return {"coords": coords, "conn": pos[gconn], "nodesets": nodesets, "elsets": elsets}


def qp_table_cpe8():
    """library.py:100-114."""
    s = np.sqrt(0.6)
    xi = s * np.array([0, -1, -1, 1, 1, -1, 0, 1, 0])
    eta = s * np.array([0, -1, 1, 1, -1, 0, 1, 0, -1])
    w1, w2, w3 = (5 / 9) ** 2, (5 / 9) * (8 / 9), (8 / 9) ** 2
    w = np.array([w3, w1, w1, w1, w1, w2, w2, w2, w2])
    return xi, eta, w


def q8_shape(p):
    """N (nQP, 8) of the serendipity quad at natural points p (nQP, 2)."""
    x, y = p[:, 0:1], p[:, 1:2]
    a, b = NATQ8[:, 0], NATQ8[:, 1]
    N = np.empty((p.shape[0], 8))
    N[:, :4] = (1 + a[:4] * x) * (1 + b[:4] * y) * (a[:4] * x + b[:4] * y - 1) / 4
    for k in range(4, 8):
        if a[k] == 0:
            N[:, k] = ((1 - x * x) * (1 + b[k] * y) / 2)[:, 0]
        else:
            N[:, k] = ((1 + a[k] * x) * (1 - y * y) / 2)[:, 0]
    return N


def q8_derivatives(p):
    """dN/dxi (nQP, 8, 2)."""
    x, y = p[:, 0][:, None], p[:, 1][:, None]
    a, b = NATQ8[None, :4, 0], NATQ8[None, :4, 1]
    d = np.empty((p.shape[0], 8, 2))
    d[:, :4, 0] = a * (1 + b * y) * (2 * a * x + b * y) / 4
    d[:, :4, 1] = b * (1 + a * x) * (a * x + 2 * b * y) / 4
    x, y = x[:, 0], y[:, 0]
    for k in range(4, 8):
        ak, bk = NATQ8[k]
        if ak == 0:
            d[:, k, 0] = -x * (1 + bk * y)
            d[:, k, 1] = bk * (1 - x * x) / 2
        else:
            d[:, k, 0] = ak * (1 - y * y) / 2
            d[:, k, 1] = -y * (1 + ak * x)
    return d


class QuadElements8PE:
    """All CPE8 elements of a mesh, vectorised like oracle.element.HexElements."""

    nn = 8
    dim = 2

    def __init__(self, coords, conn, thickness):
        self.conn = np.asarray(conn)
        self.coords = np.asarray(coords, dtype=float)
        self.t = float(thickness)
        self.xi, self.eta, self.w = qp_table_cpe8()
        self.nQP = self.w.size
        self.nDof = 16
        dN = q8_derivatives(np.stack([self.eta, self.xi], axis=1))  # exchanged evaluation point
        X = self.coords[self.conn]  # (E, 8, 2)
        self.J = np.einsum("qar,eac->eqrc", dN, X)
        self.detJ = np.linalg.det(self.J)
        invJ = np.linalg.inv(self.J)
        g = np.einsum("eqcr,qar->eqac", invJ, dN)  # dN/dx (E, Q, 8, 2)
        E, Q = g.shape[0], g.shape[1]
        B = np.zeros((E, Q, 3, 16))
        B[:, :, 0, 0::2] = g[..., 0]
        B[:, :, 1, 1::2] = g[..., 1]
        B[:, :, 2, 0::2] = g[..., 1]
        B[:, :, 2, 1::2] = g[..., 0]
        self.B = B

    @property
    def dof_map(self):
        return (2 * self.conn[:, :, None] + np.arange(2)[None, None, :]).reshape(self.conn.shape[0], -1)

    def compute(self, dU_e, state_ref, material, idx=None):
        B, detJ = self.B, self.detJ
        E, Q = B.shape[0], B.shape[1]
        dEps = np.zeros((E, Q, 6))
        dEps[:, :, VOIGT2D] = np.einsum("eqij,ej->eqi", B, dU_e)
        st = state_ref.copy()
        stress, C, mstate = material.compute(st[:, :, 0:6].reshape(-1, 6), st[:, :, 12:].reshape(E * Q, -1),
                                             dEps.reshape(-1, 6))
        st[:, :, 0:6] = stress.reshape(E, Q, 6)
        st[:, :, 12:] = mstate.reshape(E, Q, -1)
        C = C.reshape(E, Q, 6, 6)[:, :, VOIGT2D][:, :, :, VOIGT2D]
        f = detJ * self.w[None, :] * self.t
        Ke = np.einsum("eqki,eqkl,eqlj,eq->eij", B, C, B, f, optimize=True)
        Pe = -np.einsum("eqki,eqk,eq->ei", B, st[:, :, VOIGT2D], f)
        st[:, :, 6:12] += dEps
        return Ke, Pe, st

    def body_force(self, load):
        N = q8_shape(np.stack([self.xi, self.eta], axis=1))  # (Q, 8) at the table point
        f = self.detJ * self.w[None, :] * self.t
        Pe = np.einsum("qa,eq,c->eac", N, f, np.asarray(load, dtype=float)[:2])
        return Pe.reshape(Pe.shape[0], -1)

"""Closed-form and independent checks that pin the oracle's element."""

import os

import numpy as np
import pytest

from oracle import element, materials, mesh


def _distorted(nn, seed=0):
    nat = element.NAT8 if nn == 8 else element.NAT20
    rs = np.random.RandomState(seed)
    A = np.eye(3) * [2.0, 1.5, 1.0] + 0.2 * rs.randn(3, 3)
    X = 0.5 * (nat @ A.T) + np.array([3.0, -1.0, 2.0])
    X[:8] += 0.08 * rs.randn(8, 3)  # non-affine corner perturbation
    if nn == 20:  # keep mid-side nodes at edge midpoints of the perturbed corners
        edges = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
        for k, (a, b) in enumerate(edges):
            X[8 + k] = 0.5 * (X[a] + X[b]) + 0.02 * rs.randn(3)
    return X


def _poly_jacobian_hexa8(X, xi, eta, zeta):
    """Independent formulation of the reference's polynomial-fit Jacobian
    (_elementcomputationmatrices.py:306-366): fit x(s,t,u) = a0 + a1 s + a2 t + a3 u + a4 st + a5 tu
    + a6 su + a7 stu through the 8 nodes, then evaluate its derivative rows with the first two
    natural coordinates of the quadrature point exchanged."""
    nat = element.NAT8
    s, t, u = nat[:, 0], nat[:, 1], nat[:, 2]
    A = np.stack([np.ones(8), s, t, u, s * t, t * u, s * u, s * t * u], axis=1)
    a = np.linalg.solve(A, X)  # (8, 3)
    Js = []
    for i in range(len(xi)):
        S, T, Uu = eta[i], xi[i], zeta[i]  # exchanged evaluation point
        d_ds = a[1] + a[4] * T + a[6] * Uu + a[7] * T * Uu
        d_dt = a[2] + a[4] * S + a[5] * Uu + a[7] * S * Uu
        d_du = a[3] + a[5] * T + a[6] * S + a[7] * S * T
        Js.append(np.stack([d_ds, d_dt, d_du]))
    return np.array(Js)


def test_hexa8_jacobian_equals_polynomial_fit_at_exchanged_points():
    X = _distorted(8)
    xi, eta, zeta, _ = element.qp_table("C3D8")
    J, detJ, _ = element.geometry("C3D8", X[None])
    Jp = _poly_jacobian_hexa8(X, xi, eta, zeta)
    assert np.allclose(J[0], Jp, rtol=0, atol=1e-13)
    # and it is NOT the Jacobian at the unexchanged table points on a distorted element
    dN = element.shape_derivatives(8, np.stack([xi, eta, zeta], axis=1))
    Jt = np.einsum("qar,ac->qrc", dN, X)
    assert not np.allclose(J[0], Jt, atol=1e-6)


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20", "C3D20R"])
def test_shape_functions_partition_of_unity_and_kronecker(el_type):
    nn = element.nodes_per_element(el_type)
    nat = element.NAT8 if nn == 8 else element.NAT20
    N = element.shape_functions(nn, nat)
    assert np.allclose(N, np.eye(nn), atol=1e-14)
    p = np.random.RandomState(1).uniform(-1, 1, (20, 3))
    assert np.allclose(element.shape_functions(nn, p).sum(axis=1), 1.0)
    assert np.allclose(element.shape_derivatives(nn, p).sum(axis=1), 0.0, atol=1e-14)
    # derivatives against central differences of N
    h = 1e-6
    for r in range(3):
        e = np.zeros(3)
        e[r] = h
        fd = (element.shape_functions(nn, p + e) - element.shape_functions(nn, p - e)) / (2 * h)
        assert np.allclose(element.shape_derivatives(nn, p)[:, :, r], fd, atol=1e-8)


def _untwisted(X):
    """The hexa20 element X moved to twist coefficient a5 = d2x/(d eta d zeta) at the centre
    = (-x_9 + x_11 + x_13 - x_15) / 4 = 0 (0-based mid-side nodes (0, -+1, +-1)), by moving
    node 9 alone.  The reference's hexa20 Jacobian is exact on such elements only."""
    X = X.copy()
    a5 = (-X[9] + X[11] + X[13] - X[15]) / 4
    X[9] += 4 * a5
    return X


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20", "C3D20R"])
def test_element_stiffness_symmetry_rigid_modes_and_patch(el_type):
    nn = element.nodes_per_element(el_type)
    X = _distorted(nn, 2)
    if nn == 20:
        X = _untwisted(X)
    el = element.HexElements(el_type, X, np.arange(nn)[None])
    mat = materials.LinearElastic([210000.0, 0.3])
    st = np.zeros((1, el.nQP, 12))
    Ke, Pe, _ = el.compute(np.zeros((1, 3 * nn)), st, mat)
    K = Ke[0]
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    # six rigid-body modes in the null space
    modes = []
    for c in range(3):
        m = np.zeros((nn, 3))
        m[:, c] = 1
        modes.append(m.ravel())
    for axis in range(3):
        w = np.zeros(3)
        w[axis] = 1
        modes.append(np.cross(w, X).ravel())
    for m in modes:
        assert np.abs(K @ m).max() < 1e-9 * np.abs(K).max()
    # rank = 3 nn - 6 for full integration; reduced C3D20R has extra zero-energy modes
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    n_zero = int((np.abs(ev) < 1e-8 * ev.max()).sum())
    assert n_zero == (6 if el_type != "C3D20R" else 12)
    # patch test: linear displacement field gives the exact constant strain at every QP
    G = np.array([[1e-3, 2e-4, -3e-4], [5e-4, -2e-3, 1e-4], [0.0, 3e-4, 1.5e-3]])
    u = (X @ G.T).ravel()
    _, _, st2 = el.compute(u[None], st, mat)
    eps = np.array([G[0, 0], G[1, 1], G[2, 2], G[0, 1] + G[1, 0], G[0, 2] + G[2, 0], G[1, 2] + G[2, 1]])
    assert np.allclose(st2[0, :, 6:12], eps[None, :], rtol=1e-10, atol=1e-15)


def test_volume_and_body_force_total():
    """sum of w detJ equals the element volume; the body-force vector sums to f * volume."""
    m = mesh.box_mesh(3, 2, 2, 3.0, 2.0, 1.0, nodes_per_element=8, jitter=0.15, seed=4)
    el = element.HexElements("C3D8", m["coords"], m["conn"])
    vol = (el.detJ * el.w[None]).sum()
    assert abs(vol - 6.0) < 1e-12
    f = np.array([1.0, -2.0, 0.5])
    Pe = el.body_force(f)
    assert np.allclose(Pe.reshape(-1, 8, 3).sum(axis=(0, 1)), 6.0 * f)


def test_jittered_mesh_is_seeded_and_bounded():
    a = mesh.box_mesh(4, 4, 4, jitter=0.1, seed=0)["coords"]
    b = mesh.box_mesh(4, 4, 4, jitter=0.1, seed=0)["coords"]
    c = mesh.box_mesh(4, 4, 4, jitter=0.0)["coords"]
    assert np.array_equal(a, b)
    d = np.abs(a - c)
    assert d.max() <= 0.1 * 0.25 + 1e-15 and d.max() > 0.02
    on_boundary = (np.isclose(c, 0) | np.isclose(c, 1)).any(axis=1)
    assert np.array_equal(a[on_boundary], c[on_boundary])


FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "element_fixtures.npz")


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20", "C3D20R"])
def test_geometry_and_body_force_equal_the_reference_element_on_distorted_elements(el_type):
    """J, B (initializeElement, element.py:194-205) and computeBodyForce (:348-371) of the
    reference's own element on three general distorted elements (tests/golden/
    make_element_fixtures.py)."""
    d = np.load(FIX)
    X = d[el_type + "_X"]
    nn = X.shape[1]
    el = element.HexElements(el_type, X.reshape(-1, 3), np.arange(X.shape[0] * nn).reshape(-1, nn))
    assert _rel(el.J, d[el_type + "_J"]) < 1e-13
    assert _rel(element.b_operator(el.dNdx), d[el_type + "_B"]) < 1e-13
    assert _rel(el.body_force(d["body_load"]), d[el_type + "_bodyforce"]) < 1e-13


@pytest.mark.parametrize("el_type", ["C3D20", "C3D20R"])
def test_compute_yourself_equals_the_reference_element_on_distorted_elements(el_type):
    """computeYourself (element.py:290-346) with von Mises and nonlinear hardening over two
    plastic increments, the first accepted: Ke, Pe and the temporary QP state."""
    d = np.load(FIX)
    X = d[el_type + "_X"]
    nn = X.shape[1]
    el = element.HexElements(el_type, X.reshape(-1, 3), np.arange(X.shape[0] * nn).reshape(-1, nn))
    mat = materials.VonMises(d["vonmises_props"])
    Ke, Pe, st = el.compute(d[el_type + "_dU1"], np.zeros((X.shape[0], el.nQP, 13)), mat)
    assert (st[..., 12] > 0).all()  # every QP plastic
    assert _rel(Ke, d[el_type + "_Ke1"]) < 1e-12 and _rel(Pe, d[el_type + "_Pe1"]) < 1e-12
    assert _rel(st, d[el_type + "_state1"]) < 1e-12
    Ke, Pe, st2 = el.compute(d[el_type + "_dU2"], st, mat)
    assert _rel(Ke, d[el_type + "_Ke2"]) < 1e-12 and _rel(Pe, d[el_type + "_Pe2"]) < 1e-12
    assert _rel(st2, d[el_type + "_state2"]) < 1e-12


@pytest.mark.parametrize("el_type", ["C3D20", "C3D20R"])
def test_exact_hexa20_jacobian_does_not_match_the_reference(el_type):
    """The fixtures tell the reference's hexa20 Jacobian from the exact isoparametric one:
    with the exact J (no third-row term) J and Ke differ by far more than rounding, and so
    does the body force, through detJ."""
    d = np.load(FIX)
    X = d[el_type + "_X"]
    dN = element.shape_derivatives(20, element.evaluation_points(el_type))
    J = np.einsum("qar,eac->eqrc", dN, X)
    assert _rel(J, d[el_type + "_J"]) > 1e-2
    xi, eta, zeta, w = element.qp_table(el_type)
    Nt = element.shape_functions(20, np.stack([xi, eta, zeta], axis=1))
    bf = np.einsum("qa,eq,c->eac", Nt, np.linalg.det(J) * w[None], d["body_load"]).reshape(X.shape[0], -1)
    assert _rel(bf, d[el_type + "_bodyforce"]) > 1e-3


def test_reference_hexa20_is_not_rotation_free_on_twisted_elements():
    """A property of the reference element that the oracle keeps: on an element with a
    nonzero twist coefficient a5 a rigid rotation produces strain (the Jacobian's third row
    is not that of the isoparametric map), while on the same element with a5 = 0 it does not."""
    X = _distorted(20, 2)
    mat = materials.LinearElastic([210000.0, 0.3])
    w = np.array([0.3, -0.2, 0.5])
    for Xe, free in ((X, False), (_untwisted(X), True)):
        el = element.HexElements("C3D20", Xe, np.arange(20)[None])
        u = np.cross(w, Xe).ravel()[None] * 1e-3
        _, _, st = el.compute(u, np.zeros((1, el.nQP, 12)), mat)
        assert (np.abs(st[..., 6:12]).max() < 1e-15) == free

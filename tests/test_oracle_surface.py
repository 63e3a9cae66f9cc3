"""Oracle surface traction (oracle/surface.py).  The reference's DisplacementElement has no
computeDistributedLoad, so these checks pin the consistent nodal forces by what the
mathematics fixes:
* undistorted faces: A/4 per node of a 4-node face; -A/12 per corner and A/3 per mid-side
  node of an 8-node serendipity face (closed forms);
* planar but distorted faces: every nodal force equals a 12x12-point Gauss integration of
  N_a |x_s x x_t| (2x2 and 3x3 points are exact there: |x_s x x_t| is linear in s and t on a
  planar bilinear patch), and the forces sum to traction x polygon area (shoelace formula);
* a box of a jittered mesh: the forces on each side sum to traction x side area;
* the cantilever of 20x4x4 C3D8 on a 10x2x2 box, fixed at x = 0, traction -1 in z on x = 10
  (the CPU reference configuration): tip deflection within 10 % of Timoshenko beam theory,
  with the sign of the load.
"""

import numpy as np
import pytest

from helpers import oracle_model
from oracle import nist as onist
from oracle.surface import FACES8, FACES20, face_integrals, face_shape, surface_traction

FACE_OF = {"bottom": 1, "top": 2, "left": 3, "front": 4, "right": 5, "back": 6}


def _one_element(el_type, lx=2.0, ly=3.0, lz=4.0):
    return oracle_model((1, 1, 1), (lx, ly, lz), el_type, "linearelastic", [1.0, 0.3])


@pytest.mark.parametrize("face,area", [(1, 2.0 * 4.0), (2, 2.0 * 4.0), (3, 3.0 * 4.0), (4, 2.0 * 3.0),
                                       (5, 3.0 * 4.0), (6, 2.0 * 3.0)])
def test_undistorted_faces_closed_form(face, area):
    om = _one_element("C3D8")
    nodes, f = face_integrals(om.coords, om.elements.conn, 0, face)
    np.testing.assert_allclose(f, area / 4, rtol=1e-14)
    om = _one_element("C3D20")
    nodes, f = face_integrals(om.coords, om.elements.conn, 0, face)
    np.testing.assert_allclose(f[:4], -area / 12, rtol=1e-13)
    np.testing.assert_allclose(f[4:], area / 3, rtol=1e-13)


def test_face_nodes_lie_on_the_named_box_side():
    for el_type, faces in (("C3D8", FACES8), ("C3D20", FACES20)):
        om = _one_element(el_type)
        X = om.coords[om.elements.conn[0]]
        for name, fid in FACE_OF.items():
            axis, val = {"bottom": (1, 0.0), "top": (1, 3.0), "left": (0, 0.0), "right": (0, 2.0),
                         "back": (2, 0.0), "front": (2, 4.0)}[name]
            np.testing.assert_allclose(X[faces[fid - 1], axis], val)


def _polygon_area(P):
    n = np.cross(P[1] - P[0], P[3] - P[0])
    n /= np.linalg.norm(n)
    return 0.5 * abs(sum(np.dot(n, np.cross(P[i], P[(i + 1) % 4])) for i in range(4)))


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20"])
def test_planar_distorted_face_matches_brute_force(el_type):
    om = _one_element(el_type)
    conn = om.elements.conn
    coords = om.coords.copy()
    faces = FACES8 if el_type == "C3D8" else FACES20
    # distort face 5 (x = 2) inside its plane, keeping mid-side nodes at edge midpoints
    fn = faces[4]
    coords[conn[0, fn[0]], 1:] += (0.4, 0.3)
    coords[conn[0, fn[2]], 1:] += (-0.2, 0.5)
    if el_type == "C3D20":
        for k, (a, b) in enumerate(((0, 1), (1, 2), (2, 3), (3, 0))):
            coords[conn[0, fn[4 + k]]] = 0.5 * (coords[conn[0, fn[a]]] + coords[conn[0, fn[b]]])
    nodes, f = face_integrals(coords, conn, 0, 5)
    X = coords[conn[0, nodes]]
    g, w = np.polynomial.legendre.leggauss(12)
    brute = np.zeros(nodes.size)
    for i, s in enumerate(g):
        for j, t in enumerate(g):
            N, Ns, Nt = face_shape(nodes.size, s, t)
            brute += N * np.linalg.norm(np.cross(Ns @ X, Nt @ X)) * w[i] * w[j]
    np.testing.assert_allclose(f, brute, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(f.sum(), _polygon_area(X[:4]), rtol=1e-13)


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20R"])
def test_box_sides_of_jittered_mesh_sum_to_traction_times_area(el_type):
    om = oracle_model((3, 2, 4), (1.5, 2.0, 2.5), el_type, "linearelastic", [1.0, 0.3], jitter=0.2, seed=3)
    t = np.array([0.3, -1.2, 0.7])
    areas = {"bottom": 1.5 * 2.5, "top": 1.5 * 2.5, "left": 2.0 * 2.5, "right": 2.0 * 2.5,
             "back": 1.5 * 2.0, "front": 1.5 * 2.0}
    for name, fid in FACE_OF.items():
        P = surface_traction(om.coords, om.elements.conn, {fid: om.elsets["gen_" + name]}, t, om.nDof)
        np.testing.assert_allclose(P.reshape(-1, 3).sum(axis=0), t * areas[name], rtol=1e-12)
        # only nodes of that side are loaded
        loaded = np.flatnonzero(np.abs(P.reshape(-1, 3)).sum(axis=1))
        assert set(loaded) <= set(om.nodesets["gen_" + name])


def test_cantilever_tip_deflection_against_beam_theory():
    E, nu = 210000.0, 0.3
    om = oracle_model((20, 4, 4), (10.0, 2.0, 2.0), "C3D8", "linearelastic", [E, nu])
    st = onist.Step(dirichlets=[dict(nset="gen_left", values={0: 0.0, 1: 0.0, 2: 0.0})],
                    distributedloads=[dict(surface={5: "gen_right"}, vector=(0.0, 0.0, -1.0))])
    so = onist.NIST()
    so.solve_step(om, st)
    tip = om.U.reshape(-1, 3)[om.nodesets["gen_right"], 2].mean()
    P, L, b, h = 4.0, 10.0, 2.0, 2.0
    I, G = b * h ** 3 / 12, E / (2 * (1 + nu))
    timoshenko = P * L ** 3 / (3 * E * I) + P * L / (5 / 6 * G * b * h)
    assert tip < 0
    assert abs(-tip / timoshenko - 1) < 0.10

"""Write element_fixtures.npz: outputs of the reference's own DisplacementElement on
distorted elements, for the element functions that no golden vector of the reference
tells apart (every reference test mesh is undistorted, where the exchanged evaluation
point and the body-force pairing leave no trace):

* hexa20 with 27 QPs (C3D20) and 8 QPs (C3D20R): J and B per QP
  (_elementcomputationmatrices.py:369-540 _J3D20, :820-1060 _B3D20, evaluated by
  DisplacementElement.initializeElement, element.py:194-205), and computeYourself
  (element.py:290-346) with the von Mises material (vonmises.py:186-254, nonlinear
  hardening) over two increments, the first accepted (element.py:373-379): Ke, Pe and the
  temporary QP state of each increment;
* computeBodyForce (element.py:348-371: N from computeNOperator, :106-212, at the table
  point, detJ at the exchanged point) on distorted hexa8, hexa20 and hexa20R elements.

Run in this container only, with the reference checkout at EDELWEISSFE_REF (default
/root/reference); the modules it loads need numpy only.  Only the inputs and outputs are
written (tests/golden/element_fixtures.npz); nothing of the reference is stored.
"""

import os
import sys

import numpy as np

REF = os.environ.get("EDELWEISSFE_REF", "/root/reference")
HERE = os.path.dirname(os.path.abspath(__file__))

# node natural coordinates (xi, eta, zeta) of hexa8 / hexa20 in the reference's node order
NAT20 = np.array([(-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1), (-1, 1, -1), (-1, 1, 1), (1, 1, 1),
                  (1, 1, -1), (-1, -1, 0), (0, -1, 1), (1, -1, 0), (0, -1, -1), (-1, 1, 0), (0, 1, 1),
                  (1, 1, 0), (0, 1, -1), (-1, 0, -1), (-1, 0, 1), (1, 0, 1), (1, 0, -1)], dtype=float)

# stresses kept small (E = 2100, fy0 = 2.5) so that the local Newton's absolute 1e-12 tolerance
# (vonmises.py:55-56) is far above the rounding of its residual; nonlinear hardening
VONMISES = np.array([2100.0, 0.3, 2.5, 10.0, 0.8, 300.0])
BODY_LOAD = np.array([0.3, -9.81, 1.7])


def distorted_element(nn, rs):
    """A general distorted hexahedron: affine map of the natural cube plus independent
    perturbations of every node (mid-side nodes off the edge midpoints too)."""
    A = np.diag([1.2, 0.9, 1.1]) + 0.15 * rs.randn(3, 3)
    X = 0.5 * (NAT20[:nn] @ A.T) + rs.uniform(-2, 2, 3)
    X += 0.06 * rs.randn(nn, 3)
    return X


def main():
    sys.path.insert(0, REF)
    from edelweissfe.elements.displacementelement.element import DisplacementElement
    from edelweissfe.materials.vonmises.vonmises import VonMisesMaterial
    from edelweissfe.points.node import Node

    rs = np.random.RandomState(20241203)
    out = {"vonmises_props": VONMISES, "body_load": BODY_LOAD}
    for el_type, nn in (("C3D8", 8), ("C3D20", 20), ("C3D20R", 20)):
        n_el = 3
        X = np.stack([distorted_element(nn, rs) for _ in range(n_el)])
        out[el_type + "_X"] = X
        Js, Bs, Pbf = [], [], []
        Ke1, Pe1, S1, Ke2, Pe2, S2, dU1s, dU2s = [], [], [], [], [], [], [], []
        for e in range(n_el):
            el = DisplacementElement(el_type, e + 1)
            el.setNodes([Node(a + 1, X[e, a].copy()) for a in range(nn)])
            el.setProperties(np.array([]))
            el.initializeElement()
            el.setMaterial(VonMisesMaterial(VONMISES))
            Js.append(np.array(el.J))
            Bs.append(np.array(el.B))
            nd = el.nDof
            # body force
            Pb = np.zeros(nd)
            el.computeBodyForce(Pb, np.zeros(nd * nd), BODY_LOAD, np.zeros(nd), np.zeros(2), 1.0)
            Pbf.append(Pb)
            if nn == 8:
                continue  # hexa8 computeYourself is pinned by the reference's golden vectors
            # two increments: elastic-plastic, accepted; then a further plastic increment
            dU1 = 4e-3 * rs.randn(nd)
            dU2 = 3e-3 * rs.randn(nd)
            K = np.zeros(nd * nd)
            P = np.zeros(nd)
            el.computeYourself(K, P, np.zeros(nd), dU1, np.zeros(2), 1.0)
            Ke1.append(K.reshape(nd, nd).copy())
            Pe1.append(P.copy())
            S1.append(np.array(el._stateVarsTemp))
            el.acceptLastState()
            K = np.zeros(nd * nd)
            P = np.zeros(nd)
            el.computeYourself(K, P, dU1 + dU2, dU2, np.zeros(2), 1.0)
            Ke2.append(K.reshape(nd, nd).copy())
            Pe2.append(P.copy())
            S2.append(np.array(el._stateVarsTemp))
            dU1s.append(dU1)
            dU2s.append(dU2)
        out[el_type + "_J"] = np.array(Js)
        out[el_type + "_B"] = np.array(Bs)
        out[el_type + "_bodyforce"] = np.array(Pbf)
        if nn == 20:
            out[el_type + "_dU1"] = np.array(dU1s)
            out[el_type + "_dU2"] = np.array(dU2s)
            out[el_type + "_Ke1"] = np.array(Ke1)
            out[el_type + "_Pe1"] = np.array(Pe1)
            out[el_type + "_state1"] = np.array(S1)
            out[el_type + "_Ke2"] = np.array(Ke2)
            out[el_type + "_Pe2"] = np.array(Pe2)
            out[el_type + "_state2"] = np.array(S2)
    path = os.path.join(HERE, "element_fixtures.npz")
    np.savez_compressed(path, **out)
    for k, v in out.items():
        print(k, v.shape)
    print("wrote", path)


if __name__ == "__main__":
    main()

"""Surface traction on hexahedron faces (TEST INFRASTRUCTURE ONLY).

The reference applies distributed loads element by element
(NIST.computeDistributedLoads, nonlinearimplicitstatic.py:460-508: for every face ID and
element of a surface, el.computeDistributedLoad(...) and PExt[el] += Pe).  Its Python
DisplacementElement has no computeDistributedLoad (element.py:288), so the element side
here has no reference implementation: it is the consistent nodal force of a uniform
traction t on a face, f_a = t * integral N_a dA, and is pinned in the tests by what the
mathematics fixes (the forces sum to t * face area; on an undistorted face the 4-node face
gives A/4 per node and the 8-node serendipity face -A/12 per corner and A/3 per mid-side
node; a traction on all six faces of a box balances).

Faces are numbered by local node numbers as in Abaqus (face 1: nodes 1-2-3-4, 2: 5-8-7-6,
3: 1-5-6-2, 4: 2-6-7-3, 5: 3-7-8-4, 6: 4-8-5-1; hexa20 adds the mid-side nodes between
consecutive corners), the numbering the reference's box generator uses for its surfaces
(boxgen.py:438-446: bottom 1, top 2, left 3, front 4, right 5, back 6).  Integration: 2x2
Gauss points on 4-node faces, 3x3 on 8-node faces.
"""

import numpy as np

# 0-based local nodes of each face: corners around the face, then (hexa20) the mid-side
# nodes c0-c1, c1-c2, c2-c3, c3-c0
FACES8 = np.array([[0, 1, 2, 3], [4, 7, 6, 5], [0, 4, 5, 1], [1, 5, 6, 2], [2, 6, 7, 3], [3, 7, 4, 0]])
FACES20 = np.array([[0, 1, 2, 3, 8, 9, 10, 11], [4, 7, 6, 5, 15, 14, 13, 12], [0, 4, 5, 1, 16, 12, 17, 8],
                    [1, 5, 6, 2, 17, 13, 18, 9], [2, 6, 7, 3, 18, 14, 19, 10], [3, 7, 4, 0, 19, 15, 16, 11]])
# natural coordinates (s, t) of the face nodes in that order
FACE_NAT = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1], [0, -1], [1, 0], [0, 1], [-1, 0]], dtype=float)


def face_shape(nf, s, t):
    """N, dN/ds, dN/dt of the 4-node bilinear or 8-node serendipity face at (s, t)."""
    si, ti = FACE_NAT[:nf, 0], FACE_NAT[:nf, 1]
    if nf == 4:
        N = (1 + s * si) * (1 + t * ti) / 4
        Ns = si * (1 + t * ti) / 4
        Nt = ti * (1 + s * si) / 4
        return N, Ns, Nt
    N, Ns, Nt = np.empty(8), np.empty(8), np.empty(8)
    c = slice(0, 4)
    N[c] = (1 + s * si[c]) * (1 + t * ti[c]) * (s * si[c] + t * ti[c] - 1) / 4
    Ns[c] = si[c] * (1 + t * ti[c]) * (2 * s * si[c] + t * ti[c]) / 4
    Nt[c] = ti[c] * (1 + s * si[c]) * (s * si[c] + 2 * t * ti[c]) / 4
    for k in range(4, 8):
        if si[k] == 0:
            N[k] = (1 - s * s) * (1 + t * ti[k]) / 2
            Ns[k] = -s * (1 + t * ti[k])
            Nt[k] = ti[k] * (1 - s * s) / 2
        else:
            N[k] = (1 + s * si[k]) * (1 - t * t) / 2
            Ns[k] = si[k] * (1 - t * t) / 2
            Nt[k] = -t * (1 + s * si[k])
    return N, Ns, Nt


def _gauss(n):
    if n == 2:
        return np.array([-1, 1]) / np.sqrt(3), np.array([1.0, 1.0])
    return np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]), np.array([5 / 9, 8 / 9, 5 / 9])


def face_integrals(coords, conn, element, face):
    """integral N_a dA for the local nodes a of `face` (1..6) of `element`: (local nodes, values)."""
    nn = conn.shape[1]
    nodes = (FACES8 if nn == 8 else FACES20)[face - 1]
    nf = nodes.size
    X = coords[conn[element, nodes]]
    g, w = _gauss(2 if nf == 4 else 3)
    out = np.zeros(nf)
    for i, s in enumerate(g):
        for j, t in enumerate(g):
            N, Ns, Nt = face_shape(nf, s, t)
            dA = np.linalg.norm(np.cross(Ns @ X, Nt @ X)) * w[i] * w[j]
            out += N * dA
    return nodes, out


def surface_traction(coords, conn, surface, traction, n_dof):
    """PExt contribution of a uniform traction (3,) on surface {face ID: element indices}, in
    the reference's loop order (face IDs, then elements, nonlinearimplicitstatic.py:497-504)."""
    P = np.zeros(n_dof)
    t = np.asarray(traction, dtype=float)
    for face, elements in surface.items():
        for e in elements:
            nodes, f = face_integrals(coords, conn, int(e), int(face))
            for a, v in zip(nodes, f):
                P[3 * conn[e, a]:3 * conn[e, a] + 3] += v * t
    return P

"""Structured box meshes in boxGen order (oracle side).

Follows edelweissfe/generators/boxgen.py:
  * lattice coordinates by ``np.linspace`` per axis (:125-127);
  * node loop ix -> iy -> iz, hexa20 keeps lattice nodes with fewer than two odd
    indices (:133-141); node position = insertion order of the kept nodes;
  * hexa8 connectivity (:171-185), hexa20 connectivity (:186-299);
  * face / edge / vertex node sets in the order numpy's boolean indexing of the
    lattice gives (:321-366); element sets (:368-383).

Interior-node jitter is this project's synthetic-input definition, not part of
the reference: every node not on a boundary face moves by
(2u-1) * frac * h per axis, u = rng.uniform(seed, 3*pos + axis), h = l/n.
"""

import numpy as np

from . import rng


def _lattice_index(nNY, nNZ, ix, iy, iz):
    return ix * (nNY * nNZ) + iy * nNZ + iz


def box_mesh(nX, nY, nZ, lX=1.0, lY=1.0, lZ=1.0, x0=0.0, y0=0.0, z0=0.0, nodes_per_element=8,
             jitter=0.0, seed=0, name="gen"):
    if nodes_per_element == 8:
        nNX, nNY, nNZ = nX + 1, nY + 1, nZ + 1
    elif nodes_per_element == 20:
        nNX, nNY, nNZ = 2 * nX + 1, 2 * nY + 1, 2 * nZ + 1
    else:
        raise ValueError("hexa8 or hexa20 only")

    xL = np.linspace(x0, x0 + lX, nNX)
    yL = np.linspace(y0, y0 + lY, nNY)
    zL = np.linspace(z0, z0 + lZ, nNZ)
    IX, IY, IZ = np.meshgrid(np.arange(nNX), np.arange(nNY), np.arange(nNZ), indexing="ij")
    if nodes_per_element == 8:
        keep = np.ones(IX.shape, dtype=bool)
    else:
        keep = (IX % 2 + IY % 2 + IZ % 2) < 2
    pos = -np.ones(IX.shape, dtype=np.int64)
    pos[keep] = np.arange(int(keep.sum()))
    coords = np.stack([xL[IX[keep]], yL[IY[keep]], zL[IZ[keep]]], axis=1)

    if jitter:
        h = np.array([lX / nX, lY / nY, lZ / nZ])
        interior = ((IX > 0) & (IX < nNX - 1) & (IY > 0) & (IY < nNY - 1) & (IZ > 0) & (IZ < nNZ - 1))[keep]
        p = np.arange(coords.shape[0], dtype=np.uint64)
        for c in range(3):
            u = rng.uniform(seed, 3 * p + np.uint64(c))
            coords[:, c] = np.where(interior, coords[:, c] + (2.0 * u - 1.0) * jitter * h[c], coords[:, c])

    ex, ey, ez = np.meshgrid(np.arange(nX), np.arange(nY), np.arange(nZ), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    if nodes_per_element == 8:
        offs = [(0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)]
        s = 1
    else:
        offs = [(0, 0, 0), (0, 0, 2), (2, 0, 2), (2, 0, 0), (0, 2, 0), (0, 2, 2), (2, 2, 2), (2, 2, 0),
                (0, 0, 1), (1, 0, 2), (2, 0, 1), (1, 0, 0), (0, 2, 1), (1, 2, 2), (2, 2, 1), (1, 2, 0),
                (0, 1, 0), (0, 1, 2), (2, 1, 2), (2, 1, 0)]
        s = 2
    conn = np.empty((ex.size, nodes_per_element), dtype=np.int64)
    for a, (ox, oy, oz) in enumerate(offs):
        conn[:, a] = pos[s * ex + ox, s * ey + oy, s * ez + oz]
    assert (conn >= 0).all()

    def nset(sl):
        sub = pos[sl]
        return sub[sub >= 0].ravel()

    ns = {}
    ns[name + "_top"] = nset(np.s_[:, -1, :])
    ns[name + "_bottom"] = nset(np.s_[:, 0, :])
    ns[name + "_right"] = nset(np.s_[-1, :, :])
    ns[name + "_left"] = nset(np.s_[0, :, :])
    ns[name + "_front"] = nset(np.s_[:, :, -1])
    ns[name + "_back"] = nset(np.s_[:, :, 0])
    ns[name + "_bottomRight"] = nset(np.s_[-1, 0, :])
    ns[name + "_bottomLeft"] = nset(np.s_[0, 0, :])
    ns[name + "_bottomFront"] = nset(np.s_[:, 0, -1])
    ns[name + "_bottomBack"] = nset(np.s_[:, 0, 0])
    ns[name + "_topRight"] = nset(np.s_[-1, -1, :])
    ns[name + "_topLeft"] = nset(np.s_[0, -1, :])
    ns[name + "_topFront"] = nset(np.s_[:, -1, -1])
    ns[name + "_topBack"] = nset(np.s_[:, -1, 0])
    ns[name + "_rightBack"] = nset(np.s_[-1, :, 0])
    ns[name + "_rightFront"] = nset(np.s_[-1, :, -1])
    ns[name + "_leftBack"] = nset(np.s_[0, :, 0])
    ns[name + "_leftFront"] = nset(np.s_[0, :, -1])
    ns["all"] = np.arange(coords.shape[0])

    elgrid = np.arange(ex.size).reshape(nX, nY, nZ)
    es = {
        "all": np.arange(ex.size),
        name + "_all": np.arange(ex.size),
        name + "_bottom": elgrid[:, 0, :].ravel(),
        name + "_top": elgrid[:, -1, :].ravel(),
        name + "_right": elgrid[-1, :, :].ravel(),
        name + "_left": elgrid[0, :, :].ravel(),
        name + "_front": elgrid[:, :, -1].ravel(),
        name + "_back": elgrid[:, :, 0].ravel(),
    }
    return {"coords": coords, "conn": conn, "nodesets": ns, "elsets": es,
            "nodes_per_element": nodes_per_element, "shape": (nX, nY, nZ)}

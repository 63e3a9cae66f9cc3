"""Structured hexahedral boxes in the reference's boxGen order.

Mirrors edelweissfe/generators/boxgen.py:91-448 so that node positions, element
order, DOF numbering and set contents equal the reference's for the same
options: nodes in lattice order ix -> iy -> iz (hexa20 keeps lattice nodes with
fewer than two odd indices, :133-141), hexa8 / hexa20 node lists (:171-299),
face/edge node sets in numpy boolean-index order (:321-366), element sets
(:368-383).

Seeded interior-node jitter (this project's synthetic input definition): every
node strictly inside the box moves by (2u - 1) * jitter * h along each axis,
with u the SplitMix64 uniform of counter 3*pos + axis and h = l/n.
"""

import numpy as np

from ..model import FEModel
from ..utils import rng

_H8 = ((0, 0, 0), (0, 0, 1), (1, 0, 1), (1, 0, 0), (0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0))
_H20 = ((0, 0, 0), (0, 0, 2), (2, 0, 2), (2, 0, 0), (0, 2, 0), (0, 2, 2), (2, 2, 2), (2, 2, 0),
        (0, 0, 1), (1, 0, 2), (2, 0, 1), (1, 0, 0), (0, 2, 1), (1, 2, 2), (2, 2, 1), (1, 2, 0),
        (0, 1, 0), (0, 1, 2), (2, 1, 2), (2, 1, 0))


def generateBox(nX=1, nY=1, nZ=1, lX=1.0, lY=1.0, lZ=1.0, x0=0.0, y0=0.0, z0=0.0, elType="C3D8", name="gen",
                jitter=0.0, seed=0, material=None):
    elType = elType.upper()
    quadratic = elType.startswith("C3D20")
    step = 2 if quadratic else 1
    nx, ny, nz = step * nX + 1, step * nY + 1, step * nZ + 1
    ix, iy, iz = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    keep = np.ones(ix.shape, bool) if not quadratic else ((ix & 1) + (iy & 1) + (iz & 1) < 2)
    position = np.full(ix.shape, -1, dtype=np.int64)
    position[keep] = np.arange(np.count_nonzero(keep))
    xs = np.linspace(x0, x0 + lX, nx)
    ys = np.linspace(y0, y0 + lY, ny)
    zs = np.linspace(z0, z0 + lZ, nz)
    coords = np.column_stack([xs[ix[keep]], ys[iy[keep]], zs[iz[keep]]])
    if jitter:
        inner = ((ix > 0) & (ix < nx - 1) & (iy > 0) & (iy < ny - 1) & (iz > 0) & (iz < nz - 1))[keep]
        h = (lX / nX, lY / nY, lZ / nZ)
        counter = np.arange(coords.shape[0], dtype=np.uint64) * np.uint64(3)
        for axis in range(3):
            u = rng.uniform(seed, counter + np.uint64(axis))
            coords[inner, axis] += ((2.0 * u - 1.0) * jitter * h[axis])[inner]

    ex, ey, ez = (a.ravel() for a in np.meshgrid(np.arange(nX), np.arange(nY), np.arange(nZ), indexing="ij"))
    offsets = _H20 if quadratic else _H8
    conn = np.stack([position[step * ex + o[0], step * ey + o[1], step * ez + o[2]] for o in offsets], axis=1)

    def ns(sl):
        p = position[sl]
        return p[p >= 0].ravel()

    nodeSets = {
        "all": np.arange(coords.shape[0]),
        name + "_top": ns(np.s_[:, -1, :]), name + "_bottom": ns(np.s_[:, 0, :]),
        name + "_right": ns(np.s_[-1, :, :]), name + "_left": ns(np.s_[0, :, :]),
        name + "_front": ns(np.s_[:, :, -1]), name + "_back": ns(np.s_[:, :, 0]),
        name + "_bottomRight": ns(np.s_[-1, 0, :]), name + "_bottomLeft": ns(np.s_[0, 0, :]),
        name + "_bottomFront": ns(np.s_[:, 0, -1]), name + "_bottomBack": ns(np.s_[:, 0, 0]),
        name + "_topRight": ns(np.s_[-1, -1, :]), name + "_topLeft": ns(np.s_[0, -1, :]),
        name + "_topFront": ns(np.s_[:, -1, -1]), name + "_topBack": ns(np.s_[:, -1, 0]),
        name + "_rightBack": ns(np.s_[-1, :, 0]), name + "_rightFront": ns(np.s_[-1, :, -1]),
        name + "_leftBack": ns(np.s_[0, :, 0]), name + "_leftFront": ns(np.s_[0, :, -1]),
    }
    grid = np.arange(ex.size).reshape(nX, nY, nZ)
    elementSets = {
        "all": np.arange(ex.size), name + "_all": np.arange(ex.size),
        name + "_bottom": grid[:, 0, :].ravel(), name + "_top": grid[:, -1, :].ravel(),
        name + "_right": grid[-1, :, :].ravel(), name + "_left": grid[0, :, :].ravel(),
        name + "_front": grid[:, :, -1].ravel(), name + "_back": grid[:, :, 0].ravel(),
    }
    # surfaces as boxgen.py:438-446: face IDs by Abaqus local node numbering
    surfaces = {name + "_" + side: {fid: elementSets[name + "_" + side]}
                for side, fid in (("bottom", 1), ("top", 2), ("left", 3), ("front", 4), ("right", 5), ("back", 6))}
    return FEModel(coords, elType, conn, nodeSets, elementSets, material=material, surfaces=surfaces)

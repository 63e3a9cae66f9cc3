"""Seeded batches of independent J2 cube specimens of uneven size.

Specimen s of a batch with seed `seed` draws three SplitMix64 uniforms of that stream,
u0 = u(3 s), u1 = u(3 s + 1), u2 = u(3 s + 2):
  n_s   = nmin + floor(u0 (nmax - nmin + 1))   elements per edge of a unit cube (C3D8),
  E_s   = E_lo + u1 (E_hi - E_lo),
  fy_s  = fy_lo + u2 (fy_hi - fy_lo),
with nu = 0.3 and linear hardening H = 1000 (J2 props [E, nu, fy, H, 0, 0]).  Interior
nodes are jittered by 10 % of h with the specimen index as jitter seed.  Every specimen
is compressed uniaxially along y (u_y = -strain on top, 0 on bottom, symmetry u_x = 0
on x = 0 and u_z = 0 on z = 0).
"""

import numpy as np

from ..steps import AdaptiveStep, Dirichlet
from ..utils import rng
from .boxgen import generateBox


def specimenParameters(count=512, seed=2, nmin=8, nmax=24, E=(150000.0, 250000.0), fy=(150.0, 350.0)):
    k = np.arange(count, dtype=np.uint64) * np.uint64(3)
    u0, u1, u2 = (rng.uniform(seed, k + np.uint64(i)) for i in range(3))
    n = nmin + np.minimum(np.floor(u0 * (nmax - nmin + 1)).astype(np.int64), nmax - nmin)
    return n, E[0] + u1 * (E[1] - E[0]), fy[0] + u2 * (fy[1] - fy[0])


def specimenBatch(count=512, seed=2, nmin=8, nmax=24, strain=0.01, increments=10, jitter=0.1, maxIter=15):
    """Returns (models, steps, parameters) of the batch."""
    n, E, fy = specimenParameters(count, seed, nmin, nmax)
    models, steps = [], []
    for s in range(count):
        m = generateBox(int(n[s]), int(n[s]), int(n[s]), 1.0, 1.0, 1.0, elType="C3D8", jitter=jitter, seed=s)
        m.assignMaterial("vonmises", [E[s], 0.3, fy[s], 1000.0, 0.0, 0.0])
        acts = {"dirichlet": {
            "top": Dirichlet("top", "gen_top", {1: -strain}),
            "bottom": Dirichlet("bottom", "gen_bottom", {1: 0.0}),
            "left": Dirichlet("left", "gen_left", {0: 0.0}),
            "back": Dirichlet("back", "gen_back", {2: 0.0})}}
        models.append(m)
        steps.append(AdaptiveStep(1, m, acts, maxInc=1.0 / increments, minInc=1e-6, maxNumInc=1000, maxIter=maxIter))
    return models, steps, {"n": n, "E": E, "fy": fy}

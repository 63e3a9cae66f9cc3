"""Shared builders: the same case as an oracle model and as a device model."""

import numpy as np

from edelweissfe_amd.generators.boxgen import generateBox
from edelweissfe_amd.steps import AdaptiveStep, BodyForce, Dirichlet, NodeForces
from oracle import materials as omat
from oracle import mesh as omesh
from oracle import nist as onist

NPE = {"C3D8": 8, "C3D20": 20, "C3D20R": 20}


def oracle_model(n, l, el_type, material, props, jitter=0.0, seed=0):
    m = omesh.box_mesh(*n, *l, nodes_per_element=NPE[el_type], jitter=jitter, seed=seed)
    mat = omat.MATERIALS[material](props)
    return onist.Model(m["coords"], el_type, m["conn"], mat, m["nodesets"], m["elsets"])


def device_model(n, l, el_type, material, props, jitter=0.0, seed=0):
    pm = generateBox(*n, *l, elType=el_type, jitter=jitter, seed=seed)
    pm.assignMaterial(material, props)
    return pm


def uniaxial_bcs(strain, ly=1.0):
    """u_y = -strain*ly on top, u_y = 0 on bottom, symmetry u_x = 0 on left, u_z = 0 on back."""
    return [
        ("top", "gen_top", {1: -strain * ly}),
        ("bottom", "gen_bottom", {1: 0.0}),
        ("left", "gen_left", {0: 0.0}),
        ("back", "gen_back", {2: 0.0}),
    ]


def oracle_step(bcs, nodeforces=(), bodyforces=(), **kw):
    return onist.Step(dirichlets=[dict(nset=s, values=v) for _, s, v in bcs],
                      nodeforces=[dict(nset=s, vector=vec) for s, vec in nodeforces],
                      bodyforces=[dict(elset=s, vector=vec) for s, vec in bodyforces], **kw)


def device_step(model, bcs, nodeforces=(), bodyforces=(), extrapolation="linear", **kw):
    actions = {
        "dirichlet": {name: Dirichlet(name, s, v) for name, s, v in bcs},
        "nodeforces": {"nf%d" % i: NodeForces("nf%d" % i, s, vec) for i, (s, vec) in enumerate(nodeforces)},
        "bodyforce": {"bf%d" % i: BodyForce("bf%d" % i, s, vec) for i, (s, vec) in enumerate(bodyforces)},
        "options": {"NISTSolver": {"extrapolation": extrapolation}},
    }
    return AdaptiveStep(1, model, actions, **kw)


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)

"""Surface traction on the device (efe_surface_traction) against the oracle
(oracle/surface.py), and the CPU reference configuration end to end: a 20x4x4 C3D8
cantilever on a 10x2x2 box, linear elastic (E 210000, nu 0.3), fixed at x = 0, uniform
traction -1 in z on the face x = 10, one increment, FP64."""

import numpy as np
import pytest

from helpers import device_model, oracle_model, rel_err
from oracle import nist as onist
from oracle.surface import surface_traction

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("el_type,n", [("C3D8", (4, 3, 5)), ("C3D20", (2, 3, 2)), ("C3D20R", (3, 2, 2))])
def test_surface_traction_matches_oracle(el_type, n):
    from edelweissfe_amd.device import DeviceSystem

    om = oracle_model(n, (1.0, 1.5, 2.0), el_type, "linearelastic", [1.0, 0.3], jitter=0.15, seed=5)
    pm = device_model(n, (1.0, 1.5, 2.0), el_type, "linearelastic", [1.0, 0.3], jitter=0.15, seed=5)
    sysm = DeviceSystem(pm.coordinates, pm.connectivity, pm.elementType, "linearelastic", [1.0, 0.3])
    t = (0.4, -1.3, 0.9)
    # two surfaces sharing edge elements (an element with two loaded faces)
    surf = {}
    for name in ("gen_right", "gen_top"):
        for f, els in pm.surfaces[name].items():
            surf.setdefault(f, [])
            surf[f] = np.concatenate([surf[f], els]).astype(np.int64)
    PExt = sysm.zeros()
    PExt += 0.5  # the entry point adds into PExt
    sysm.surface_traction(surf, t, PExt)
    want = 0.5 + surface_traction(om.coords, om.elements.conn, surf, t, om.nDof)
    got = PExt.cpu().numpy()
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)
    # an empty surface adds nothing
    before = PExt.clone()
    sysm.surface_traction({}, t, PExt)
    assert torch.equal(before, PExt)
    sysm.close()


def test_cantilever_with_end_traction_matches_oracle():
    from edelweissfe_amd.solvers import NISTGPU
    from edelweissfe_amd.steps import AdaptiveStep, Dirichlet, DistributedLoad

    E, nu = 210000.0, 0.3
    n, l = (20, 4, 4), (10.0, 2.0, 2.0)
    om = oracle_model(n, l, "C3D8", "linearelastic", [E, nu])
    so = onist.NIST()
    so.solve_step(om, onist.Step(dirichlets=[dict(nset="gen_left", values={0: 0.0, 1: 0.0, 2: 0.0})],
                                 distributedloads=[dict(surface={5: "gen_right"}, vector=(0.0, 0.0, -1.0))]))
    pm = device_model(n, l, "C3D8", "linearelastic", [E, nu])
    actions = {"dirichlet": {"left": Dirichlet("left", "gen_left", {0: 0.0, 1: 0.0, 2: 0.0})},
               "distributedload": {"tip": DistributedLoad("tip", "gen_right", (0.0, 0.0, -1.0))},
               "options": {"NISTSolver": {"extrapolation": "linear"}}}
    s = NISTGPU(precision="fp64")
    s.solveStep(AdaptiveStep(1, pm, actions), pm)
    assert [h[2] for h in so.history] == [r["iters"] for r in s.history]
    assert rel_err(pm.U, om.U) < 1e-9
    tip = pm.U.reshape(-1, 3)[pm.nodeSets["gen_right"], 2].mean()
    assert tip < 0

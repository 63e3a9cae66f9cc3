"""Increment by increment against the oracle: after every converged increment of a plastic
load history, the device U and QP stresses equal the oracle's to 1e-9 relative, and every
increment takes the oracle's Newton iterations (failed increments included).  The device
values are taken by an output manager, called where the reference calls its output managers
(nonlinearimplicitstatic.py:304-315), so this also checks that the solver reports every
increment with the state it accepted.  Cases cover the element types and materials of the
path: hexa20 (27 QP) and hexa20R with J2, hexa8 and hexa20 with non-associative
Drucker-Prager (GMRES + ILU(0))."""

import numpy as np
import pytest

from helpers import device_model, device_step, oracle_model, oracle_step, rel_err, uniaxial_bcs
from oracle import nist as onist

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]


class _Recorder:
    """Output manager that keeps U and the accepted QP stresses of every converged increment."""

    def __init__(self, model, solver):
        self.model, self.solver, self.rows = model, solver, []

    def finalizeIncrement(self, statusInfoDict=None, **kwargs):
        st = self.solver.system.qp_state(0)
        self.rows.append((statusInfoDict["inc"], np.array(self.model.U), st[..., 0:6].copy()))

    def finalizeFailedIncrement(self, **kwargs):
        pass


@pytest.mark.parametrize("el_type,material,props,n,strain,linsolver", [
    ("C3D20", "vonmises", J2, (2, 2, 1), 0.01, "pcg"),
    ("C3D20R", "vonmises", J2, (2, 1, 2), 0.01, "pcg"),
    ("C3D8", "druckerprager", DP, (3, 3, 2), 0.004, "gmres"),
    ("C3D20", "druckerprager", DP, (1, 2, 1), 0.004, "gmres"),
])
def test_every_increment_matches_the_oracle(el_type, material, props, n, strain, linsolver):
    from edelweissfe_amd.solvers import NISTGPU

    kw = dict(maxInc=0.2, minInc=1e-6, maxNumInc=100, maxIter=15)
    om = oracle_model(n, (1.0, 1.0, 1.0), el_type, material, props, jitter=0.1, seed=3)
    so = onist.NIST(convention="colmajor")
    so.solve_step(om, oracle_step(uniaxial_bcs(strain), **kw))
    pm = device_model(n, (1.0, 1.0, 1.0), el_type, material, props, jitter=0.1, seed=3)
    s = NISTGPU(precision="fp64", linsolver=linsolver)
    rec = _Recorder(pm, s)  # reads the accepted state from the solver's device system
    s.solveStep(device_step(pm, uniaxial_bcs(strain), **kw), pm, outputmanagers=[rec])

    assert [r["iters"] for r in s.history] == [w["iters"] if w["converged"] else np.inf for w in so.increments]
    want = [w for w in so.increments if w["converged"]]
    assert len(rec.rows) == len(want) and (om.state[..., 12] > 0).any()
    for (inc, U, sig), w in zip(rec.rows, want):
        assert inc == w["inc"]
        assert rel_err(U, w["U"]) < 1e-9, inc
        assert rel_err(sig, w["stress"]) < 1e-9, inc


@pytest.mark.parametrize("material,props,precision,linsolver", [
    ("vonmises", J2, "fp64", "pcg"), ("vonmises", J2, "mixed", "pcg"), ("druckerprager", DP, "fp64", "gmres")])
def test_residual_first_changes_nothing_but_the_work(material, props, precision, linsolver):
    """residualFirst (the default) checks Newton iteration 0 on a residual-only pass and forms
    the tangent only when a solve follows: the same Newton counts, U and QP state bit for bit as
    with a full element pass at every check, and fewer tangent passes."""
    from edelweissfe_amd.solvers import NISTGPU

    runs = []
    for residualFirst in (True, False):
        pm = device_model((4, 3, 3), (1.0, 1.0, 1.0), "C3D8", material, props, jitter=0.1, seed=6)
        s = NISTGPU(precision=precision, linsolver=linsolver, residualFirst=residualFirst)
        s.solveStep(device_step(pm, uniaxial_bcs(0.006 if material == "vonmises" else 0.003), maxInc=0.1,
                                minInc=1e-6, maxNumInc=100, maxIter=15), pm)
        runs.append((s, pm, pm.qpState.copy()))
    (s1, m1, q1), (s0, m0, q0) = runs
    print([r["iters"] for r in s1.history], s1.elementPasses, s0.elementPasses)
    assert [r["iters"] for r in s1.history] == [r["iters"] for r in s0.history]
    assert 0 in [r["iters"] for r in s1.history]  # increments that converge at their first check
    assert np.array_equal(m1.U, m0.U) and np.array_equal(q1, q0)
    assert s1.elementPasses < s0.elementPasses

"""The drop-in solver on the reference's own model and step objects (stand-ins with their
attributes, tests/reference_standins.py), driven on the CPU with the oracle in place of the
device (tests/device_mock.py): the same increments, Newton counts, U and QP state as the
oracle's NIST run of the same model, for a node field in an order other than the node
labels, element sets, node forces and body forces on subsets, two steps, the option step
action with string values and a direct-solver name, and several sections."""

import numpy as np
import pytest

import device_mock
import reference_standins as rs
from oracle import materials as omat
from oracle import mesh as omesh
from oracle import nist as onist

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
LE = [150000.0, 0.25]


def _mesh(n=(3, 3, 2), el_type="C3D8", seed=3):
    m = omesh.box_mesh(*n, 1.0, 1.0, 1.0, nodes_per_element=8 if el_type == "C3D8" else 20, jitter=0.1, seed=seed)
    return m


def _models(m, el_type, sections, seed=5):
    coords, conn = m["coords"], m["conn"]
    N = coords.shape[0]
    order = np.random.RandomState(seed).permutation(N)  # the node field's (DOF) order
    pos = np.empty(N, dtype=np.int64)
    pos[order] = np.arange(N)
    E = conn.shape[0]
    upper = np.zeros(E, dtype=bool)
    if sections:
        yc = coords[conn][:, :, 1].mean(axis=1)
        upper = yc > 0.5
    mats = [rs.LinearElasticMaterial(LE) if u else rs.VonMisesMaterial(J2) for u in upper]
    rm = rs.reference_model(coords, conn, el_type, mats, node_order=order, node_sets=m["nodesets"],
                            element_sets=m["elsets"])
    nq = 8 if el_type != "C3D20" else 27
    if sections:
        omat_ = omat.Sections([omat.VonMises(J2), omat.LinearElastic(LE)], upper.astype(int), nq)
    else:
        omat_ = omat.VonMises(J2)
    om = onist.Model(coords[order], el_type, pos[conn], omat_, {k: pos[v] for k, v in m["nodesets"].items()},
                     dict(m["elsets"]))
    return rm, om, pos


def _actions(rm, strain, step2=False):
    ns = rm.nodeSets
    top = -strain if not step2 else -2 * strain
    a = {"dirichlet": {"top": rs.Dirichlet(ns["gen_top"], {1: top}),
                       "bottom": rs.Dirichlet(ns["gen_bottom"], {1: 0.0}),
                       "left": rs.Dirichlet(ns["gen_left"], {0: 0.0}),
                       "back": rs.Dirichlet(ns["gen_back"], {2: 0.0})},
         "options": {"NISTSolver": rs.Options(linsolver="pardiso", extrapolation="linear", residualFirst="False")}}
    return a


def _oracle_step(strain, loads, step2=False, **kw):
    top = -strain if not step2 else -2 * strain
    d = [dict(nset="gen_top", values={1: top}), dict(nset="gen_bottom", values={1: 0.0}),
         dict(nset="gen_left", values={0: 0.0}), dict(nset="gen_back", values={2: 0.0})]
    return onist.Step(dirichlets=d, **loads, **kw)


@pytest.fixture
def oracle_device(monkeypatch):
    from edelweissfe_amd.solvers import nistgpu

    monkeypatch.setattr(nistgpu, "DeviceSystem", device_mock.OracleSystem)


@pytest.mark.parametrize("sections", [False, True])
def test_dropin_two_steps_on_reference_objects_equal_oracle(oracle_device, sections):
    from edelweissfe_amd.solvers import NISTGPUDropIn
    from edelweissfe_amd.steps import AdaptiveStep

    m = _mesh()
    rm, om, pos = _models(m, "C3D8", sections)
    kw = dict(maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)
    strain = 0.004
    # loads: node forces on the top-right edge, a body force on the top element layer
    nf = rs.NodeForces(rm.nodeSets["gen_topRight"], (0.0, 0.0, -10.0))
    bf = rs.BodyForce(rm.elementSets["gen_top"], (30.0, 0.0, 0.0))
    solver = NISTGPUDropIn(None, None)
    so = onist.NIST()
    for k, step2 in enumerate((False, True)):
        acts = _actions(rm, strain, step2)
        acts["nodeforces"] = {"nf": nf}
        acts["bodyforce"] = {"bf": bf}
        solver.solveStep(AdaptiveStep(k + 1, rm, acts, **kw), rm, None, [])
        loads = dict(nodeforces=[dict(nset="gen_topRight", vector=(0.0, 0.0, 0.0 if step2 else -10.0),
                                      start=(0.0, 0.0, -10.0) if step2 else 0.0)],
                     bodyforces=[dict(elset="gen_top", vector=(0.0 if step2 else 30.0, 0.0, 0.0),
                                      start=(30.0, 0.0, 0.0) if step2 else 0.0)])
        so.solve_step(om, _oracle_step(strain, loads, step2, **kw))
        U = rm.nodeFields["displacement"]["U"]  # rows in the node field's order = DOF order
        assert np.abs(U.ravel() - om.U).max() <= 1e-9 * np.abs(om.U).max(), k
    assert [r["iters"] for r in solver.device.history] == [h[2] for h in so.history]
    assert (om.state[..., 12] > 0).any()
    # the accepted QP state went back to the reference elements, the time to the model
    st = np.array([el._stateVarsRef[:, :12] for el in rm.elements.values()])
    assert np.abs(st - om.state[..., :12]).max() <= 1e-9 * np.abs(om.state[..., :12]).max()
    kap = [el._stateVarsRef[:, 12] for el in rm.elements.values() if el._stateVarsRef.shape[1] > 12]
    assert np.abs(np.array(kap) - om.state[[k for k, el in enumerate(rm.elements.values())
                                            if el._stateVarsRef.shape[1] > 12], :, 12]).max() < 1e-12
    assert rm.time == pytest.approx(2.0)
    # the direct-solver name chose the device's PCG; the string option was converted
    assert solver.device.options["linsolver"] == "pcg" and solver.device.options["residualFirst"] is False


def test_dropin_geostatic_step_keeps_the_prestress_for_the_next_step(oracle_device):
    """A geostatic step (prestress in equilibrium with gravity, displacements reset at its end
    by the reference's own action) followed by a loading step: the state of the first is the
    start of the second, as the reference's elements keep theirs."""
    from edelweissfe_amd.solvers import NISTGPUDropIn
    from edelweissfe_amd.steps import AdaptiveStep

    GEO = dict(p1=0.0, h1=1.0, p2=-10.0, h2=0.0, xLateral=0.5, zLateral=0.7)
    slope = (GEO["p2"] - GEO["p1"]) / (GEO["h2"] - GEO["h1"])
    m = _mesh((2, 2, 2), seed=4)
    rm, om, pos = _models(m, "C3D8", False)
    ns = rm.nodeSets
    rollers = lambda: {"bottom": rs.Dirichlet(ns["gen_bottom"], {1: 0.0}),  # noqa: E731
                       "left": rs.Dirichlet(ns["gen_left"], {0: 0.0}), "right": rs.Dirichlet(ns["gen_right"], {0: 0.0}),
                       "back": rs.Dirichlet(ns["gen_back"], {2: 0.0}), "front": rs.Dirichlet(ns["gen_front"], {2: 0.0})}
    grav = rs.BodyForce(rm.elementSets["all"], (0.0, -slope, 0.0))
    grav._amplitude = lambda t: 1.0
    solver = NISTGPUDropIn(None, None)
    kw = dict(maxInc=1.0, minInc=1e-6, maxNumInc=100, maxIter=15)
    a1 = {"dirichlet": rollers(), "bodyforce": {"g": grav},
          "geostatic": {"geo": rs.Geostatic(rm.elementSets["all"], **GEO)}}
    solver.solveStep(AdaptiveStep(1, rm, a1, **kw), rm, None, [])
    assert np.abs(rm.nodeFields["displacement"]["U"]).max() == 0.0
    a2 = {"dirichlet": dict(rollers(), top=rs.Dirichlet(ns["gen_top"], {1: -2e-4})), "bodyforce": {"g": grav}}
    solver.solveStep(AdaptiveStep(2, rm, a2, maxInc=0.5, minInc=1e-6, maxNumInc=100, maxIter=15), rm, None, [])

    so = onist.NIST()
    rollers_o = [dict(nset=s, values=v) for s, v in (("gen_bottom", {1: 0.0}), ("gen_left", {0: 0.0}),
                                                     ("gen_right", {0: 0.0}), ("gen_back", {2: 0.0}),
                                                     ("gen_front", {2: 0.0}))]
    st1 = onist.Step(dirichlets=rollers_o, bodyforces=[dict(elset="all", vector=(0.0, 0.0, 0.0),
                                                            start=(0.0, -slope, 0.0))], geostatic=[dict(GEO)], **kw)
    so.solve_step(om, st1)
    om.U[:] = 0.0
    st2 = onist.Step(dirichlets=rollers_o + [dict(nset="gen_top", values={1: -2e-4})],
                     bodyforces=[dict(elset="all", vector=(0.0, 0.0, 0.0), start=(0.0, -slope, 0.0))],
                     maxInc=0.5, minInc=1e-6, maxNumInc=100, maxIter=15)
    so.solve_step(om, st2)
    U = rm.nodeFields["displacement"]["U"].ravel()
    assert np.abs(U - om.U).max() <= 1e-9 * np.abs(om.U).max()
    st = np.array([el._stateVarsRef for el in rm.elements.values()])
    assert np.abs(st[..., 0:6] - om.state[..., 0:6]).max() <= 1e-9 * np.abs(om.state[..., 0:6]).max()
    assert np.abs(st[..., 1]).min() > 0.0  # the prestress is still there


def test_element_types_and_materials_of_reference_objects():
    from edelweissfe_amd.solvers.dropin import element_type_of, material_of

    for t in ("C3D8", "C3D20", "C3D20R"):
        assert element_type_of(rs.DisplacementElement(t, 1)) == t
    name, p = material_of(rs.VonMisesMaterial(J2))
    assert name == "vonmises" and np.array_equal(p, J2)
    name, p = material_of(rs.LinearElasticMaterial(LE))
    assert name == "linearelastic" and np.array_equal(p, LE)
    with pytest.raises(NotImplementedError):
        material_of(object())

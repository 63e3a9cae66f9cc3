"""Device element pass and body force on distorted elements against the reference element's
own outputs (tests/golden/element_fixtures.npz, written by make_element_fixtures.py from
DisplacementElement, element.py:194-371), and body forces on jittered meshes and element
subsets against the oracle's NIST history."""

import os

import numpy as np
import pytest

from helpers import device_model, device_step, oracle_model, oracle_step, rel_err

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "element_fixtures.npz")


def _disconnected_system(X, el_type, material, props):
    """The fixture's elements as one mesh with no shared nodes (element k owns nodes k*nn ..)."""
    from edelweissfe_amd.device import DeviceSystem

    n_el, nn = X.shape[0], X.shape[1]
    conn = np.arange(n_el * nn).reshape(n_el, nn)
    return DeviceSystem(X.reshape(-1, 3), conn, el_type, material, props)


@pytest.mark.parametrize("el_type", ["C3D20", "C3D20R"])
def test_element_pass_equals_reference_element_on_distorted_hexa20(el_type):
    d = np.load(FIX)
    X = d[el_type + "_X"]
    n_el = X.shape[0]
    sysm = _disconnected_system(X, el_type, "vonmises", d["vonmises_props"])
    els = np.arange(n_el)
    for inc in (1, 2):
        dU = torch.as_tensor(d[el_type + "_dU%d" % inc].ravel(), device="cuda")
        P, F = sysm.zeros(), sysm.zeros()
        rc, _ = sysm.compute_elements(dU, P, F)
        assert rc == 0
        Ke, Pe = sysm.element_matrices(els)
        Ke = Ke.cpu().numpy().transpose(0, 2, 1)  # column-major per element -> row-major
        assert rel_err(Ke, d[el_type + "_Ke%d" % inc]) < 1e-12
        assert rel_err(Pe.cpu().numpy(), d[el_type + "_Pe%d" % inc]) < 1e-12
        assert rel_err(P.cpu().numpy(), d[el_type + "_Pe%d" % inc].ravel()) < 1e-12
        st = sysm.qp_state(1)
        ref = d[el_type + "_state%d" % inc]
        for lo, hi in ((0, 6), (6, 12), (12, 13)):
            assert rel_err(st[..., lo:hi], ref[..., lo:hi]) < 1e-12
        sysm.accept_state()
    sysm.close()


@pytest.mark.parametrize("el_type", ["C3D20", "C3D20R"])
def test_matrix_free_operator_on_distorted_hexa20(el_type):
    """The FP32 matrix-free product forms J with the reference's hexa20 Jacobian too."""
    from edelweissfe_amd import _lib

    d = np.load(FIX)
    X = d[el_type + "_X"]
    sysm = _disconnected_system(X, el_type, "vonmises", d["vonmises_props"])
    dU = torch.as_tensor(d[el_type + "_dU1"].ravel(), device="cuda")
    P, F = sysm.zeros(), sysm.zeros()
    assert sysm.compute_elements(dU, P, F, precision=_lib.EFE_FP32_MATFREE)[0] == 0
    x = np.random.RandomState(7).uniform(-1, 1, sysm.n_dof)
    y = sysm.zeros()
    sysm.spmv(torch.as_tensor(x, device="cuda"), y)
    K = d[el_type + "_Ke1"]
    yref = np.einsum("eij,ej->ei", K, x.reshape(K.shape[0], -1)).ravel()
    assert rel_err(y.cpu().numpy(), yref) < 2e-5
    sysm.close()


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20", "C3D20R"])
def test_body_force_equals_reference_element_on_distorted_elements(el_type):
    d = np.load(FIX)
    X = d[el_type + "_X"]
    sysm = _disconnected_system(X, el_type, "linearelastic", [210000.0, 0.3])
    P = sysm.zeros()
    sysm.body_force(d["body_load"], P)
    assert rel_err(P.cpu().numpy(), d[el_type + "_bodyforce"].ravel()) < 1e-13
    # an element subset: the others receive nothing; a repeated element counts once
    P2 = sysm.zeros()
    sysm.body_force(d["body_load"], P2, elements=[2, 0, 2])
    ref = d[el_type + "_bodyforce"].copy()
    ref[1] = 0.0
    assert rel_err(P2.cpu().numpy(), ref.ravel()) < 1e-13
    sysm.close()


@pytest.mark.parametrize("case", [
    ((3, 3, 3), "C3D8", "all"),
    ((3, 3, 3), "C3D8", "gen_top"),
    ((2, 2, 2), "C3D20R", "all"),
    ((2, 3, 2), "C3D20R", "gen_left"),
])
def test_body_force_history_on_jittered_mesh_equals_oracle(case):
    """Gravity-like body force on a jittered mesh (all elements, or an element set), bottom
    fixed, J2: the device NIST history equals the oracle's (Newton counts, U, stresses)."""
    from edelweissfe_amd.solvers import NISTGPU
    from oracle import nist as onist

    n, et, elset = case
    J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
    bcs = [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0})]
    bodyforces = [(elset, (40.0, -330.0, 30.0))]
    kw = dict(maxInc=0.5, minInc=1e-6, maxNumInc=50, maxIter=15)
    om = oracle_model(n, (1.0, 1.0, 1.0), et, "vonmises", J2, jitter=0.15, seed=5)
    so = onist.NIST()
    so.solve_step(om, oracle_step(bcs, bodyforces=bodyforces, **kw))
    pm = device_model(n, (1.0, 1.0, 1.0), et, "vonmises", J2, jitter=0.15, seed=5)
    sd = NISTGPU()
    sd.solveStep(device_step(pm, bcs, bodyforces=bodyforces, **kw), pm)
    assert [h[2] for h in so.history] == [r["iters"] for r in sd.history]
    assert np.abs(om.U).max() > 1e-5
    assert rel_err(pm.U, om.U) < 1e-9
    assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 1e-9

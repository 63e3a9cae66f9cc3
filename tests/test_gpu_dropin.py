"""The drop-in solver on the GPU: the reference's own model and step objects (stand-ins,
tests/reference_standins.py) solved by NISTGPUDropIn against the oracle's NIST run; models with
several sections (materials per element set) against the oracle with the same sections; and
the linear-solver plugin's contract f(csr_matrix, ndarray) -> ndarray against SciPy's direct
solve on the matrices NIST hands over (applyDirichletK, nonlinearimplicitstatic.py:559-593)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import reference_standins as rs
from helpers import device_model, device_step, oracle_step, rel_err, uniaxial_bcs
from oracle import assembly as oasm
from oracle import materials as omat
from oracle import mesh as omesh
from oracle import nist as onist
from test_dropin import J2, LE, _actions, _models, _mesh, _oracle_step

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]


@pytest.mark.parametrize("sections,precision", [(False, "fp64"), (True, "fp64"), (True, "mixed")])
def test_dropin_on_reference_objects_matches_oracle(sections, precision):
    from edelweissfe_amd.solvers import NISTGPUDropIn
    from edelweissfe_amd.steps import AdaptiveStep

    m = _mesh((5, 5, 5), seed=2)
    rm, om, pos = _models(m, "C3D8", sections)
    kw = dict(maxInc=0.2, minInc=1e-6, maxNumInc=100, maxIter=15)
    acts = _actions(rm, 0.004)
    acts["options"]["NISTSolver"].options["precision"] = precision
    acts["bodyforce"] = {"bf": rs.BodyForce(rm.elementSets["gen_top"], (30.0, 0.0, 0.0))}
    solver = NISTGPUDropIn(None, None)
    solver.solveStep(AdaptiveStep(1, rm, acts, **kw), rm, None, [])
    so = onist.NIST()
    so.solve_step(om, _oracle_step(0.004, dict(bodyforces=[dict(elset="gen_top", vector=(30.0, 0.0, 0.0))]), **kw))
    U = rm.nodeFields["displacement"]["U"].ravel()
    st = np.array([el._stateVarsRef[:, :12] for el in rm.elements.values()])
    assert (om.state[..., 12] > 0).any()
    tol = 1e-9 if precision == "fp64" else 1e-8
    assert rel_err(U, om.U) < tol
    assert rel_err(st[..., 0:6], om.state[..., 0:6]) < 10 * tol
    if precision == "fp64":
        assert [r["iters"] for r in solver.device.history] == [h[2] for h in so.history]


@pytest.mark.parametrize("el_type,mats,linsolver", [
    ("C3D8", ("vonmises", "linearelastic"), "pcg"),
    ("C3D20R", ("vonmises", "druckerprager"), "gmres"),
    ("C3D8", ("druckerprager", "linearelastic"), "gmres"),
])
def test_sections_history_matches_oracle(el_type, mats, linsolver):
    """Two sections (lower and upper half of the box) with different materials in one device
    system: one element launch per material, one assembled operator."""
    from edelweissfe_amd.solvers import NISTGPU

    props = {"vonmises": J2, "linearelastic": LE, "druckerprager": DP}
    n = (3, 3, 3) if el_type == "C3D8" else (2, 2, 2)
    strain = 0.004 if "vonmises" in mats else 0.003
    pm = device_model(n, (1.0, 1.0, 1.0), el_type, mats[0], props[mats[0]], jitter=0.1, seed=3)
    yc = pm.coordinates[pm.connectivity][:, :, 1].mean(axis=1)
    upper = np.flatnonzero(yc > 0.5)
    pm.elementSets["upper"] = upper
    pm.assignMaterial(mats[1], props[mats[1]], elementSet="upper")
    om_ = omesh.box_mesh(*n, 1.0, 1.0, 1.0, nodes_per_element=8 if el_type == "C3D8" else 20, jitter=0.1, seed=3)
    of = np.zeros(pm.nElements, dtype=int)
    of[upper] = 1
    nq = 8
    sec = omat.Sections([omat.MATERIALS[mats[0]](props[mats[0]]), omat.MATERIALS[mats[1]](props[mats[1]])], of, nq)
    om = onist.Model(om_["coords"], el_type, om_["conn"], sec, om_["nodesets"], om_["elsets"])
    kw = dict(maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)
    so = onist.NIST(convention="colmajor")
    so.solve_step(om, oracle_step(uniaxial_bcs(strain), **kw))
    s = NISTGPU(precision="fp64", linsolver=linsolver)
    s.solveStep(device_step(pm, uniaxial_bcs(strain), **kw), pm)
    assert (om.state[of == 0][..., 12] > 0).any() or (om.state[of == 1][..., 12] > 0).any()
    assert [r["iters"] for r in s.history] == [h[2] for h in so.history]
    assert rel_err(pm.U, om.U) < 1e-9
    assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 1e-9
    assert rel_err(pm.qpState[..., 12], om.state[..., 12]) < 1e-8


def test_sections_matrix_free_operator_matches_assembled():
    """Sections of J2 and Drucker-Prager share the 11-value compact tangent layout of the
    matrix-free operator (cn1 = c1n = 0 for J2): the FP32 product equals the FP64 operator."""
    from edelweissfe_amd import _lib
    from edelweissfe_amd.device import DeviceSystem

    pm = device_model((3, 3, 3), (1.0, 1.0, 1.0), "C3D8", "vonmises", J2, jitter=0.1, seed=2)
    names = np.where(np.arange(pm.nElements) % 2 == 0, "vonmises", "druckerprager")
    props = np.array([J2 if nm == "vonmises" else [30000.0, 0.2, 30.0, 30.0, 10.0, 0.0] for nm in names])
    sysm = DeviceSystem(pm.coordinates, pm.connectivity, "C3D8", names, props)
    rng = np.random.default_rng(5)
    dU = torch.as_tensor(rng.uniform(-1, 1, sysm.n_dof) * 4e-3, device="cuda")
    x = torch.as_tensor(rng.standard_normal(sysm.n_dof), device="cuda")
    P, F = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, P, F)
    y64 = sysm.zeros()
    sysm.spmv(x, y64)
    assert (sysm.qp_state(1)[..., 12] > 0).any()
    Pm, Fm = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, Pm, Fm, precision=_lib.EFE_FP32_MATFREE)
    ym = sysm.zeros()
    sysm.spmv(x, ym)
    assert torch.equal(Pm, P)
    assert rel_err(ym.cpu().numpy(), y64.cpu().numpy()) < 3e-6
    sysm.close()


def _reference_system(el_type, material, props, n, seed, dirichlet=True):
    """K after applyDirichletK (row replacement, columns kept, eliminate_zeros) and a
    right-hand side, as NIST hands them to its linear solver, from the oracle element pass."""
    from helpers import oracle_model

    om = oracle_model(n, (1.0, 1.0, 1.0), el_type, material, props, jitter=0.1, seed=seed)
    so = onist.NIST(convention="colmajor")
    so._setup(om)
    rng = np.random.default_rng(seed)
    dU = rng.uniform(-1, 1, om.nDof) * 3e-3
    V, P, F, st = so.compute_elements(om, dU)
    K = oasm.csr_matrix(oasm.update_csr(V, so.x, so.nnz), so.indptr, so.indices)
    dirs = np.unique(np.r_[3 * om.nodesets["gen_bottom"][:, None] + np.arange(3)[None, :]].ravel())
    if dirichlet:
        K = so.apply_dirichlet_K(K, dirs)
    b = rng.standard_normal(om.nDof)
    if dirichlet:
        b[dirs] = rng.uniform(-1e-3, 1e-3, dirs.size)
    return sp.csr_matrix(K), b


@pytest.mark.parametrize("case", [
    ("C3D8", "vonmises", J2, (5, 4, 4), "pcg"),
    ("C3D20R", "linearelastic", [210000.0, 0.3], (2, 2, 3), "pcg"),
    ("C3D8", "druckerprager", DP, (4, 3, 3), "bicgstab"),
    # (C3D20R DP at this state is singular, cond ~4e20: no solver, direct ones included, solves it)
    ("C3D20", "druckerprager", DP, (2, 2, 2), "gmres"),
    ("C3D8", "druckerprager", DP, (4, 3, 3), "gmres"),
])
def test_linsolve_plugin_on_reference_csr_matches_direct_solve(case):
    from edelweissfe_amd.config.linsolve import getLinSolverByName

    el_type, material, props, n, name = case
    K, b = _reference_system(el_type, material, props, n, 7)
    x_ref = spla.spsolve(K.tocsc(), b, use_umfpack=False)
    solve = getLinSolverByName(name, {"rtol": 1e-13})
    x = solve(K, b)
    assert isinstance(x, np.ndarray) and x.shape == b.shape
    assert np.linalg.norm(K @ x - b) <= 1e-11 * np.linalg.norm(b)
    assert rel_err(x, x_ref) < 1e-8
    # a right-hand side of two columns (the reference's arc-length systems pass [D, 2])
    B = np.stack([b, 2.0 * b + 1.0], axis=1)
    X = solve(K, B)
    assert X.shape == B.shape
    assert rel_err(X[:, 1], spla.spsolve(K.tocsc(), B[:, 1], use_umfpack=False)) < 1e-8


def test_csr_solve_on_device_arrays_matches_host_entry_point():
    from edelweissfe_amd import _lib
    from edelweissfe_amd.config.linsolve import csr_solve
    import ctypes

    K, b = _reference_system("C3D8", "vonmises", J2, (4, 4, 4), 3)
    x_host, infos = csr_solve(K, b, rtol=1e-12)
    lib = _lib.load()
    dev = torch.device("cuda")
    ip = torch.as_tensor(K.indptr.astype(np.int32), device=dev)
    ix = torch.as_tensor(K.indices.astype(np.int32), device=dev)
    v = torch.as_tensor(K.data, device=dev)
    tb = torch.as_tensor(b, device=dev)
    tx = torch.empty_like(tb)
    o = _lib.SolveOpts(solver=_lib.EFE_SOLVER_CG, precond=0, precision=0, max_iter=50000, restart=0, check_every=32,
                       rtol=1e-12, atol=0.0)
    info = _lib.SolveInfo()
    torch.cuda.synchronize()
    _lib.check(lib.efe_csr_linsolve_device(K.shape[0], ip.data_ptr(), ix.data_ptr(), v.data_ptr(), tb.data_ptr(),
                                           tx.data_ptr(), None, ctypes.byref(o), ctypes.byref(info)),
               "efe_csr_linsolve_device")
    assert info.iterations == infos[0].iterations
    assert np.array_equal(tx.cpu().numpy(), x_host)
    # an out-of-range column index is refused before any kernel walks it
    bad = ix.clone()
    bad[5] = K.shape[0] + 3
    rc = lib.efe_csr_linsolve_device(K.shape[0], ip.data_ptr(), bad.data_ptr(), v.data_ptr(), tb.data_ptr(),
                                     tx.data_ptr(), None, ctypes.byref(o), ctypes.byref(info))
    assert rc == _lib.EFE_ERR_ARG

"""GPU parity of the device path against the oracle (runs through the C ABI)."""

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from helpers import NPE, device_model, device_step, oracle_model, oracle_step, rel_err, uniaxial_bcs
from oracle import assembly as oasm

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
J2NL = [210000.0, 0.15, 355.0, 1000.0, 200.0, 1400.0]
DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]


def _system(pm):
    from edelweissfe_amd.device import DeviceSystem

    return DeviceSystem(pm.coordinates, pm.connectivity, pm.elementType, pm.material["name"],
                        pm.material["properties"])


def _oracle_pattern(om):
    I, J = oasm.vij_pattern(om.elements.dof_map)
    indptr, indices = oasm.csr_pattern(I, J, om.nDof)
    return I, J, indptr, indices


@pytest.mark.parametrize("case", [
    ((20, 4, 4), (10.0, 2.0, 2.0), "C3D8", 0.0),
    ((4, 4, 4), (1.0, 1.0, 1.0), "C3D8", 0.1),
    ((2, 2, 2), (1.0, 1.0, 1.0), "C3D20R", 0.0),
    ((3, 2, 2), (1.0, 2.0, 1.5), "C3D20", 0.0),
])
def test_csr_pattern_bit_exact(case):
    n, l, et, jit = case
    om = oracle_model(n, l, et, "linearelastic", [210000.0, 0.3], jitter=jit)
    pm = device_model(n, l, et, "linearelastic", [210000.0, 0.3], jitter=jit)
    assert np.array_equal(om.coords, pm.coordinates)
    sysm = _system(pm)
    _, _, indptr, indices = _oracle_pattern(om)
    gi, gj = sysm.csr_pattern()
    assert gi.dtype == np.int32 and gj.dtype == np.int32
    assert np.array_equal(gi, indptr)
    assert np.array_equal(gj, indices)
    sysm.close()


def _element_parity(n, l, et, mat, props, scale, jitter, seed, transpose_convention):
    om = oracle_model(n, l, et, mat, props, jitter=jitter, seed=seed)
    pm = device_model(n, l, et, mat, props, jitter=jitter, seed=seed)
    sysm = _system(pm)
    rs = np.random.RandomState(seed)
    dU = scale * rs.uniform(-1.0, 1.0, om.nDof)
    # oracle element pass
    Ke, Pe, st = om.elements.compute(dU[om.elements.dof_map], om.state, om.material)
    I, J, indptr, indices = _oracle_pattern(om)
    x = oasm.x_map(I, J, indptr, indices)
    V = oasm.vij_values_reference(Ke) if transpose_convention else oasm.vij_values_colmajor(Ke)
    Kref = oasm.update_csr(V, x, indices.size)
    Pref = np.zeros(om.nDof)
    Fref = np.zeros(om.nDof)
    np.add.at(Pref, om.elements.dof_map.ravel(), Pe.ravel())
    np.add.at(Fref, om.elements.dof_map.ravel(), np.abs(Pe).ravel())
    # device pass
    dev = torch.device("cuda")
    tdU = torch.as_tensor(dU, device=dev)
    P = sysm.zeros()
    F = sysm.zeros()
    rc, _ = sysm.compute_elements(tdU, P, F)
    assert rc == 0
    Kg = sysm.export_csr_values().cpu().numpy()
    if transpose_convention:
        # the reference assembles Ke^T (element.py:318 with dofmanager.py:551-555); the device
        # assembles Ke.  Compare against the transpose of the oracle matrix.
        Kref = sp.csr_matrix((Kref, indices, indptr), shape=(om.nDof,) * 2).T.tocsr()
        Kref.sort_indices()
        Kref = Kref.data
    stg = sysm.qp_state(1)
    sysm.close()
    return Kg, Kref, P.cpu().numpy(), Pref, F.cpu().numpy(), Fref, stg, st


@pytest.mark.parametrize("case", [
    ((4, 4, 4), "C3D8", "linearelastic", [210000.0, 0.3], 1e-3),
    ((4, 3, 5), "C3D8", "vonmises", J2, 3e-3),
    ((3, 3, 3), "C3D8", "vonmises", J2NL, 3e-3),
    ((2, 2, 2), "C3D20R", "vonmises", J2, 3e-3),
    ((2, 2, 2), "C3D20", "vonmises", J2, 3e-3),
])
def test_element_assembly_parity_fp64(case):
    n, et, mat, props, scale = case
    Kg, Kref, P, Pref, F, Fref, stg, st = _element_parity(n, (1.0, 1.0, 1.0), et, mat, props, scale, 0.1, 3, True)
    # element R/K to 1e-12 relative (per field, relative to its largest entry)
    assert rel_err(Kg, Kref) < 1e-12
    assert rel_err(P, Pref) < 1e-12
    assert rel_err(F, Fref) < 1e-12
    assert rel_err(stg[..., 0:6], st[..., 0:6]) < 1e-12
    assert rel_err(stg[..., 6:12], st[..., 6:12]) < 1e-12
    if st.shape[-1] > 12:
        assert rel_err(stg[..., 12], st[..., 12]) < 1e-12
        if mat == "vonmises":
            assert (st[..., 12] > 0).any(), "case must be plastic somewhere"


def test_element_assembly_parity_drucker_prager():
    # non-symmetric tangent: column-major (native element ABI) convention on both sides
    Kg, Kref, P, Pref, F, Fref, stg, st = _element_parity((2, 2, 2), (1.0, 1.0, 1.0), "C3D20R", "druckerprager", DP,
                                                          2e-3, 0.0, 5, False)
    assert (st[..., 12] > 0).any()
    assert rel_err(Kg, Kref) < 1e-12
    assert rel_err(P, Pref) < 1e-12
    assert rel_err(stg[..., 0:6], st[..., 0:6]) < 1e-12


def test_element_assembly_fp32_contractions():
    from edelweissfe_amd import _lib

    n, et = (4, 4, 4), "C3D8"
    om = oracle_model(n, (1, 1, 1), et, "vonmises", J2, jitter=0.1, seed=3)
    pm = device_model(n, (1, 1, 1), et, "vonmises", J2, jitter=0.1, seed=3)
    sysm = _system(pm)
    dU = 3e-3 * np.random.RandomState(3).uniform(-1, 1, om.nDof)
    Ke, Pe, st = om.elements.compute(dU[om.elements.dof_map], om.state, om.material)
    I, J, indptr, indices = _oracle_pattern(om)
    Kref = oasm.update_csr(oasm.vij_values_colmajor(Ke), oasm.x_map(I, J, indptr, indices), indices.size)
    Pref = np.zeros(om.nDof)
    np.add.at(Pref, om.elements.dof_map.ravel(), Pe.ravel())
    P = sysm.zeros()
    F = sysm.zeros()
    sysm.compute_elements(torch.as_tensor(dU, device="cuda"), P, F, precision=_lib.EFE_FP32)
    Kg = sysm.export_csr_values().cpu().numpy()
    sysm.close()
    assert rel_err(Kg, Kref) < 2e-6          # FP32 contractions
    assert rel_err(P.cpu().numpy(), Pref) < 1e-12  # residual stays FP64


def test_pcg_matches_direct_solve():
    from edelweissfe_amd import _lib

    n = (6, 5, 4)
    om = oracle_model(n, (1, 1, 1), "C3D8", "linearelastic", [210000.0, 0.3], jitter=0.1)
    pm = device_model(n, (1, 1, 1), "C3D8", "linearelastic", [210000.0, 0.3], jitter=0.1)
    sysm = _system(pm)
    bcs = uniaxial_bcs(0.01)
    dirs = np.concatenate([pm.dofsOfNodeSet(s, sorted(v)) for _, s, v in bcs])
    sysm.set_dirichlet(dirs)
    delta = np.zeros(om.nDof)
    for _, s, v in bcs:
        for c, val in v.items():
            delta[pm.dofsOfNodeSet(s, [c])] = val
    dev = torch.device("cuda")
    P, F, corr = sysm.zeros(), sysm.zeros(), sysm.zeros()
    dirv = torch.as_tensor(delta, device=dev)
    sysm.compute_elements(sysm.zeros(), P, F, dir_values=dirv, dir_correction=corr)
    R = P.clone()
    R[torch.as_tensor(dirs, device=dev)] = dirv[torch.as_tensor(dirs, device=dev)]
    b = R - corr
    x = sysm.zeros()
    info = sysm.linsolve(b, x, precision=_lib.EFE_FP64, rtol=1e-12)
    assert info.status == 0 and info.iterations > 0
    # reference: row-replacement Dirichlet and SuperLU (nonlinearimplicitstatic.py:577-588, linsolve.py:47-50)
    Ke, Pe, _ = om.elements.compute(np.zeros((om.elements.conn.shape[0], 24)), om.state, om.material)
    I, J, indptr, indices = _oracle_pattern(om)
    K = oasm.csr_matrix(oasm.update_csr(oasm.vij_values_reference(Ke), oasm.x_map(I, J, indptr, indices),
                                        indices.size), indptr, indices)
    for row in dirs:
        K.data[K.indptr[row]:K.indptr[row + 1]] = 0.0
    d = K.diagonal()
    d[dirs] = 1.0
    K.setdiag(d)
    K.eliminate_zeros()
    Rr = np.zeros(om.nDof)
    Rr[dirs] = delta[dirs]
    xr = spla.spsolve(K, Rr, use_umfpack=False)
    xg = x.cpu().numpy()
    sysm.close()
    assert rel_err(xg, xr) < 1e-9
    assert np.array_equal(xg[dirs], delta[dirs])


def _run_device(pm, step_kw, bcs, precision="fp64", nodeforces=(), bodyforces=(), extrapolation="linear", **opts):
    from edelweissfe_amd.solvers import NISTGPU

    st = device_step(pm, bcs, nodeforces, bodyforces, extrapolation=extrapolation, **step_kw)
    s = NISTGPU(precision=precision, **opts)
    s.solveStep(st, pm)
    return s


@pytest.mark.parametrize("name,n,l,et,E,nu,bcs,nfs,bfs", [
    ("TensionBarHexa8", (4, 4, 30), (5, 5, 500), "C3D8", 1.8e4, 0.22,
     [("front", "gen_front", {0: 0.0, 1: 0.0, 2: 0.0})], [("gen_back", (0, 0, -4.0))], []),
    ("WallShearHexa8", (20, 20, 2), (1000, 1000, 5), "C3D8", 2.1e4, 0.22,
     [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0})], [("gen_top", (1.0, 0, 0))],
     [("all", (0.0, -0.000077, 0.0))]),
    ("WallShearHexa20", (10, 10, 1), (1000, 1000, 5), "C3D20", 2.1e4, 0.22,
     [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0})], [("gen_top", (1.18868, 0, 0))],
     [("all", (0.0, -0.000077, 0.0))]),
])
def test_nist_reference_golden_vectors(golden, name, n, l, et, E, nu, bcs, nfs, bfs):
    pm = device_model(n, l, et, "linearelastic", [E, nu])
    s = _run_device(pm, dict(maxInc=1.0, minInc=1e-9, maxNumInc=1000, maxIter=25), bcs, nodeforces=nfs,
                    bodyforces=bfs, extrapolation="off")
    ref = golden["uref_" + name]
    print(name, s.history)
    # the reference regression criterion (_run_tests_edelweissfe.py:101-106)
    assert np.abs(pm.U - ref).max() < 1e-6
    assert rel_err(pm.U, ref) < 1e-9
    assert [r["iters"] for r in s.history] == [1, 2]


@pytest.mark.parametrize("precond", ["ilu0", "bjacobi"])
def test_nist_reference_golden_vector_gmres(golden, precond):
    """testfiles/WallShearHexa8GMRES: the 10 x 10 x 1 wall of C3D8N (the C3D8 element,
    elements/library.py:541) that the reference solves with GMRES (linsolver=gmres), here with
    the device GMRES(50) and block ILU(0) or block Jacobi, FP64."""
    pm = device_model((10, 10, 1), (1000, 1000, 5), "C3D8", "linearelastic", [2.1e4, 0.22])
    s = _run_device(pm, dict(maxInc=1.0, minInc=1e-9, maxNumInc=1000, maxIter=25),
                    [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0})], nodeforces=[("gen_top", (1.0, 0, 0))],
                    bodyforces=[("all", (0.0, -0.000077, 0.0))], extrapolation="off", linsolver="gmres",
                    preconditioner=precond)
    ref = golden["uref_WallShearHexa8GMRES"]
    print(precond, s.history)
    assert np.abs(pm.U - ref).max() < 1e-6
    assert rel_err(pm.U, ref) < 1e-8  # U.ref holds the reference's GMRES error (2e-9 relative)
    assert [r["iters"] for r in s.history] == [1, 2]


def test_nist_j2_load_history_matches_oracle():
    n, l = (5, 5, 5), (1.0, 1.0, 1.0)
    om = oracle_model(n, l, "C3D8", "vonmises", J2, jitter=0.1, seed=0)
    pm = device_model(n, l, "C3D8", "vonmises", J2, jitter=0.1, seed=0)
    bcs = uniaxial_bcs(0.01)
    kw = dict(maxInc=0.1, minInc=1e-6, maxNumInc=100, maxIter=15)
    so = __import__("oracle.nist", fromlist=["NIST"]).NIST()
    so.solve_step(om, oracle_step(bcs, **kw))
    sd = _run_device(pm, kw, bcs)
    assert [h[2] for h in so.history] == [r["iters"] for r in sd.history]
    assert rel_err(pm.U, om.U) < 1e-9
    assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 1e-9
    assert (om.state[..., 12] > 0).any()


@pytest.mark.parametrize("operator", ["matfree", "assembled", "auto"])
def test_mixed_precision_error_vs_fp64(operator):
    n, l = (5, 5, 5), (1.0, 1.0, 1.0)
    bcs = uniaxial_bcs(0.01)
    kw = dict(maxInc=0.1, minInc=1e-6, maxNumInc=100, maxIter=15)
    p64 = device_model(n, l, "C3D8", "vonmises", J2, jitter=0.1)
    p32 = device_model(n, l, "C3D8", "vonmises", J2, jitter=0.1)
    _run_device(p64, kw, bcs)
    _run_device(p32, kw, bcs, precision="mixed", mixedOperator=operator)
    assert rel_err(p32.U, p64.U) < 1e-8
    assert rel_err(p32.qpState[..., 0:6], p64.qpState[..., 0:6]) < 1e-7


def test_device_path_is_bitwise_reproducible():
    n, l = (4, 4, 4), (1.0, 1.0, 1.0)
    bcs = uniaxial_bcs(0.01)
    kw = dict(maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)
    a = device_model(n, l, "C3D8", "vonmises", J2, jitter=0.1)
    b = device_model(n, l, "C3D8", "vonmises", J2, jitter=0.1)
    _run_device(a, kw, bcs)
    _run_device(b, kw, bcs)
    assert np.array_equal(a.U, b.U)
    assert np.array_equal(a.qpState, b.qpState)


def _dirichlet_system(om, pm, sysm, bcs, dU, precision):
    """Assemble at dU with the Dirichlet set of `bcs`, iteration-0 right-hand side; returns
    (b, oracle K with row replacement, oracle rhs, Dirichlet DOFs, prescribed values)."""
    dirs = np.concatenate([pm.dofsOfNodeSet(s, sorted(v)) for _, s, v in bcs])
    sysm.set_dirichlet(dirs)
    delta = np.zeros(om.nDof)
    for _, s, v in bcs:
        for c, val in v.items():
            delta[pm.dofsOfNodeSet(s, [c])] = val
    dev = torch.device("cuda")
    P, F, corr = sysm.zeros(), sysm.zeros(), sysm.zeros()
    dirv = torch.as_tensor(delta, device=dev)
    sysm.compute_elements(torch.as_tensor(dU, device=dev), P, F, dir_values=dirv, dir_correction=corr,
                          precision=precision)
    idx = torch.as_tensor(dirs, device=dev)
    R = P.clone()
    R[idx] = dirv[idx]
    b = R - corr
    Ke, Pe, _ = om.elements.compute(dU[om.elements.dof_map], om.state, om.material)
    I, J, indptr, indices = _oracle_pattern(om)
    K = oasm.csr_matrix(oasm.update_csr(oasm.vij_values_colmajor(Ke), oasm.x_map(I, J, indptr, indices),
                                        indices.size), indptr, indices)
    for row in dirs:
        K.data[K.indptr[row]:K.indptr[row + 1]] = 0.0
    d = K.diagonal()
    d[dirs] = 1.0
    K.setdiag(d)
    K.eliminate_zeros()
    Rr = np.zeros(om.nDof)
    np.add.at(Rr, om.elements.dof_map.ravel(), Pe.ravel())
    Rr[dirs] = delta[dirs]
    return b, K, Rr, dirs, delta


@pytest.mark.parametrize("precond", ["ilu0", "bjacobi"])
def test_gmres_matches_direct_solve_on_nonsymmetric_tangent(precond):
    from edelweissfe_amd import _lib

    n = (3, 3, 3)
    om = oracle_model(n, (1, 1, 1), "C3D20R", "druckerprager", DP)
    pm = device_model(n, (1, 1, 1), "C3D20R", "druckerprager", DP)
    sysm = _system(pm)
    # a plastic state: load first, then assemble the non-symmetric consistent tangent there
    dU = np.zeros(om.nDof)
    coords = om.coords
    dU[1::3] = -2e-3 * coords[:, 1]
    dU[0::3] = 1e-3 * coords[:, 1]
    bcs = [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0}), ("top", "gen_top", {1: -1e-3})]
    b, K, Rr, dirs, delta = _dirichlet_system(om, pm, sysm, bcs, dU, _lib.EFE_FP64)
    x = sysm.zeros()
    pc = _lib.EFE_PRECOND_ILU0 if precond == "ilu0" else _lib.EFE_PRECOND_BLOCK_JACOBI
    info = sysm.linsolve(b, x, solver=_lib.EFE_SOLVER_GMRES, precond=pc, precision=_lib.EFE_FP64, rtol=1e-12,
                         max_iter=5000, restart=50)
    assert info.status == 0, (info.iterations, info.residual)
    xr = spla.spsolve(K, Rr, use_umfpack=False)
    xg = x.cpu().numpy()
    assert abs(K - K.T).max() > 1e-6 * abs(K).max(), "tangent must be non-symmetric"
    assert rel_err(xg, xr) < 1e-9
    # FP32 GMRES on the FP32 operator reaches its own tolerance
    b32, _, _, _, _ = _dirichlet_system(om, pm, sysm, bcs, dU, _lib.EFE_FP32)
    x32 = sysm.zeros()
    info32 = sysm.linsolve(b32, x32, solver=_lib.EFE_SOLVER_GMRES, precond=pc, precision=_lib.EFE_FP32, rtol=1e-5,
                           max_iter=5000, restart=50)
    sysm.close()
    assert info32.status == 0
    assert rel_err(x32.cpu().numpy(), xr) < 1e-3  # rtol 1e-5 in FP32 times the conditioning


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nist_drucker_prager_hexa20r_matches_oracle(precision):
    from oracle import nist as onist

    n, l = (3, 3, 3), (1.0, 1.0, 1.0)
    om = oracle_model(n, l, "C3D20R", "druckerprager", DP)
    pm = device_model(n, l, "C3D20R", "druckerprager", DP)
    bcs = [("bottom", "gen_bottom", {0: 0.0, 1: 0.0, 2: 0.0}), ("top", "gen_top", {1: -4e-3, 0: 1e-3})]
    kw = dict(maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)
    so = onist.NIST(convention="colmajor")
    so.solve_step(om, oracle_step(bcs, **kw))
    sd = _run_device(pm, kw, bcs, precision=precision, linsolver="gmres")
    assert (om.state[..., 12] > 0).any()
    if precision == "fp64":
        assert [h[2] for h in so.history] == [r["iters"] for r in sd.history]
        assert rel_err(pm.U, om.U) < 1e-9
        assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 1e-8
    else:
        assert rel_err(pm.U, om.U) < 1e-8


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nist_j2_per_element_lognormal_yield_matches_oracle(precision):
    """Per-element material properties (sigma_y lognormal per element, the 96^3 workload's
    material field at 5^3) through the device path against the oracle with the same
    properties repeated per QP."""
    from edelweissfe_amd.utils import rng

    n, l = (5, 5, 5), (1.0, 1.0, 1.0)
    E = n[0] * n[1] * n[2]
    props = np.tile(np.asarray(J2), (E, 1))
    props[:, 2] = rng.lognormal(1, E, 250.0, 0.1)
    om = oracle_model(n, l, "C3D8", "vonmises", np.repeat(props, 8, axis=0), jitter=0.1, seed=0)
    pm = device_model(n, l, "C3D8", "vonmises", props, jitter=0.1, seed=0)
    bcs = uniaxial_bcs(0.02)
    kw = dict(maxInc=1.0 / 6, minInc=1e-6, maxNumInc=100, maxIter=15)
    so = __import__("oracle.nist", fromlist=["NIST"]).NIST()
    so.solve_step(om, oracle_step(bcs, **kw))
    sd = _run_device(pm, kw, bcs, precision=precision)
    # the yield stresses spread, and so do the plastic strains they leave
    kap = om.state[..., 12]
    assert np.ptp(props[:, 2]) > 50.0 and kap.min() < 0.9 * kap.max()
    if precision == "fp64":
        assert [h[2] for h in so.history] == [r["iters"] for r in sd.history]
        tol = 1e-9
    else:
        tol = 1e-8
    assert rel_err(pm.U, om.U) < tol
    assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 10 * tol
    assert rel_err(pm.qpState[..., 12], kap) < 100 * tol


GEO = dict(p1=0.0, h1=1.0, p2=-10.0, h2=0.0, xLateral=0.5, zLateral=0.7)
ROLLERS = [("bottom", "gen_bottom", {1: 0.0}), ("left", "gen_left", {0: 0.0}), ("right", "gen_right", {0: 0.0}),
           ("back", "gen_back", {2: 0.0}), ("front", "gen_front", {2: 0.0})]


@pytest.mark.parametrize("el_type", ["C3D8", "C3D20R", "C3D20"])
def test_geostatic_initial_state_matches_oracle(el_type):
    from oracle import nist as onist

    n, l = (3, 2, 2), (1.0, 2.0, 1.0)
    om = oracle_model(n, l, el_type, "druckerprager", DP, jitter=0.1, seed=5)
    pm = device_model(n, l, el_type, "druckerprager", DP, jitter=0.1, seed=5)
    sysm = _system(pm)
    sub = np.array([0, 2, 5, 7])
    om.elsets["sub"] = sub
    onist.NIST.apply_geostatic(om, dict(GEO, h1=2.0, elset="sub"))
    sysm.set_geostatic([GEO["p1"], 2.0, GEO["p2"], GEO["h2"], GEO["xLateral"], GEO["zLateral"]], sub)
    got = sysm.qp_state(0)
    assert np.abs(got[sub, :, 0:3]).min() > 0
    np.testing.assert_allclose(got[..., 0:6], om.state[..., 0:6], rtol=1e-14, atol=1e-12)
    onist.NIST.apply_geostatic(om, dict(GEO))
    sysm.set_geostatic([GEO[k] for k in ("p1", "h1", "p2", "h2", "xLateral", "zLateral")])
    np.testing.assert_allclose(sysm.qp_state(0)[..., 0:6], om.state[..., 0:6], rtol=1e-14, atol=1e-12)


def _footing_nodes(coords, top):
    x, z = coords[top, 0], coords[top, 2]
    return top[(np.abs(x - 0.5) <= 0.25 + 1e-12) & (np.abs(z - 0.5) <= 0.25 + 1e-12)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_geostatic_then_footing_drucker_prager_matches_oracle(precision):
    """The non-symmetric workload's two steps at 3^3: geostatic prestress in equilibrium with
    gravity (displacements reset afterwards), then a rigid footing pressed into the top-face
    centre quarter; GMRES + ILU(0) on the device, SuperLU in the oracle."""
    from edelweissfe_amd.solvers import NISTGPU
    from edelweissfe_amd.steps import AdaptiveStep, BodyForce, Dirichlet, Geostatic
    from oracle import nist as onist

    n, l = (3, 3, 3), (1.0, 1.0, 1.0)
    slope = (GEO["p2"] - GEO["p1"]) / (GEO["h2"] - GEO["h1"])
    om = oracle_model(n, l, "C3D20R", "druckerprager", DP, jitter=0.1, seed=3)
    pm = device_model(n, l, "C3D20R", "druckerprager", DP, jitter=0.1, seed=3)
    foot = _footing_nodes(pm.coordinates, pm.nodeSets["gen_top"])
    assert foot.size == 8
    om.nodesets["footing"] = foot
    pm.nodeSets["footing"] = foot
    footing = [("footing", "footing", {1: -4e-3})]
    kw1 = dict(maxInc=1.0, minInc=1e-6, maxNumInc=100, maxIter=15)
    kw2 = dict(maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)

    so = onist.NIST(convention="colmajor")
    st1 = oracle_step(ROLLERS, bodyforces=[("all", (0.0, 0.0, 0.0))], **kw1)
    st1.bodyforces[0]["start"] = (0.0, -slope, 0.0)
    st1.geostatic = [dict(GEO)]
    so.solve_step(om, st1)
    st2 = oracle_step(ROLLERS + footing, bodyforces=[("all", (0.0, 0.0, 0.0))], **kw2)
    st2.bodyforces[0]["start"] = (0.0, -slope, 0.0)
    so.solve_step(om, st2)

    grav = BodyForce("gravity", "all", (0.0, -slope, 0.0), amplitude=lambda t: 1.0)
    s = NISTGPU(precision=precision, linsolver="gmres")
    a1 = {"dirichlet": {nm: Dirichlet(nm, ns, v) for nm, ns, v in ROLLERS}, "bodyforce": {"gravity": grav},
          "geostatic": {"geo": Geostatic("geo", **GEO)}}
    s.solveStep(AdaptiveStep(1, pm, a1, **kw1), pm)
    assert np.abs(pm.U).max() == 0.0
    a2 = {"dirichlet": {nm: Dirichlet(nm, ns, v) for nm, ns, v in ROLLERS + footing},
          "bodyforce": {"gravity": grav}}
    s.solveStep(AdaptiveStep(2, pm, a2, **kw2), pm)

    assert (om.state[..., 12] > 0).any()
    print([h[2] for h in so.history], [r["iters"] for r in s.history])
    if precision == "fp64":
        assert [h[2] for h in so.history] == [r["iters"] for r in s.history]
        assert rel_err(pm.U, om.U) < 1e-9
        assert rel_err(pm.qpState[..., 0:6], om.state[..., 0:6]) < 1e-8
    else:
        assert rel_err(pm.U, om.U) < 1e-8


@pytest.mark.parametrize("el_type,material,props,n", [
    ("C3D8", "vonmises", J2, (4, 3, 3)),
    ("C3D8", "druckerprager", DP, (3, 3, 2)),
    ("C3D20R", "druckerprager", DP, (2, 3, 2)),
    ("C3D20", "linearelastic", [210000.0, 0.3], (2, 2, 2)),
    ("C3D20R", "vonmises", J2NL, (2, 2, 3)),
])
def test_matfree_operator_matches_assembled(el_type, material, props, n):
    """The FP32 matrix-free operator (per-QP J^-1 and compact tangent) against the FP64
    assembled, Dirichlet-eliminated operator at a plastic state; P and F from the FP64
    residual path are identical; the right-hand-side correction K[:, dir] dir_values too."""
    from edelweissfe_amd import _lib

    l = (1.0, 1.0, 1.0)
    pm = device_model(n, l, el_type, material, props, jitter=0.1, seed=2)
    sysm = _system(pm)
    dirs = np.r_[pm.dofsOfNodeSet("gen_bottom"), pm.dofsOfNodeSet("gen_top", [1])]
    sysm.set_dirichlet(dirs)
    rng = np.random.default_rng(11)
    dU = torch.as_tensor(rng.uniform(-1, 1, sysm.n_dof) * 4e-3, device="cuda")
    dv = torch.zeros(sysm.n_dof, dtype=torch.float64, device="cuda")
    dv[torch.as_tensor(np.unique(dirs), device="cuda")] = torch.as_tensor(
        rng.uniform(-1, 1, np.unique(dirs).size) * 1e-3, device="cuda")
    x = torch.as_tensor(rng.standard_normal(sysm.n_dof), device="cuda")
    P64, F64, c64 = sysm.zeros(), sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, P64, F64, dir_values=dv, dir_correction=c64)
    y64 = sysm.zeros()
    sysm.spmv(x, y64)
    if material != "linearelastic":
        assert (sysm.qp_state(1)[..., 12] > 0).any()
    Pm, Fm, cm = sysm.zeros(), sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, Pm, Fm, dir_values=dv, dir_correction=cm, precision=_lib.EFE_FP32_MATFREE)
    ym = sysm.zeros()
    sysm.spmv(x, ym)
    assert torch.equal(Pm, P64) and torch.equal(Fm, F64)
    assert rel_err(ym.cpu().numpy(), y64.cpu().numpy()) < 3e-6
    assert rel_err(cm.cpu().numpy(), c64.cpu().numpy()) < 3e-6
    # Dirichlet rows are identity rows of the FP32 operator
    d = torch.as_tensor(np.unique(dirs), device="cuda")
    assert torch.equal(ym[d], x[d].float().double())


def test_matfree_operator_on_a_mesh_without_element_locality():
    """Elements in random order: 256 consecutive elements touch more nodes than a workgroup
    brick stages, so the hexa8 matrix-free operator keeps one force slot per (element, node)
    (mf_apply_hex8_kernel and the slot gather).  Against the FP64 assembled operator of the same
    context at a plastic state, as test_matfree_operator_matches_assembled does."""
    from edelweissfe_amd import _lib
    from edelweissfe_amd.device import DeviceSystem

    pm = device_model((12, 12, 12), (1.0, 1.0, 1.0), "C3D8", "vonmises", J2, jitter=0.1, seed=2)
    perm = np.random.default_rng(4).permutation(pm.connectivity.shape[0])
    sysm = DeviceSystem(pm.coordinates, np.ascontiguousarray(pm.connectivity[perm]), "C3D8", "vonmises",
                        pm.material["properties"])
    dirs = np.r_[pm.dofsOfNodeSet("gen_bottom"), pm.dofsOfNodeSet("gen_top", [1])]
    sysm.set_dirichlet(dirs)
    rng = np.random.default_rng(11)
    dU = torch.as_tensor(rng.uniform(-1, 1, sysm.n_dof) * 4e-3, device="cuda")
    x = torch.as_tensor(rng.standard_normal(sysm.n_dof), device="cuda")
    P64, F64 = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, P64, F64)
    assert (sysm.qp_state(1)[..., 12] > 0).any()
    y64 = sysm.zeros()
    sysm.spmv(x, y64)
    Pm, Fm = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, Pm, Fm, precision=_lib.EFE_FP32_MATFREE)
    ym = sysm.zeros()
    sysm.spmv(x, ym)
    b = Pm.clone()
    b[torch.as_tensor(dirs, device="cuda")] = 0.0
    xm = sysm.zeros()
    info = sysm.linsolve(b, xm, precision=_lib.EFE_FP32_MATFREE, rtol=1e-6)
    sysm.close()
    assert torch.equal(Pm, P64) and torch.equal(Fm, F64)
    assert rel_err(ym.cpu().numpy(), y64.cpu().numpy()) < 3e-6
    assert info.status == 0 and info.iterations > 0


def test_matfree_solve_after_an_assembled_fp32_pass_keeps_its_preconditioner():
    """The matrix-free PCG is replayed from a captured graph that holds the address of the FP32
    block-Jacobi inverse; a first assembled FP32 pass must not move that buffer.  A matrix-free
    solve at a plastic state, after matrix-free and assembled FP32 passes and solves at another
    state (the mixed operator cost probe's sequence), equals bit for bit the same solve on a
    fresh context."""
    from edelweissfe_amd import _lib

    pm = device_model((4, 4, 3), (1.0, 1.0, 1.0), "C3D8", "vonmises", J2, jitter=0.1, seed=2)
    dirs = np.r_[pm.dofsOfNodeSet("gen_bottom"), pm.dofsOfNodeSet("gen_top", [1])]
    rng = np.random.default_rng(3)
    dU = torch.as_tensor(rng.uniform(-1, 1, 3 * pm.nNodes) * 4e-3, device="cuda")
    b0 = torch.as_tensor(rng.standard_normal(3 * pm.nNodes), device="cuda")

    def solve(probe):
        sysm = _system(pm)
        sysm.set_dirichlet(dirs)
        P, F, x = sysm.zeros(), sysm.zeros(), sysm.zeros()
        if probe:
            for prec in (_lib.EFE_FP32_MATFREE, _lib.EFE_FP32):
                sysm.compute_elements(sysm.zeros(), P, F, precision=prec)
                sysm.linsolve(b0, x, precision=prec, rtol=1e-30, max_iter=64, check_every=32)
        sysm.compute_elements(dU, P, F, precision=_lib.EFE_FP32_MATFREE)
        assert (sysm.qp_state(1)[..., 12] > 0).any()
        b = P.clone()
        b[torch.as_tensor(dirs, device="cuda")] = 0.0
        x = sysm.zeros()
        info = sysm.linsolve(b, x, precision=_lib.EFE_FP32_MATFREE, rtol=1e-6, check_every=32)
        sysm.close()
        assert info.status == 0
        return info.iterations, x

    its_a, xa = solve(False)
    its_b, xb = solve(True)
    assert its_a == its_b
    assert torch.equal(xa, xb)


@pytest.mark.parametrize("case", [
    ((4, 3, 5), "C3D8", "vonmises", J2, 3e-3),
    ((2, 2, 2), "C3D20R", "druckerprager", None, 3e-3),
])
def test_element_matrices_in_native_element_layout(case):
    """efe_get_element_matrices: Ke column-major and Pe per element, as
    MarmotElement::computeYourself returns them, equal to the oracle element pass."""
    n, et, mat, props, scale = case
    if props is None:
        props = DP
    om = oracle_model(n, (1.0, 1.0, 1.0), et, mat, props, jitter=0.1, seed=4)
    pm = device_model(n, (1.0, 1.0, 1.0), et, mat, props, jitter=0.1, seed=4)
    sysm = _system(pm)
    dU = scale * np.random.RandomState(4).uniform(-1.0, 1.0, om.nDof)
    Ke, Pe, _ = om.elements.compute(dU[om.elements.dof_map], om.state, om.material)
    P, F = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(torch.as_tensor(dU, device="cuda"), P, F)
    els = np.array([0, om.elements.conn.shape[0] - 1, 3, 1])
    Kd, Pd = (t.cpu().numpy() for t in sysm.element_matrices(els))
    sysm.close()
    for k, e in enumerate(els):
        # column-major: Kd[k][col][row]
        assert rel_err(Kd[k].T, Ke[e]) < 1e-12
        assert rel_err(Pd[k], Pe[e]) < 1e-12

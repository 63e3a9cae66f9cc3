"""efe_csr_linsolve_device_noalloc: the CSR solve with its workspace in the library's static
device scratch instead of an allocation, against efe_csr_linsolve_device (the same kernels
with an allocated workspace) through the C ABI: the same solution bit for bit and the same
iteration counts on two calls in a row, the second with a smaller system, for CG and
BiCGStab; a system beyond the scratch is refused with an error before any device work."""

import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import oracle_model
from oracle import assembly as oasm
from oracle import nist as onist

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]
DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]


def _system(n, material, props, seed):
    om = oracle_model(n, (1.0, 1.0, 1.0), "C3D8", material, props, jitter=0.1, seed=seed)
    so = onist.NIST(convention="colmajor")
    so._setup(om)
    rng = np.random.default_rng(seed)
    V, _, _, _ = so.compute_elements(om, rng.uniform(-1, 1, om.nDof) * 3e-3)
    K = oasm.csr_matrix(oasm.update_csr(V, so.x, so.nnz), so.indptr, so.indices)
    dirs = np.unique((3 * om.nodesets["gen_bottom"][:, None] + np.arange(3)[None, :]).ravel())
    K = sp.csr_matrix(so.apply_dirichlet_K(K, dirs))
    b = rng.standard_normal(om.nDof)
    b[dirs] = rng.uniform(-1e-3, 1e-3, dirs.size)
    return K, b


def _call(fn, K, b, solver):
    from edelweissfe_amd import _lib

    dev = torch.device("cuda")
    ip = torch.as_tensor(K.indptr.astype(np.int32), device=dev)
    ix = torch.as_tensor(K.indices.astype(np.int32), device=dev)
    v = torch.as_tensor(K.data, device=dev)
    tb = torch.as_tensor(b, device=dev)
    tx = torch.full_like(tb, np.nan)
    o = _lib.SolveOpts(solver=solver, precond=0, precision=0, max_iter=20000, restart=0, check_every=16, rtol=1e-12,
                       atol=0.0)
    info = _lib.SolveInfo()
    torch.cuda.synchronize()
    rc = fn(K.shape[0], ctypes.c_void_p(ip.data_ptr()), ctypes.c_void_p(ix.data_ptr()), ctypes.c_void_p(v.data_ptr()),
            ctypes.c_void_p(tb.data_ptr()), ctypes.c_void_p(tx.data_ptr()), None, ctypes.byref(o), ctypes.byref(info))
    _lib.check(rc, "csr solve")
    return tx.cpu().numpy(), info.iterations, info.status


@pytest.mark.parametrize("material,props,solver_name", [("vonmises", J2, "cg"), ("druckerprager", DP, "bicgstab")])
def test_noalloc_solve_equals_allocating_solve_twice(material, props, solver_name):
    from edelweissfe_amd import _lib

    lib = _lib.load()
    noalloc = lib.efe_csr_linsolve_device_noalloc
    noalloc.restype = ctypes.c_int32
    noalloc.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 6 + [ctypes.POINTER(_lib.SolveOpts),
                                                                    ctypes.POINTER(_lib.SolveInfo)]
    solver = _lib.EFE_SOLVER_CG if solver_name == "cg" else _lib.EFE_SOLVER_BICGSTAB
    for n, seed in (((5, 4, 4), 1), ((3, 3, 2), 2)):  # the second system is smaller
        K, b = _system(n, material, props, seed)
        xa, ita, sta = _call(lib.efe_csr_linsolve_device, K, b, solver)
        xn, itn, stn = _call(noalloc, K, b, solver)
        print("n_dof", K.shape[0], "iterations", ita, itn, flush=True)
        assert sta == stn == _lib.EFE_OK
        assert ita == itn and ita > 0
        assert np.array_equal(xa, xn)
        assert np.linalg.norm(K @ xn - b) <= 1e-10 * np.linalg.norm(b)


def test_noalloc_solve_refuses_a_system_beyond_its_scratch():
    from edelweissfe_amd import _lib

    lib = _lib.load()
    noalloc = lib.efe_csr_linsolve_device_noalloc
    noalloc.restype = ctypes.c_int32
    noalloc.argtypes = [ctypes.c_int32] + [ctypes.c_void_p] * 6 + [ctypes.POINTER(_lib.SolveOpts),
                                                                    ctypes.POINTER(_lib.SolveInfo)]
    o = _lib.SolveOpts(solver=_lib.EFE_SOLVER_CG, max_iter=10, rtol=1e-8)
    info = _lib.SolveInfo()
    one = ctypes.c_void_p(1)  # never dereferenced: the size checks come first
    rc = noalloc((1 << 20) + 1, one, one, one, one, one, None, ctypes.byref(o), ctypes.byref(info))
    assert rc == _lib.EFE_ERR_ARG
    assert b"scratch" in lib.efe_last_error()
    o.solver = _lib.EFE_SOLVER_BICGSTAB  # 9 vectors: beyond the scratch well below 2^20 unknowns
    rc = noalloc(800000, one, one, one, one, one, None, ctypes.byref(o), ctypes.byref(info))
    assert rc == _lib.EFE_ERR_ARG
    assert b"scratch" in lib.efe_last_error()
    o.solver = _lib.EFE_SOLVER_GMRES
    rc = noalloc(100, one, one, one, one, one, None, ctypes.byref(o), ctypes.byref(info))
    assert rc == _lib.EFE_ERR_ARG

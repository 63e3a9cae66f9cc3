"""Linear-solver registry with the reference's contract (edelweissfe/config/linsolve.py:46-86):
``getLinSolverByName(name, opts)`` returns ``f(A, b) -> x``.

* ``A`` a SciPy CSR matrix (int32 indices) and ``b`` a NumPy array [D] or [D, k], as
  NIST.linearSolve passes them (nonlinearimplicitstatic.py:727-751): the matrix after
  applyDirichletK, whose constrained rows are unit rows.  The system goes to the GPU
  (efe_csr_linsolve): rows that are unit rows are prescribed, the Krylov method solves the free
  block.  ``x`` has the shape of ``b``; a NaN in x (also the answer of a solve that did not
  converge) is left to the caller, which raises DivergingSolution (nonlinearimplicitstatic.py:748-749).
* ``A`` a DeviceSystem (its assembled, Dirichlet-eliminated operator) and ``b`` a CUDA float64
  tensor: the context's own block solvers.

Names: "pcg" (also "cg", "pcg-fp64"): Jacobi CG for symmetric K (on a DeviceSystem: block
Jacobi); "pcg-fp32": on a DeviceSystem, the FP32 solve; "gmres": restarted GMRES(m) for
non-symmetric K (Jacobi on a CSR matrix, ILU(0) on a DeviceSystem); "bicgstab": BiCGStab with
Jacobi on a CSR matrix (GMRES on a DeviceSystem).
opts: rtol (default 1e-12 on CSR matrices, which stand in for the reference's direct solvers;
1e-10 / 1e-5 on a DeviceSystem), atol (0), maxIter (50000), restart (GMRES, 50)."""

import ctypes

import numpy as np

from .. import _lib
from ..exceptions import DivergingSolution

_CG = ("pcg", "cg", "pcg-fp64", "pcg-fp32")
_NONSYM = ("bicgstab", "gmres")


def csr_solve(A, b, solver=_lib.EFE_SOLVER_CG, rtol=1e-12, atol=0.0, max_iter=50000, check_every=32, restart=50):
    """x with A x = b for a SciPy CSR matrix and a host vector or matrix b, on the GPU."""
    import scipy.sparse as sp

    lib = _lib.load()
    A = sp.csr_matrix(A)
    n = A.shape[0]
    if A.shape != (n, n):
        raise ValueError("A must be square")
    indptr = np.ascontiguousarray(A.indptr, dtype=np.int32)
    indices = np.ascontiguousarray(A.indices, dtype=np.int32)
    data = np.ascontiguousarray(A.data, dtype=np.float64)
    B = np.asarray(b, dtype=np.float64)
    cols = B.reshape(n, -1)
    X = np.empty_like(cols)
    infos = []
    for k in range(cols.shape[1]):
        bk = np.ascontiguousarray(cols[:, k])
        xk = np.empty(n)
        o = _lib.SolveOpts(solver=solver, precond=_lib.EFE_PRECOND_BLOCK_JACOBI, precision=_lib.EFE_FP64,
                           max_iter=int(max_iter), restart=int(restart), check_every=int(check_every),
                           rtol=float(rtol), atol=float(atol))
        info = _lib.SolveInfo()
        _lib.check(lib.efe_csr_linsolve(n, indptr.ctypes.data, indices.ctypes.data, data.ctypes.data, bk.ctypes.data,
                                        xk.ctypes.data, ctypes.byref(o), ctypes.byref(info)), "efe_csr_linsolve")
        # a solve that did not converge returns NaN: the caller raises DivergingSolution for it
        # (nonlinearimplicitstatic.py:748-749), as for a failed direct factorisation
        X[:, k] = xk if info.status == _lib.EFE_OK else np.nan
        infos.append(info)
    return X.reshape(B.shape), infos


def getLinSolverByName(linsolverName, opts=None):
    opts = dict(opts or {})
    name = linsolverName.lower()
    if name not in _CG + _NONSYM:
        raise AttributeError("invalid linear solver {:} requested".format(linsolverName))
    max_iter = int(opts.get("maxIter", 50000))

    def solve(A, b):
        if hasattr(A, "linsolve"):  # a DeviceSystem and a CUDA tensor
            import torch

            if name in _NONSYM:
                solver, precond, precision = _lib.EFE_SOLVER_GMRES, _lib.EFE_PRECOND_ILU0, _lib.EFE_FP64
            else:
                solver, precond = _lib.EFE_SOLVER_CG, _lib.EFE_PRECOND_BLOCK_JACOBI
                precision = _lib.EFE_FP32 if name == "pcg-fp32" else _lib.EFE_FP64
            rtol = float(opts.get("rtol", 1e-10 if precision == _lib.EFE_FP64 else 1e-5))
            x = torch.empty_like(b)
            info = A.linsolve(b, x, solver=solver, precond=precond, precision=precision, rtol=rtol,
                              atol=float(opts.get("atol", 0.0)), max_iter=max_iter,
                              restart=int(opts.get("restart", 50)))
            if info.status == _lib.EFE_NAN:
                raise DivergingSolution("Obtained NaN in linear solve")
            return x
        solver = {"bicgstab": _lib.EFE_SOLVER_BICGSTAB, "gmres": _lib.EFE_SOLVER_GMRES}.get(name, _lib.EFE_SOLVER_CG)
        x, _ = csr_solve(A, b, solver=solver, rtol=float(opts.get("rtol", 1e-12)), atol=float(opts.get("atol", 0.0)),
                         max_iter=max_iter, restart=int(opts.get("restart", 50)))
        return x

    return solve

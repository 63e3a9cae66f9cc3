"""Linear-solver registry (edelweissfe/config/linsolve.py:46-86 semantics):
``getLinSolverByName(name, opts)`` returns ``f(A, b) -> x``.  The device
solvers act on the system a DeviceSystem has assembled, so ``A`` is that
DeviceSystem and ``b`` a CUDA float64 tensor."""

import torch

from .. import _lib
from ..exceptions import DivergingSolution


def getLinSolverByName(linsolverName, opts=None):
    opts = dict(opts or {})
    name = linsolverName.lower()
    if name in ("pcg", "pcg-fp64", "pcg-fp32"):
        precision = _lib.EFE_FP32 if name == "pcg-fp32" else _lib.EFE_FP64
        rtol = float(opts.get("rtol", 1e-10 if precision == _lib.EFE_FP64 else 1e-5))

        def solve(A, b):
            x = torch.empty_like(b)
            info = A.linsolve(b, x, solver=_lib.EFE_SOLVER_CG, precond=_lib.EFE_PRECOND_BLOCK_JACOBI,
                              precision=precision, rtol=rtol, max_iter=int(opts.get("maxIter", 50000)))
            if info.status == _lib.EFE_NAN:
                raise DivergingSolution("Obtained NaN in linear solve")
            return x

        return solve
    raise AttributeError("invalid linear solver {:} requested".format(linsolverName))

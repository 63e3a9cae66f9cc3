"""ctypes binding of oracle/lib/liboracle_nist.so (the C restatement; test infrastructure)."""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "lib", "liboracle_nist.so")
_lib = None

_P = ctypes.c_void_p


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle C library missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB)
        lib.oracle_elements_hexa8.restype = ctypes.c_int
        lib.oracle_elements_hexa8.argtypes = [ctypes.c_int, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, _P]
        lib.oracle_csr_pattern.restype = ctypes.c_int64
        lib.oracle_csr_pattern.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _P, _P, ctypes.POINTER(_P)]
        lib.oracle_x_map.restype = None
        lib.oracle_x_map.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P]
        lib.oracle_update_csr.restype = None
        lib.oracle_update_csr.argtypes = [ctypes.c_int64, _P, _P, _P, ctypes.c_int64]
        lib.oracle_ilu0.restype = ctypes.c_int
        lib.oracle_ilu0.argtypes = [ctypes.c_int, _P, _P, _P, _P, ctypes.c_double]
        lib.oracle_gmres_phase_times.restype = None
        lib.oracle_gmres_phase_times.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int, _P]
        lib.oracle_gather_pf.restype = None
        lib.oracle_gather_pf.argtypes = [ctypes.c_int, ctypes.c_int, _P, _P, _P, _P, ctypes.c_int]
        lib.oracle_pcg.restype = None
        lib.oracle_pcg.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                   ctypes.c_double, _P]
        lib.oracle_free.restype = None
        lib.oracle_free.argtypes = [_P]
        lib.oracle_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data


def elements_hexa8(coords, conn, props, material, dU, state_ref):
    """material: 0 linear elastic, 1 von Mises.  state AoS [E][8][ns].  Returns V (E*576, row-major
    blocks as the reference writes them), Pe (E, 24), new state, number of failed local Newtons."""
    lib = load()
    E = conn.shape[0]
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    props = np.ascontiguousarray(props, dtype=np.float64)
    dU = np.ascontiguousarray(dU, dtype=np.float64)
    sref = np.ascontiguousarray(state_ref, dtype=np.float64)
    stmp = np.empty_like(sref)
    V = np.empty(E * 576)
    Pe = np.empty((E, 24))
    fails = lib.oracle_elements_hexa8(E, _p(coords), _p(conn), _p(props), 1 if props.ndim == 2 else 0, material,
                                      _p(dU), _p(sref), _p(stmp), _p(V), _p(Pe))
    return V, Pe, stmp, fails


def csr_pattern(nDof, dofmap):
    lib = load()
    dofmap = np.ascontiguousarray(dofmap, dtype=np.int32)
    E, n = dofmap.shape
    indptr = np.empty(nDof + 1, dtype=np.int32)
    ptr = _P()
    nnz = lib.oracle_csr_pattern(nDof, E, n, _p(dofmap), _p(indptr), ctypes.byref(ptr))
    buf = (ctypes.c_int32 * nnz).from_address(ptr.value)
    indices = np.frombuffer(buf, dtype=np.int32).copy()
    lib.oracle_free(ptr)
    return indptr, indices


def x_map(dofmap, indptr, indices):
    lib = load()
    dofmap = np.ascontiguousarray(dofmap, dtype=np.int32)
    E, n = dofmap.shape
    x = np.empty(E * n * n, dtype=np.int32)
    lib.oracle_x_map(E, n, _p(dofmap), _p(indptr), _p(indices), _p(x))
    return x


def update_csr(V, x, nnz):
    lib = load()
    data = np.empty(nnz)
    lib.oracle_update_csr(V.size, _p(x), _p(V), _p(data), nnz)
    return data


def gather_pf(dofmap, Pe, nDof):
    lib = load()
    dofmap = np.ascontiguousarray(dofmap, dtype=np.int32)
    P = np.empty(nDof)
    F = np.empty(nDof)
    lib.oracle_gather_pf(dofmap.shape[0], dofmap.shape[1], _p(dofmap), _p(np.ascontiguousarray(Pe)), _p(P), _p(F),
                         nDof)
    return P, F


def pcg(indptr, indices, data, b, rtol, atol, max_iter, budget_s):
    lib = load()
    x = np.empty_like(b)
    out = np.zeros(3)
    lib.oracle_pcg(b.size, _p(indptr), _p(indices), _p(data), _p(b), _p(x), rtol, atol, max_iter, budget_s, _p(out))
    return x, int(out[0]), out[1], out[2]


def threads():
    return load().oracle_threads()


def ilu0(indptr, indices, data, budget_s):
    """Scalar ILU(0) of the leading rows within budget_s: returns (lu, diag, rows done)."""
    lib = load()
    lu = np.array(data, dtype=np.float64)
    diag = np.empty(indptr.size - 1, dtype=np.int32)
    rows = lib.oracle_ilu0(indptr.size - 1, _p(indptr), _p(indices), _p(lu), _p(diag), budget_s)
    return lu, diag, rows


def gmres_phase_times(indptr, indices, data, lu, diag, j, reps):
    lib = load()
    out = np.zeros(4)
    lib.oracle_gmres_phase_times(indptr.size - 1, _p(indptr), _p(indices), _p(data), _p(lu), _p(diag), j, reps,
                                 _p(out))
    return {"spmv": out[0], "ilu_apply": out[1], "gram_schmidt": out[2], "normalize": out[3]}

"""The C ABI's argument checks (include/efe_gpu.h): every rejected call returns its error
code and message before anything reaches the GPU, and leaves the context usable.  The
reference raises Python exceptions at the same places (a DOF outside the DOF vector, an
element outside the element list, an unknown face or quadrature point, a solver that does
not fit the operator); the device path reports them through efe_last_error."""

import ctypes

import numpy as np
import pytest

from helpers import device_model

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def _err(lib):
    return lib.efe_last_error().decode()


def _desc(lib_mod, coords, conn, et="C3D8", material="linearelastic", props=(1.0, 0.3), spec=None, nspec=1):
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    conn = np.ascontiguousarray(conn, dtype=np.int32)
    props = np.ascontiguousarray(props, dtype=np.float64)
    keep = [coords, conn, props]
    if spec is not None:
        spec = np.ascontiguousarray(spec, dtype=np.int32)
        keep.append(spec)
    d = lib_mod.MeshDesc(n_nodes=coords.shape[0], n_elements=conn.shape[0], element_type=lib_mod.ELEMENT_TYPES[et],
                         material=lib_mod.MATERIALS[material], coords=coords.ctypes.data, conn=conn.ctypes.data,
                         n_props=props.size, props_per_element=0, props=props.ctypes.data,
                         specimen_of_node=spec.ctypes.data if spec is not None else None, n_specimens=nspec,
                         stream=torch.cuda.current_stream().cuda_stream)
    return d, keep


def test_context_creation_rejects_bad_meshes():
    from edelweissfe_amd import _lib

    lib = _lib.load()
    pm = device_model((2, 2, 2), (1.0, 1.0, 1.0), "C3D8", "linearelastic", [1.0, 0.3])
    X, conn = pm.coordinates, pm.connectivity
    ctx = ctypes.c_void_p()

    def create(d):
        return lib.efe_ctx_create(ctypes.byref(d), ctypes.byref(ctx))

    d, keep = _desc(_lib, X, conn)
    d.n_nodes = 0
    assert create(d) == _lib.EFE_ERR_ARG and "invalid mesh descriptor" in _err(lib)
    d, keep = _desc(_lib, X, conn)
    d.element_type = 99
    assert create(d) == _lib.EFE_ERR_ARG and "unknown element type" in _err(lib)
    d, keep = _desc(_lib, X, conn, material="vonmises", props=(1.0, 0.3))
    assert create(d) == _lib.EFE_ERR_ARG and "n_props" in _err(lib)
    bad = conn.copy()
    bad[3, 5] = X.shape[0]
    d, keep = _desc(_lib, X, bad)
    assert create(d) == _lib.EFE_ERR_ARG and "outside [0, n_nodes)" in _err(lib)
    bad[3, 5] = -1
    d, keep = _desc(_lib, X, bad)
    assert create(d) == _lib.EFE_ERR_ARG and "outside [0, n_nodes)" in _err(lib)
    spec = np.zeros(X.shape[0], dtype=np.int32)
    spec[::2] = 1  # not contiguous node ranges
    d, keep = _desc(_lib, X, conn, spec=spec, nspec=2)
    assert create(d) == _lib.EFE_ERR_ARG and "contiguous" in _err(lib)
    # a good descriptor still works after the failures
    d, keep = _desc(_lib, X, conn)
    assert create(d) == _lib.EFE_OK
    lib.efe_ctx_destroy(ctx)


def test_entry_points_reject_bad_arguments_and_keep_the_context_usable():
    from edelweissfe_amd import _lib
    from edelweissfe_amd.device import DeviceSystem

    pm = device_model((2, 2, 2), (1.0, 1.0, 1.0), "C3D8", "vonmises", [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0])
    sysm = DeviceSystem(pm.coordinates, pm.connectivity, "C3D8", "vonmises", [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0])
    lib, ctx = sysm.lib, sysm.ctx
    D, E = sysm.n_dof, sysm.n_elements
    dU, P, F = sysm.zeros(), sysm.zeros(), sysm.zeros()
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    i32 = lambda *v: (ctypes.c_int32 * len(v))(*v)  # noqa: E731

    # solve before any element pass
    b, x = sysm.zeros(), sysm.zeros()
    o = _lib.SolveOpts(solver=_lib.EFE_SOLVER_CG, precond=_lib.EFE_PRECOND_BLOCK_JACOBI, precision=_lib.EFE_FP64,
                       max_iter=10, restart=0, check_every=4, rtol=1e-8, atol=0.0)
    info = _lib.SolveInfo()
    assert lib.efe_linsolve(ctx, ptr(b), ptr(x), ctypes.byref(o), ctypes.byref(info)) != _lib.EFE_OK
    assert "call efe_compute_elements first" in _err(lib)
    # Dirichlet DOFs outside the DOF vector
    assert lib.efe_set_dirichlet(ctx, i32(0, D), 2) == _lib.EFE_ERR_ARG and "out of range" in _err(lib)
    assert lib.efe_set_dirichlet(ctx, i32(-1), 1) == _lib.EFE_ERR_ARG
    # element pass: unknown precision, NULL vector, prescribed values without their correction
    t = (ctypes.c_double * 2)(0.0, 0.0)
    pnew = ctypes.c_double()
    args = lambda d_, prec, dv=None, dc=None: (ctx, None, d_, t, 0.0, ptr(P), ptr(F), dv, dc, prec,  # noqa: E731
                                                 ctypes.byref(pnew))
    assert lib.efe_compute_elements(*args(ptr(dU), 7)) == _lib.EFE_ERR_ARG and "invalid arguments" in _err(lib)
    assert lib.efe_compute_elements(*args(None, _lib.EFE_FP64)) == _lib.EFE_ERR_ARG
    assert lib.efe_compute_elements(*args(ptr(dU), _lib.EFE_FP64, dv=ptr(b))) == _lib.EFE_ERR_ARG
    # surface traction: face outside 1..6, element outside the list
    tr = (ctypes.c_double * 3)(0.0, 0.0, -1.0)
    PExt = sysm.zeros()
    assert lib.efe_surface_traction(ctx, i32(0), i32(7), 1, tr, ptr(PExt)) == _lib.EFE_ERR_ARG
    assert "face not in 1..6" in _err(lib)
    assert lib.efe_surface_traction(ctx, i32(E), i32(1), 1, tr, ptr(PExt)) == _lib.EFE_ERR_ARG
    assert float(PExt.abs().max()) == 0.0
    # geostatic: h1 == h2, element outside the list
    geo = (ctypes.c_double * 6)(-1.0, 0.0, -2.0, 0.0, 0.5, 0.5)
    assert lib.efe_set_geostatic(ctx, geo, None, 0) == _lib.EFE_ERR_ARG
    geo = (ctypes.c_double * 6)(-1.0, 0.0, -2.0, 1.0, 0.5, 0.5)
    assert lib.efe_set_geostatic(ctx, geo, i32(E), 1) == _lib.EFE_ERR_ARG and "out of range" in _err(lib)
    # results: quadrature point and element outside their ranges, components past the state
    out = torch.zeros(64, dtype=torch.float64, device=sysm.device)
    assert lib.efe_collect_results(ctx, 0, i32(0), 1, i32(8), 1, 0, 6, ptr(out)) == _lib.EFE_ERR_ARG
    assert "quadrature point out of range" in _err(lib)
    assert lib.efe_collect_results(ctx, 0, i32(E), 1, i32(0), 1, 0, 6, ptr(out)) == _lib.EFE_ERR_ARG
    assert "element out of range" in _err(lib)
    assert lib.efe_collect_results(ctx, 0, i32(0), 1, i32(0), 1, 10, 6, ptr(out)) == _lib.EFE_ERR_ARG
    # batched entry points on an ordinary context
    u8 = (ctypes.c_uint8 * 1)(1)
    assert lib.efe_get_specimen_status(ctx, u8) != _lib.EFE_OK and "specimens" in _err(lib)
    assert lib.efe_accept_state_batched(ctx, u8) != _lib.EFE_OK
    assert lib.efe_set_active_specimens(ctx, u8) == _lib.EFE_ERR_STATE and "specimens" in _err(lib)
    tol = (ctypes.c_double * 1)(1e-8)
    assert lib.efe_linsolve_batched(ctx, ptr(b), ptr(x), ctypes.byref(o), tol, None, None, None,
                                    ctypes.byref(info)) != _lib.EFE_OK

    # the context still works: an element pass, a solve and the element matrices
    sysm.set_dirichlet(np.arange(3 * 9))
    st, _ = sysm.compute_elements(dU, P, F)
    assert st == _lib.EFE_OK
    # a solve in another precision than K was assembled in, and GMRES restart too long
    o.precision = _lib.EFE_FP32
    assert lib.efe_linsolve(ctx, ptr(b), ptr(x), ctypes.byref(o), ctypes.byref(info)) != _lib.EFE_OK
    assert "precision" in _err(lib)
    o.precision, o.solver, o.restart = _lib.EFE_FP64, _lib.EFE_SOLVER_GMRES, 1000
    assert lib.efe_linsolve(ctx, ptr(b), ptr(x), ctypes.byref(o), ctypes.byref(info)) != _lib.EFE_OK
    assert "restart" in _err(lib)
    b.fill_(1.0)
    info = sysm.linsolve(b, x, rtol=1e-10)
    assert info.status == _lib.EFE_OK and info.iterations > 0
    Ke, Pe = sysm.element_matrices([0, E - 1])
    assert float(Ke.abs().max()) > 0
    with pytest.raises(Exception):
        sysm.element_matrices([E])
    # after a matrix-free pass there are no element matrices to export
    sysm.compute_elements(dU, P, F, precision=_lib.EFE_FP32_MATFREE)
    with pytest.raises(Exception, match="matrix-free"):
        sysm.element_matrices([0])
    sysm.close()


def test_newton_round_rejects_bad_arguments():
    """efe_newton_round_batched: a context without specimens, a missing array, an unknown
    precision and negative limits are refused before any device work; the context stays usable."""
    from edelweissfe_amd import _lib
    from edelweissfe_amd.device import DeviceSystem

    lib = _lib.load()
    pm = device_model((2, 2, 2), (1.0, 1.0, 1.0), "C3D8", "linearelastic", [1.0, 0.3])
    dev = torch.device("cuda")
    S = 2
    spec = np.r_[np.zeros(pm.nNodes, np.int32), np.ones(pm.nNodes, np.int32)]
    coords = np.vstack([pm.coordinates, pm.coordinates + 2.0])
    conn = np.vstack([pm.connectivity, pm.connectivity + pm.nNodes])
    batch = DeviceSystem(coords, conn, "C3D8", "linearelastic", [1.0, 0.3], specimen_of_node=spec, n_specimens=S)
    single = DeviceSystem(pm.coordinates, pm.connectivity, "C3D8", "linearelastic", [1.0, 0.3])
    o = _lib.NewtonOpts(precision=_lib.EFE_FP64, krylov_max_iter=100, check_every=16, alt_after=15, rtol=1e-10,
                        rtol_first=1e-10, atol_factor=1e-2, flux_tol=1e-8, flux_tol_alt=5e-3, correction_tol=1e-8,
                        flux_floor=1e-7, avg_floor=1e-10)
    u8 = np.ones(S, np.uint8)
    i32 = np.full(S, 5, np.int32)
    na = np.ones(S)
    outs = [np.zeros(S, np.int32) for _ in range(3)] + [np.zeros((S, 3), np.int32)]

    def call(sysm, dU, live=u8, maxit=i32, nacc=na, opts=o):
        dT = torch.zeros_like(dU)
        return lib.efe_newton_round_batched(
            sysm.ctx, ctypes.c_void_p(dU.data_ptr()), None, None, None, ctypes.c_void_p(dT.data_ptr()),
            live.ctypes.data if live is not None else None, u8.ctypes.data, u8.ctypes.data, maxit.ctypes.data,
            i32.ctypes.data, nacc.ctypes.data, ctypes.byref(opts), *[a.ctypes.data for a in outs])

    dU1 = torch.zeros(single.n_dof, dtype=torch.float64, device=dev)
    assert call(single, dU1) == _lib.EFE_ERR_STATE and "specimens" in _err(lib)
    dU = torch.zeros(batch.n_dof, dtype=torch.float64, device=dev)
    assert call(batch, dU, live=None) == _lib.EFE_ERR_ARG
    bad = _lib.NewtonOpts(precision=7)
    assert call(batch, dU, opts=bad) == _lib.EFE_ERR_ARG
    assert call(batch, dU, maxit=np.array([5, -1], np.int32)) == _lib.EFE_ERR_ARG and "max_iter" in _err(lib)
    assert call(batch, dU, nacc=np.array([1.0, 0.0])) == _lib.EFE_ERR_ARG
    # still usable: a round with no live specimen takes one pass and reports every specimen idle
    assert call(batch, dU, live=np.zeros(S, np.uint8)) == _lib.EFE_OK
    assert (outs[0] == _lib.EFE_NEWTON_IDLE).all() and (outs[3] == 0).all()
    batch.close()
    single.close()

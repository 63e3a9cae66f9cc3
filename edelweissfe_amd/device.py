"""Device-side system of one homogeneous element set: a thin owner of an efe_ctx.

Vectors are torch float64 tensors on the GPU; they are passed to the library
by address and all work is ordered on torch's current stream.
"""

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check


class DeviceSystem:
    def __init__(self, coords, conn, el_type, material, props, specimen_of_node=None, n_specimens=1, device=None):
        """material: a material name for all elements, or one name per element (sections);
        props: one row, or one row per element (with several materials: 6 values per row,
        linear elastic reading the first 2)."""
        if not torch.cuda.is_available():
            raise RuntimeError("DeviceSystem needs a GPU (torch.cuda.is_available() is False)")
        self.lib = _lib.load()
        self.device = torch.device(device or "cuda")
        coords = np.ascontiguousarray(coords, dtype=np.float64)
        conn = np.ascontiguousarray(conn, dtype=np.int32)
        props = np.ascontiguousarray(props, dtype=np.float64)
        self.el_type = el_type.upper()
        self.n_nodes = coords.shape[0]
        self.n_elements = conn.shape[0]
        if self.el_type not in _lib.ELEMENT_TYPES:
            raise ValueError("element type {} not supported (C3D8, C3D20, C3D20R)".format(el_type))
        mat_of = None
        if isinstance(material, str):
            self.material = material.lower()
            if self.material not in _lib.MATERIALS:
                raise ValueError("material {} not supported".format(material))
        else:  # several sections: a material per element
            names = [str(m).lower() for m in np.asarray(material).ravel()]
            if len(names) != self.n_elements:
                raise ValueError("one material name per element needed")
            bad = sorted(set(names) - set(_lib.MATERIALS))
            if bad:
                raise ValueError("materials {} not supported".format(bad))
            self.material = names[0] if len(set(names)) == 1 else "sections"
            self.materials = names
            mat_of = np.ascontiguousarray([_lib.MATERIALS[m] for m in names], dtype=np.int32)
            full = np.zeros((self.n_elements, 6))
            p2 = np.atleast_2d(props)
            if p2.shape[0] == 1:
                p2 = np.repeat(p2, self.n_elements, axis=0)
            if p2.shape[0] != self.n_elements:
                raise ValueError("per-element properties need one row per element")
            full[:, :min(6, p2.shape[1])] = p2[:, :6]
            props = np.ascontiguousarray(full)
        nn = 8 if self.el_type == "C3D8" else 20
        if conn.shape[1] != nn:
            raise ValueError("connectivity has {} nodes per element, {} needs {}".format(conn.shape[1], el_type, nn))
        if conn.min() < 0 or conn.max() >= self.n_nodes:
            raise ValueError("connectivity references nodes outside [0, n_nodes)")
        per_el = props.ndim == 2
        if per_el and props.shape[0] != self.n_elements:
            raise ValueError("per-element properties need one row per element")
        spec = None
        if specimen_of_node is not None:
            spec = np.ascontiguousarray(specimen_of_node, dtype=np.int32)
        self._keep = (coords, conn, props, spec, mat_of)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        desc = _lib.MeshDesc(
            n_nodes=self.n_nodes, n_elements=self.n_elements, element_type=_lib.ELEMENT_TYPES[self.el_type],
            material=_lib.MATERIALS.get(self.material, 0), coords=coords.ctypes.data, conn=conn.ctypes.data,
            n_props=props.shape[-1], props_per_element=1 if (per_el or mat_of is not None) else 0,
            props=props.ctypes.data, specimen_of_node=spec.ctypes.data if spec is not None else None,
            n_specimens=n_specimens, stream=stream,
            material_of_element=mat_of.ctypes.data if mat_of is not None else None)
        self.ctx = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            check(self.lib.efe_ctx_create(ctypes.byref(desc), ctypes.byref(self.ctx)), "efe_ctx_create")
        sizes = (ctypes.c_int64 * 4)()
        check(self.lib.efe_get_sizes(self.ctx, sizes), "efe_get_sizes")
        self.n_dof, self.nnz, self.n_qp, self.n_state = (int(v) for v in sizes)
        self.n_specimens = int(n_specimens) if specimen_of_node is not None else 1

    # -- helpers
    def _ptr(self, t):
        if t is None:
            return None
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and t.numel() == self.n_dof):
            raise ValueError("expected a contiguous float64 CUDA tensor of length n_dof")
        return ctypes.c_void_p(t.data_ptr())

    def zeros(self):
        return torch.zeros(self.n_dof, dtype=torch.float64, device=self.device)

    def close(self):
        if getattr(self, "ctx", None) is not None and self.ctx.value:
            self.lib.efe_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- entry points
    def csr_pattern(self):
        indptr = np.empty(self.n_dof + 1, dtype=np.int32)
        indices = np.empty(self.nnz, dtype=np.int32)
        check(self.lib.efe_get_csr_pattern(self.ctx, indptr.ctypes.data, indices.ctypes.data), "efe_get_csr_pattern")
        return indptr, indices

    def set_dirichlet(self, dofs):
        dofs = np.ascontiguousarray(np.unique(np.asarray(dofs, dtype=np.int64)).astype(np.int32))
        check(self.lib.efe_set_dirichlet(self.ctx, dofs.ctypes.data, int(dofs.size)), "efe_set_dirichlet")

    def compute_elements(self, dU, P, F, dir_values=None, dir_correction=None, precision=_lib.EFE_FP64, time=(0.0, 0.0),
                         dT=0.0, U=None):
        """Returns the status (EFE_OK or EFE_CUTBACK) and pNewDT."""
        t = (ctypes.c_double * 2)(*time)
        pnew = ctypes.c_double(1.0)
        rc = self.lib.efe_compute_elements(self.ctx, self._ptr(U) if U is not None else None, self._ptr(dU), t,
                                           float(dT), self._ptr(P), self._ptr(F), self._ptr(dir_values),
                                           self._ptr(dir_correction), int(precision), ctypes.byref(pnew))
        check(rc, "efe_compute_elements")
        return rc, pnew.value

    def compute_residual(self, dU, P, F, precision=_lib.EFE_FP64, time=(0.0, 0.0), dT=0.0, U=None):
        """Residual-only element pass (efe_compute_residual): the status and pNewDT."""
        t = (ctypes.c_double * 2)(*time)
        pnew = ctypes.c_double(1.0)
        rc = self.lib.efe_compute_residual(self.ctx, self._ptr(U) if U is not None else None, self._ptr(dU), t,
                                           float(dT), self._ptr(P), self._ptr(F), int(precision), ctypes.byref(pnew))
        check(rc, "efe_compute_residual")
        return rc, pnew.value

    def body_force(self, force, PExt, elements=None):
        """Body force (3,) per volume on `elements` (all if None), added into PExt."""
        f = (ctypes.c_double * 3)(*[float(v) for v in force])
        if elements is None:
            check(self.lib.efe_body_force(self.ctx, None, 0, f, self._ptr(PExt)), "efe_body_force")
            return
        el = np.ascontiguousarray(elements, dtype=np.int32)
        check(self.lib.efe_body_force(self.ctx, el.ctypes.data, int(el.size), f, self._ptr(PExt)), "efe_body_force")

    def surface_traction(self, surface, traction, PExt):
        """Uniform traction (3,) on a surface {face ID (1..6): element indices}, added into PExt."""
        el = np.concatenate([np.asarray(v, dtype=np.int32) for v in surface.values()]) if surface else \
            np.zeros(0, np.int32)
        fc = np.concatenate([np.full(len(v), int(f), dtype=np.int32) for f, v in surface.items()]) if surface else \
            np.zeros(0, np.int32)
        el, fc = np.ascontiguousarray(el), np.ascontiguousarray(fc)
        t = (ctypes.c_double * 3)(*[float(v) for v in traction])
        check(self.lib.efe_surface_traction(self.ctx, el.ctypes.data, fc.ctypes.data, int(el.size), t, self._ptr(PExt)),
              "efe_surface_traction")

    def set_geostatic(self, definition, elements=None):
        """Geostatic initial stress [p1, h1, p2, h2, xLateral, zLateral] on `elements` (all if None)."""
        d = (ctypes.c_double * 6)(*[float(v) for v in definition])
        if elements is None:
            check(self.lib.efe_set_geostatic(self.ctx, d, None, 0), "efe_set_geostatic")
        else:
            el = np.ascontiguousarray(elements, dtype=np.int32)
            check(self.lib.efe_set_geostatic(self.ctx, d, el.ctypes.data, int(el.size)), "efe_set_geostatic")

    def element_matrices(self, elements):
        """Ke [n][nDof][nDof] (column-major per element, i.e. Ke[k][col][row]) and Pe [n][nDof] of
        the last assembled element pass, as device tensors (the native element ABI layout)."""
        el = np.ascontiguousarray(elements, dtype=np.int32)
        nd = 3 * (8 if self.el_type == "C3D8" else 20)
        Ke = torch.empty((el.size, nd, nd), dtype=torch.float64, device=self.device)
        Pe = torch.empty((el.size, nd), dtype=torch.float64, device=self.device)
        check(self.lib.efe_get_element_matrices(self.ctx, el.ctypes.data, int(el.size), ctypes.c_void_p(Ke.data_ptr()),
                                                ctypes.c_void_p(Pe.data_ptr())), "efe_get_element_matrices")
        return Ke, Pe

    def element_compute(self, elements, dQ, stateRef, stateTemp, Pe, Ke=None, time=(0.0, 0.0), dT=0.0, QTotal=None):
        """The native element ABI for the listed elements (efe_element_compute): dQ [n][nDof],
        stateRef / stateTemp [n][nQP][nState], Pe [n][nDof] and Ke [n][nDof][nDof] (column-major,
        or None) are float64 device tensors in the caller's element order.  Returns pNewDT
        (1e36, or at most 0.5 when a material asked for a smaller increment)."""
        el = np.ascontiguousarray(elements, dtype=np.int32)
        nd = 3 * (8 if self.el_type == "C3D8" else 20)
        n = el.size

        def chk(t, shape, name):
            if t is None:
                return None
            if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous() and tuple(t.shape) == shape):
                raise ValueError("{} must be a contiguous float64 CUDA tensor of shape {}".format(name, shape))
            return ctypes.c_void_p(t.data_ptr())

        args = [chk(dQ, (n, nd), "dQ"), chk(stateRef, (n, self.n_qp, self.n_state), "stateRef"),
                chk(stateTemp, (n, self.n_qp, self.n_state), "stateTemp"), chk(Pe, (n, nd), "Pe"),
                chk(Ke, (n, nd, nd), "Ke")]
        t = (ctypes.c_double * 2)(*time)
        pnew = ctypes.c_double(1e36)
        rc = self.lib.efe_element_compute(self.ctx, el.ctypes.data, n, chk(QTotal, (n, nd), "QTotal"), args[0],
                                          args[1], args[2], args[3], args[4], t, float(dT), ctypes.byref(pnew))
        check(rc, "efe_element_compute")
        return pnew.value

    def export_csr_values(self):
        out = torch.empty(self.nnz, dtype=torch.float64, device=self.device)
        check(self.lib.efe_export_csr_values(self.ctx, ctypes.c_void_p(out.data_ptr())), "efe_export_csr_values")
        return out

    def spmv(self, x, y):
        check(self.lib.efe_spmv(self.ctx, self._ptr(x), self._ptr(y)), "efe_spmv")

    def linsolve(self, b, x, solver=_lib.EFE_SOLVER_CG, precond=_lib.EFE_PRECOND_BLOCK_JACOBI,
                 precision=_lib.EFE_FP64, rtol=1e-10, atol=0.0, max_iter=20000, restart=50, check_every=32):
        o = _lib.SolveOpts(solver=solver, precond=precond, precision=precision, max_iter=max_iter, restart=restart,
                           check_every=check_every, rtol=rtol, atol=atol)
        info = _lib.SolveInfo()
        check(self.lib.efe_linsolve(self.ctx, self._ptr(b), self._ptr(x), ctypes.byref(o), ctypes.byref(info)),
              "efe_linsolve")
        return info

    def norms(self, R, ddU, F):
        out = (ctypes.c_double * 5)()
        check(self.lib.efe_convergence_norms(self.ctx, self._ptr(R), self._ptr(ddU) if ddU is not None else None,
                                             self._ptr(F), out), "efe_convergence_norms")
        return np.array(out[:])

    def linsolve_batched(self, b, x, spec_atol, spec_active=None, precision=_lib.EFE_FP64, rtol=0.0, max_iter=20000,
                         check_every=16, spec_rtol=None):
        """One block-Jacobi PCG per specimen, all at once; returns (info, iterations per specimen).
        spec_rtol: per-specimen relative tolerances (default: rtol for all)."""
        o = _lib.SolveOpts(solver=_lib.EFE_SOLVER_CG, precond=_lib.EFE_PRECOND_BLOCK_JACOBI, precision=precision,
                           max_iter=max_iter, restart=0, check_every=check_every, rtol=rtol, atol=0.0)
        info = _lib.SolveInfo()
        tol = np.ascontiguousarray(spec_atol, dtype=np.float64)
        act = None if spec_active is None else np.ascontiguousarray(spec_active, dtype=np.uint8)
        its = np.zeros(self.n_specimens, dtype=np.int32)
        rt = None if spec_rtol is None else np.ascontiguousarray(spec_rtol, dtype=np.float64)
        check(self.lib.efe_linsolve_batched(self.ctx, self._ptr(b), self._ptr(x), ctypes.byref(o), tol.ctypes.data,
                                            rt.ctypes.data if rt is not None else None,
                                            act.ctypes.data if act is not None else None, its.ctypes.data,
                                            ctypes.byref(info)), "efe_linsolve_batched")
        return info, its

    def newton_round_batched(self, dU, PExt, deltaT, live, extrapolated, has_dirichlet, max_iter, max_grow, n_acc,
                             opts, P=None, F=None):
        """The Newton iterations of one lockstep round on the device (efe_newton_round_batched):
        dU (updated in place), PExt (or None), deltaT device [n_dof]; per-specimen host arrays;
        opts a _lib.NewtonOpts.  Returns (state, iterations, reason, work [n_specimens, 3])."""
        S = self.n_specimens
        u8 = lambda a: np.ascontiguousarray(a, dtype=np.uint8)  # noqa: E731
        i32 = lambda a: np.ascontiguousarray(a, dtype=np.int32)  # noqa: E731
        lv, ex, hd = u8(live), u8(extrapolated), u8(has_dirichlet)
        mi, mg = i32(max_iter), i32(max_grow)
        na = np.ascontiguousarray(n_acc, dtype=np.float64)
        for a in (lv, ex, hd, mi, mg, na):
            if a.shape != (S,):
                raise ValueError("per-specimen arrays must have n_specimens entries")
        state, its, reason = (np.zeros(S, dtype=np.int32) for _ in range(3))
        work = np.zeros((S, 3), dtype=np.int32)
        check(self.lib.efe_newton_round_batched(
            self.ctx, self._ptr(dU), self._ptr(P) if P is not None else None, self._ptr(F) if F is not None else None,
            self._ptr(PExt) if PExt is not None else None, self._ptr(deltaT), lv.ctypes.data, ex.ctypes.data,
            hd.ctypes.data, mi.ctypes.data, mg.ctypes.data, na.ctypes.data, ctypes.byref(opts), state.ctypes.data,
            its.ctypes.data, reason.ctypes.data, work.ctypes.data), "efe_newton_round_batched")
        return state, its, reason, work

    def norms_batched(self, R, ddU, F):
        """[n_specimens, 4]: max|R|, ||ddU||_inf, ||F||_1, NaN flag."""
        out = np.empty((self.n_specimens, 4))
        check(self.lib.efe_convergence_norms_batched(self.ctx, self._ptr(R), self._ptr(ddU) if ddU is not None else None,
                                                     self._ptr(F), out.ctypes.data), "efe_convergence_norms_batched")
        return out

    def specimen_status(self):
        """Per specimen: True where the last element pass requested a cutback."""
        out = np.zeros(self.n_specimens, dtype=np.uint8)
        check(self.lib.efe_get_specimen_status(self.ctx, out.ctypes.data), "efe_get_specimen_status")
        return out.astype(bool)

    def set_active_specimens(self, active):
        """Specimens the following element passes cover (None: all); see efe_set_active_specimens."""
        if active is None:
            check(self.lib.efe_set_active_specimens(self.ctx, None), "efe_set_active_specimens")
            return
        a = np.ascontiguousarray(active, dtype=np.uint8)
        check(self.lib.efe_set_active_specimens(self.ctx, a.ctypes.data), "efe_set_active_specimens")

    def accept_state_batched(self, accept):
        a = np.ascontiguousarray(accept, dtype=np.uint8)
        check(self.lib.efe_accept_state_batched(self.ctx, a.ctypes.data), "efe_accept_state_batched")

    def accept_state(self):
        check(self.lib.efe_accept_state(self.ctx), "efe_accept_state")

    def qp_state(self, which=0):
        out = np.empty((self.n_elements, self.n_qp, self.n_state))
        check(self.lib.efe_get_qp_state(self.ctx, int(which), out.ctypes.data), "efe_get_qp_state")
        return out

    def set_qp_state(self, state):
        state = np.ascontiguousarray(state, dtype=np.float64)
        check(self.lib.efe_set_qp_state(self.ctx, state.ctypes.data), "efe_set_qp_state")

    def reset_state(self):
        """Zero reference QP state (a fresh model on the same mesh), cleared on the device."""
        check(self.lib.efe_set_qp_state(self.ctx, None), "efe_set_qp_state")

    def set_launch(self, elements_per_workgroup=0, krylov_block=0):
        check(self.lib.efe_set_launch(self.ctx, int(elements_per_workgroup), int(krylov_block)), "efe_set_launch")

    def phase_times(self):
        """Device ms of the last call of each phase (the autotuner's probes)."""
        out = (ctypes.c_double * 16)()
        check(self.lib.efe_get_phase_times(self.ctx, out), "efe_get_phase_times")
        return {"elements": out[0], "assembly": out[1], "krylov_iteration": out[2], "norms": out[3]}

    def phase_totals(self):
        """Device time totals since the system was created and the history averages (ms per
        call): element kernel and assembly per tangent pass, residual-only pass, Krylov
        iteration (CG or GMRES, without preconditioner set-up), ILU(0) factorisation per solve,
        norms per call."""
        out = (ctypes.c_double * 16)()
        check(self.lib.efe_get_phase_times(self.ctx, out), "efe_get_phase_times")
        t = list(out[4:15])
        tot = dict(zip(("elements_ms", "element_passes", "assembly_ms", "residual_ms", "residual_passes",
                        "krylov_ms", "krylov_iterations", "solves", "ilu_factor_ms", "norms_ms", "norm_calls"), t))

        def avg(a, n):
            return tot[a] / tot[n] if tot[n] else 0.0

        tot["avg_ms"] = {"elements": avg("elements_ms", "element_passes"),
                         "assembly": avg("assembly_ms", "element_passes"),
                         "residual_pass": avg("residual_ms", "residual_passes"),
                         "krylov_iteration": avg("krylov_ms", "krylov_iterations"),
                         "ilu_factor": avg("ilu_factor_ms", "solves"),
                         "norms": avg("norms_ms", "norm_calls")}
        return tot

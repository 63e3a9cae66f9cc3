"""A stand-in for DeviceSystem under NISTGPU, computed with the oracle (CPU only).

It lets NISTGPU's control flow, and the drop-in solver that adapts the reference's own model
and step objects to it, be checked on the CPU against the oracle's NIST run: element routines,
loads and geostatic prestress are the oracle's, linear solves are direct, with the prescribed
DOFs eliminated as on the device (rows and columns zeroed, unit diagonal, the prescribed
values' column contribution returned as dir_correction).
"""

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla
import torch

from edelweissfe_amd import _lib
from oracle import assembly
from oracle import materials as omat
from oracle import nist as onist
from oracle.element import HexElements
from oracle.exceptions import CutbackRequest
from oracle.surface import surface_traction


class _Info:
    def __init__(self, iterations, status=_lib.EFE_OK):
        self.iterations, self.status = iterations, status


def oracle_material(material, props, nqp):
    """The oracle material of a device material description: one name, or one name per element
    (sections) with per-element rows of properties."""
    props = np.asarray(props, dtype=float)
    if isinstance(material, str):
        if props.ndim == 2:
            props = np.repeat(props, nqp, axis=0)
        return omat.MATERIALS[material](props)
    names = np.asarray(material)
    kinds = sorted(set(names))
    # one section per (material, property row)
    rows = {}
    of = np.empty(names.size, dtype=np.int64)
    mats = []
    for e, (nm, p) in enumerate(zip(names, props)):
        key = (nm, tuple(p[:2] if nm == "linearelastic" else p))
        if key not in rows:
            rows[key] = len(mats)
            mats.append(omat.MATERIALS[nm](np.array(key[1])))
        of[e] = rows[key]
    assert kinds
    return omat.Sections(mats, of, nqp)


class OracleSystem:
    def __init__(self, coords, conn, el_type, material, props, **kw):
        self.elements = HexElements(el_type, np.asarray(coords, float), np.asarray(conn))
        self.material = oracle_material(material, props, self.elements.nQP)
        self.model = onist.Model(coords, el_type, conn, self.material, {}, {}, elements=self.elements)
        self.nist = onist.NIST(convention="colmajor")
        self.nist._setup(self.model)
        self.n_nodes = self.model.coords.shape[0]
        self.n_dof = self.model.nDof
        self.n_elements = self.elements.conn.shape[0]
        self.device = torch.device("cpu")
        self.ref = self.model.state.copy()
        self.tmp = self.ref.copy()
        self.K = None
        self.dir = np.zeros(0, dtype=np.int64)
        self.el_type = el_type

    def close(self):
        pass

    def set_dirichlet(self, dofs):
        self.dir = np.unique(np.asarray(dofs, dtype=np.int64))

    def _pass(self, dU):
        self.model.state = self.ref
        try:
            V, p, f, st = self.nist.compute_elements(self.model, dU.numpy())
        except CutbackRequest as e:
            return None, e.cutbackSize
        self.tmp = st
        so = self.nist
        self.K = assembly.csr_matrix(assembly.update_csr(V, so.x, so.nnz), so.indptr, so.indices)
        return (p, f), 1.0

    def compute_elements(self, dU, P, F, dir_values=None, dir_correction=None, precision=_lib.EFE_FP64, **kw):
        r, pnew = self._pass(dU)
        if r is None:
            return _lib.EFE_CUTBACK, pnew
        P.copy_(torch.from_numpy(r[0]))
        F.copy_(torch.from_numpy(r[1]))
        if dir_values is not None:
            delta = np.zeros(self.n_dof)
            delta[self.dir] = dir_values.numpy()[self.dir]
            c = self.K @ delta
            c[self.dir] = 0.0
            dir_correction.copy_(torch.from_numpy(c))
        return _lib.EFE_OK, 1.0

    def compute_residual(self, dU, P, F, precision=_lib.EFE_FP64, **kw):
        return self.compute_elements(dU, P, F, precision=precision)

    def norms(self, R, ddU, F):
        Rh = R.numpy()
        a = np.abs(Rh)
        return np.array([a.max(), float(a.argmax()), 0.0 if ddU is None else np.abs(ddU.numpy()).max(),
                         np.abs(F.numpy()).sum(), float(np.isnan(Rh).any())])

    def linsolve(self, b, x, **kw):
        keep = np.ones(self.n_dof)
        keep[self.dir] = 0.0
        D = sp.diags(keep)
        A = (D @ self.K @ D + sp.diags(1.0 - keep)).tocsc()
        xs = spla.spsolve(A, b.numpy(), use_umfpack=False)
        x.copy_(torch.from_numpy(xs))
        return _Info(1, _lib.EFE_NAN if np.isnan(xs).any() else _lib.EFE_OK)

    def accept_state(self):
        self.ref = self.tmp.copy()

    def qp_state(self, which=0):
        return (self.ref if which == 0 else self.tmp).copy()

    def set_qp_state(self, state):
        self.ref = np.array(state, dtype=float)

    def reset_state(self):
        self.ref = np.zeros_like(self.ref)

    def body_force(self, force, PExt, elements=None):
        els = np.arange(self.n_elements) if elements is None else np.unique(np.asarray(elements))
        Pe = self.elements.body_force(np.asarray(force, dtype=float))[els]
        out = np.zeros(self.n_dof)
        np.add.at(out, self.elements.dof_map[els].ravel(), Pe.ravel())
        PExt += torch.from_numpy(out)

    def surface_traction(self, surface, traction, PExt):
        PExt += torch.from_numpy(surface_traction(self.model.coords, self.elements.conn, surface,
                                                  np.asarray(traction, dtype=float), self.n_dof))

    def set_geostatic(self, definition, elements=None):
        om = self.model
        om.state = self.ref
        om.elsets = {"_geo": np.arange(self.n_elements) if elements is None else np.asarray(elements)}
        keys = ("p1", "h1", "p2", "h2", "xLateral", "zLateral")
        onist.NIST.apply_geostatic(om, dict(zip(keys, (float(v) for v in definition)), elset="_geo"))
        self.ref = om.state


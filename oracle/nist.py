"""Nonlinear implicit static Newton step and adaptive time stepping (oracle side).

Restates, for displacement-only hexahedral models:
  * NIST.solveStep / solveIncrement   edelweissfe/solvers/nonlinearimplicitstatic.py:110-458
  * applyDirichletK / applyDirichlet  :559-623
  * checkConvergence, spatial fluxes  :625-725, :771-792, divergence :988-1006
  * computeElements, loads            :794-849, :903-945, body forces :510-557
  * extrapolateLastIncrement          :947-986
  * AdaptiveTimeStepper               edelweissfe/timesteppers/adaptivetimestepper.py:73-183
  * AdaptiveStep defaults             edelweissfe/steps/adaptivestep.py:56-113
  * dirichlet.getDelta                edelweissfe/stepactions/dirichlet.py:105-112
  * nodeforces.getCurrentLoad         edelweissfe/stepactions/nodeforces.py:136-143
  * bodyforce.getCurrentLoad          edelweissfe/stepactions/bodyforce.py
Tolerances from edelweissfe/config/phenomena.py:59-94 (displacement field).
The linear solve defaults to SciPy SuperLU, the reference's ``superlu`` option
(edelweissfe/config/linsolve.py:47-50).
"""

from dataclasses import dataclass, field

import numpy as np
import scipy.sparse.linalg as spla

from . import assembly
from .element import HexElements
from .surface import surface_traction
from .exceptions import (CutbackRequest, DivergingSolution, ReachedMaxIncrements, ReachedMaxIterations,
                         ReachedMinIncrementSize, StepFailed)

FIELD_CORRECTION_TOL = 1e-8
FLUX_RESIDUAL_TOL = 1e-8
FLUX_RESIDUAL_TOL_ALT = 5e-3


@dataclass
class TimeStep:
    number: int
    stepProgressIncrement: float
    stepProgress: float
    timeIncrement: float
    stepTime: float
    totalTime: float


class AdaptiveTimeStepper:
    def __init__(self, currentTime, stepLength, startIncrement, maxIncrement, minIncrement, maxNumberIncrements,
                 increaseFactor=1.1):
        self.nPassedGoodIncrements = 0
        self.incrementCounter = 0
        self.startIncrement = startIncrement
        self.maxIncrement = maxIncrement
        self.minIncrement = minIncrement
        self.maxNumberIncrements = maxNumberIncrements
        self.finishedStepProgress = 0.0
        self.increment = min(startIncrement, maxIncrement)
        self.allowedToIncreasedNext = True
        self.currentTime = currentTime
        self.stepLength = stepLength
        self.increaseFactor = increaseFactor

    def generateTimeStep(self):
        while self.finishedStepProgress < (1.0 - 1e-15):
            remainder = 1.0 - self.finishedStepProgress
            if remainder < self.increment:
                self.increment = remainder
            theIncrement = self.increment if self.incrementCounter > 0 else 0.0
            dT = self.stepLength * theIncrement
            self.finishedStepProgress += theIncrement
            endInStep = self.stepLength * self.finishedStepProgress
            yield TimeStep(self.incrementCounter, theIncrement, self.finishedStepProgress, dT, endInStep,
                           self.currentTime + endInStep)
            if self.incrementCounter > self.maxNumberIncrements:
                raise ReachedMaxIncrements()
            if (self.nPassedGoodIncrements >= 3) and self.allowedToIncreasedNext:
                self.increment *= self.increaseFactor
                if self.increment > self.maxIncrement:
                    self.increment = self.maxIncrement
            self.allowedToIncreasedNext = True
            self.incrementCounter += 1
            self.nPassedGoodIncrements += 1

    def preventIncrementIncrease(self):
        self.allowedToIncreasedNext = False

    def reduceNextIncrement(self, scaleFactor):
        if self.increment == self.minIncrement:
            raise ReachedMinIncrementSize()
        newIncrement = self.increment * scaleFactor
        if newIncrement > self.maxIncrement:
            self.increment = self.maxIncrement
        elif newIncrement < self.minIncrement:
            self.increment = self.minIncrement
        else:
            self.increment = newIncrement

    def discardAndChangeIncrement(self, scaleFactor):
        if self.incrementCounter == 0:
            raise ReachedMinIncrementSize()
        self.finishedStepProgress -= self.increment
        self.incrementCounter -= 1
        self.nPassedGoodIncrements = 0
        self.reduceNextIncrement(scaleFactor)


@dataclass
class Step:
    """AdaptiveStep options (adaptivestep.py:56-113) and its step actions."""
    startInc: float = 1.0
    maxInc: float = 1.0
    minInc: float = 1e-4
    maxNumInc: int = 1000
    maxIter: int = 10
    criticalIter: int = 5
    maxGrowIter: int = 10
    cutbackFactor: float = 0.25
    stepLength: float = 1.0
    extrapolation: str = "linear"
    dirichlets: list = field(default_factory=list)   # dict(nset=str, values={comp: value})
    nodeforces: list = field(default_factory=list)   # dict(nset=str, vector=(3,))
    bodyforces: list = field(default_factory=list)   # dict(elset=str, vector=(3,))
    # dict(surface={face ID: elset name}, vector=(3,)): uniform surface traction,
    # stepactions/distributedload.py:52-110 with type "surface traction"
    distributedloads: list = field(default_factory=list)
    # dict(p1, h1, p2, h2, xLateral, zLateral, elset="all"): geostatic initial stress,
    # stepactions/geostatic.py:41-101
    geostatic: list = field(default_factory=list)


class Model:
    """Nodes, one homogeneous hexahedral element set with its material, sets."""

    def __init__(self, coords, el_type, conn, material, nodesets, elsets, elements=None):
        """elements: an element set object (default: HexElements of el_type); its `dim` is the
        number of displacement DOFs per node (3 unless given)."""
        self.coords = np.asarray(coords, dtype=float)
        self.elements = elements if elements is not None else HexElements(el_type, self.coords, conn)
        self.dim = getattr(self.elements, "dim", 3)
        self.material = material
        self.nodesets = nodesets
        self.elsets = elsets
        self.nDof = self.dim * self.coords.shape[0]
        E, Q = self.elements.conn.shape[0], self.elements.nQP
        self.state = np.zeros((E, Q, 12 + material.n_state))
        self.U = np.zeros(self.nDof)
        self.time = 0.0


def superlu(A, b):
    return spla.spsolve(A, b, use_umfpack=False)


class NIST:
    """Oracle NIST solver.  ``linsolver(A_csr, b) -> x``."""

    def __init__(self, linsolver=superlu, element_kernel=None, convention="reference"):
        """convention: "reference" assembles Ke^T like DisplacementElement (element.py:318 read
        column-major by dofmanager.py:551-555); "colmajor" assembles Ke (the native element ABI,
        marmotelement/element.pxd:97-103), required for non-symmetric tangents."""
        self.linsolver = linsolver
        self.element_kernel = element_kernel
        self.convention = convention
        self.history = []   # (increment number, stepProgress, iterations or None on failure)
        # per increment, the fields of the reference's status-file line (statusInfoDict,
        # nonlinearimplicitstatic.py:207-215, :250-253, :267-270, :304-305)
        self.increments = []

    # -- setup, as solveStep :133-164
    def _setup(self, model):
        el = model.elements
        self.dof_map = el.dof_map
        I, J = assembly.vij_pattern(self.dof_map)
        self.indptr, self.indices = assembly.csr_pattern(I, J, model.nDof)
        self.x = assembly.x_map(I, J, self.indptr, self.indices)
        self.nnz = self.indices.size
        self.nAccFlux = el.conn.shape[0] * el.nn * model.dim

    def dirichlet_indices(self, model, d):
        nodes = model.nodesets[d["nset"]]
        comps = sorted(d["values"].keys())
        idx = (model.dim * np.asarray(nodes)[:, None] + np.arange(model.dim)[None, :])[:, comps]
        return idx.ravel()

    def dirichlet_delta(self, model, d, ts):
        nodes = model.nodesets[d["nset"]]
        comps = sorted(d["values"].keys())
        vals = np.tile([float(d["values"][c]) for c in comps], (len(nodes), 1))
        if not d.get("active", True):
            return (vals * 0.0).ravel()
        amp = d.get("amplitude", lambda t: t)
        return (vals * (amp(ts.stepProgress) - amp(ts.stepProgress - ts.stepProgressIncrement))).ravel()

    def compute_elements(self, model, dU):
        el = model.elements
        if self.element_kernel is not None:
            return self.element_kernel(model, dU)
        Ke, Pe, st = el.compute(dU[self.dof_map], model.state, model.material)
        V = assembly.vij_values_reference(Ke) if self.convention == "reference" else assembly.vij_values_colmajor(Ke)
        P = np.zeros(model.nDof)
        F = np.zeros(model.nDof)
        np.add.at(P, self.dof_map.ravel(), Pe.ravel())
        np.add.at(F, self.dof_map.ravel(), np.abs(Pe).ravel())
        return V, P, F, st

    def external_loads(self, model, step, ts):
        """Loads start + vector * stepProgress; `start` (default 0) is the load a previous step
        left (nodeforces.py:136-152, bodyforce.py)."""
        PExt = np.zeros(model.nDof)
        for nf in step.nodeforces:
            nodes = model.nodesets[nf["nset"]]
            idx = (model.dim * np.asarray(nodes)[:, None] + np.arange(model.dim)[None, :]).ravel()
            vec = np.asarray(nf.get("start", 0.0), dtype=float) + np.asarray(nf["vector"], dtype=float) * ts.stepProgress
            PExt[idx] += np.tile(vec, (len(nodes), 1)).ravel()
        for dl in step.distributedloads:
            vec = np.asarray(dl.get("start", 0.0), dtype=float) + np.asarray(dl["vector"], dtype=float) * ts.stepProgress
            surface = {f: model.elsets[s] for f, s in dl["surface"].items()}
            PExt += surface_traction(model.coords, model.elements.conn, surface, vec, model.nDof)
        for bf in step.bodyforces:
            els = model.elsets[bf["elset"]]
            force = np.asarray(bf.get("start", 0.0), dtype=float) + np.asarray(bf["vector"], dtype=float) * ts.stepProgress
            Pe = model.elements.body_force(force)[els]
            np.add.at(PExt, self.dof_map[els].ravel(), Pe.ravel())
        return PExt

    @staticmethod
    def apply_geostatic(model, g):
        """Geostatic initial stress into the reference state of the element set, at every
        iteration start (nonlinearimplicitstatic.py:413-414 -> geostatic.py:94-101 ->
        setInitialCondition, marmotelement/element.pyx:203-213: Temp = Ref, set, Ref = Temp).
        The stress rule is the one of Marmot's displacement elements: at each QP,
        sigma_yy = p1 + (p2 - p1) (y - h1) / (h2 - h1), sigma_xx = xLateral sigma_yy,
        sigma_zz = zLateral sigma_yy; shear stresses and the other state stay."""
        els = model.elsets[g.get("elset", "all")]
        y = model.elements.qp_coordinates()[els, :, 1]
        sy = g["p1"] + (g["p2"] - g["p1"]) * (y - g["h1"]) / (g["h2"] - g["h1"])
        model.state[els, :, 0] = g.get("xLateral", 1.0) * sy
        model.state[els, :, 1] = sy
        model.state[els, :, 2] = g.get("zLateral", 1.0) * sy

    def apply_dirichlet_K(self, K, dir_idx):
        if dir_idx.size:
            for row in dir_idx:
                K.data[K.indptr[row]:K.indptr[row + 1]] = 0.0
            diag = K.diagonal()
            diag[dir_idx] = 1.0
            K.setdiag(diag)
            K.eliminate_zeros()
        return K

    def check_convergence(self, R, ddU, F, iterationCounter, history):
        avgFlux = max(1e-10, np.linalg.norm(F, 1) / self.nAccFlux)
        tolFlux = FLUX_RESIDUAL_TOL if iterationCounter < 15 else FLUX_RESIDUAL_TOL_ALT
        fluxResidual = np.max(np.abs(R))
        correction = np.linalg.norm(ddU, np.inf) if ddU is not None else 0.0
        convergedCorrection = correction < FIELD_CORRECTION_TOL
        convergedFlux = fluxResidual <= max(tolFlux * avgFlux, 1e-7)
        prev, nGrew = history
        if fluxResidual > prev:
            nGrew += 1
        return convergedCorrection and convergedFlux, (fluxResidual, nGrew)

    def solve_increment(self, model, step, ts, prev_ts, U_n, dU, dir_list):
        iterationCounter = 0
        history = (0.0, 0)
        ddU = None
        dir_all = np.concatenate([ix for ix, _ in dir_list]) if dir_list else np.zeros(0, dtype=np.int64)
        if step.extrapolation == "linear" and prev_ts is not None and prev_ts.timeIncrement:
            dU = dU * (ts.stepProgressIncrement / prev_ts.stepProgressIncrement)
            for ix, d in dir_list:
                dU[ix] = self.dirichlet_delta(model, d, ts)
            extrapolated = True
        else:
            extrapolated = False
            dU = np.zeros_like(dU)
        while True:
            for g in step.geostatic:
                self.apply_geostatic(model, g)
            U_np = U_n + dU
            V, P, F, st = self.compute_elements(model, dU)
            PExt = self.external_loads(model, step, ts)
            R = P + PExt
            if iterationCounter == 0 and not extrapolated and dir_list:
                for ix, d in dir_list:
                    R[ix] = self.dirichlet_delta(model, d, ts)
            else:
                R[dir_all] = 0.0
                converged, history = self.check_convergence(R, ddU, F, iterationCounter, history)
                if converged:
                    break
                if history[1] > step.maxGrowIter:
                    raise DivergingSolution("Residual grew {:} times, cutting back".format(step.maxGrowIter))
                if iterationCounter == step.maxIter:
                    raise ReachedMaxIterations("Reached max. iterations in current increment, cutting back")
            K = assembly.csr_matrix(assembly.update_csr(V, self.x, self.nnz), self.indptr, self.indices)
            K = self.apply_dirichlet_K(K, dir_all)
            ddU = self.linsolver(K, R)
            if np.isnan(ddU).any():
                raise DivergingSolution("Obtained NaN in linear solve")
            dU = dU + ddU
            iterationCounter += 1
        return U_np, dU, P, st, iterationCounter

    def _status(self, ts, iters, converged, notes):
        self.increments.append({"inc": ts.number, "iters": iters, "converged": converged,
                                "time inc": ts.timeIncrement, "time end": ts.totalTime, "notes": notes})

    def solve_step(self, model, step):
        self._setup(model)
        stepper = AdaptiveTimeStepper(model.time, step.stepLength, step.startInc, step.maxInc, step.minInc,
                                      step.maxNumInc)
        dir_list = [(self.dirichlet_indices(model, d), d) for d in step.dirichlets]
        U = model.U.copy()
        dU = np.zeros(model.nDof)
        prev_ts = None
        try:
            for ts in stepper.generateTimeStep():
                try:
                    U_new, dU, P, st, iters = self.solve_increment(model, step, ts, prev_ts, U, dU, dir_list)
                except CutbackRequest as e:
                    stepper.discardAndChangeIncrement(max(e.cutbackSize, step.cutbackFactor))
                    prev_ts = None
                    self.history.append((ts.number, ts.stepProgress, None))
                    self._status(ts, np.inf, False, str(e))
                except (ReachedMaxIterations, DivergingSolution) as e:
                    stepper.discardAndChangeIncrement(step.cutbackFactor)
                    prev_ts = None
                    self.history.append((ts.number, ts.stepProgress, None))
                    self._status(ts, np.inf, False, str(e))
                else:
                    prev_ts = ts
                    if iters >= step.criticalIter:
                        stepper.preventIncrementIncrease()
                    U = U_new
                    model.state = st
                    model.time = ts.totalTime
                    self.history.append((ts.number, ts.stepProgress, iters))
                    self._status(ts, iters, True, "")
                    # per converged increment: the displacements and QP stresses, for parity
                    # checks increment by increment
                    self.increments[-1]["U"] = U.copy()
                    self.increments[-1]["stress"] = st[..., 0:6].copy()
        except (ReachedMaxIncrements, ReachedMinIncrementSize):
            raise StepFailed()
        model.U = U
        self.P = P
        self.dU = dU
        if step.geostatic:
            # end of a geostatic step: displacements are reset (geostatic.py:81-92)
            model.U = np.zeros_like(U)
        return U

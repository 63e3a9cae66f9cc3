"""NISTGPU: the nonlinear implicit static Newton step on one MI355X.

Drop-in twin of the reference solver plugin ``NIST``
(edelweissfe/solvers/nonlinearimplicitstatic.py:60-458): same option names and
defaults (NISTOptions, :73-80, plus the device options below), same
``solveStep(step, model, fieldOutputController, outputmanagers)`` entry point,
same increment / Newton control flow and exceptions:

* zero increment first, linear extrapolation of dU from the previous
  increment (:947-986);
* iteration 0 of a non-extrapolated increment with Dirichlet BCs: R[dir] = delta
  and an unconditional solve (:428-430); afterwards R[dir] = 0 and the field-wise
  convergence test ||ddU||_inf < 1e-8 and max|R| <= max(1e-8 * avgFlux, 1e-7),
  5e-3 from iteration 15 on (:625-725, phenomena.py:59-94);
* divergence after maxGrowIter residual growths, ReachedMaxIterations at
  maxIter, CutbackRequest(0.5) from a failing material, DivergingSolution on NaN
  (:247-279, :443-449, :748-749);
* state accepted only on convergence (:296).

Every per-DOF and per-element operation runs on the GPU through libefe_gpu
(elements + assembly, Dirichlet elimination, norms, Krylov solve); only the
scalar control decisions are taken on the host, after one readback of the
convergence quantities per Newton iteration.

Device options (in addition to NISTOptions):
  precision       "fp64" (default) or "mixed": FP32 stiffness contractions,
                  FP32 Krylov inner solve, FP64 material update, residual and
                  Newton update (the Newton loop is the FP64 refinement);
  linsolver       "pcg" (default, block-Jacobi PCG, symmetric tangents) or "gmres"
                  (GMRES(gmresRestart) for non-symmetric tangents);
  preconditioner  "auto" (pcg: block Jacobi, gmres: multicolor block ILU(0)), "bjacobi", "ilu0";
  mixedOperator   FP32 operator of the mixed form's PCG: "matfree" (K applied element by
                  element from compact per-QP FP32 tangents, no FP32 assembly), "assembled"
                  (FP32 sliced-ELL operator) or "auto" (default: both are timed once per
                  system, edelweissfe_amd.autotune.mixed_operator_costs, and every increment
                  takes the one with the smaller element-pass + R x Krylov-iteration time,
                  R = Krylov iterations per element pass observed so far).  GMRES always
                  uses the assembled operator (ILU(0) needs it);
  gmresRestart    restart length m of GMRES(m) (default 50);
  krylovAtolFactor the Krylov solve stops once ||r||_2 <= krylovAtolFactor times the
                  Newton flux threshold max(tol * avgFlux, 1e-7) of the current
                  iteration (default 1e-2), so the next convergence check sees the
                  same residual class a direct solve would leave;
  krylovRtol      relative floor of the FP64 solve (default 1e-13);
  krylovRtolMixed relative residual of the FP32 inner solve (default 1e-4: on the J2
                  load histories the displacements then stay within a few 1e-9 of the
                  FP64 path, inside the 1e-8 target, with ~13 % fewer Krylov iterations
                  than at 1e-5);
  krylovRtolMixedFirst relative residual of the FP32 PCG solve of iteration 0 of an increment
                  that solves its prescribed increment (R[dir] = delta, not extrapolated;
                  default 1e-7): that solve carries the whole increment, and taken this far
                  its correction falls under the 1e-8 correction tolerance at the next
                  check, so an elastic increment converges in the two Newton iterations of
                  the FP64 path instead of three (64^3 history: 9 % fewer Krylov iterations);
                  GMRES keeps krylovRtolMixed;
  krylovMaxIter, krylovCheckEvery;
  residualFirst   (default True) the convergence check of Newton iteration 0 of an increment
                  that needs no prescribed-value solve runs on a residual-only element pass
                  (efe_compute_residual); the tangent pass follows only if the check fails.
                  Same residual, same decisions and results bit for bit; the tangents that
                  the reference forms and drops in increments that converge at once are
                  not formed.
"""

import time as _time

import numpy as np
import torch

from .. import _lib
from ..device import DeviceSystem
from ..exceptions import (ConditionalStop, CutbackRequest, DivergingSolution, ReachedMaxIncrements,
                          ReachedMaxIterations, ReachedMinIncrementSize, StepFailed)
from ..outputmanagers.timing import createTimingDict

FIELD_CORRECTION_TOL = 1e-8      # phenomena.py fieldCorrectionTolerance["displacement"]
FLUX_RESIDUAL_TOL = 1e-8         # phenomena.py fluxResidualTolerance["displacement"]
FLUX_RESIDUAL_TOL_ALT = 5e-3     # phenomena.py fluxResidualToleranceAlternative["displacement"]


def _cutbackMessage(model):
    """Text of the CutbackRequest a failing material raises: the von Mises local Newton
    (vonmises.py:230-231), or a Drucker-Prager trial state beyond the apex of a cone with
    zero dilatancy, which has no return."""
    names = model.materialNames()
    msgs = []
    if "vonmises" in names:
        msgs.append("Von Mises Newton failed.")
    if "druckerprager" in names:
        msgs.append("Drucker-Prager: trial stress beyond the apex with zero dilatancy, no return.")
    return " / ".join(msgs) or "Material update failed."


def _convertOption(default, value):
    """An option value in the type of its default; strings as the reference's option step
    action holds them (stepactions/options.py) convert too, "False" / "0" to False."""
    if isinstance(default, bool) and isinstance(value, str):
        v = value.strip().lower()
        if v in ("true", "1", "yes", "on"):
            return True
        if v in ("false", "0", "no", "off"):
            return False
        raise ValueError("invalid boolean option value {:}".format(value))
    return type(default)(value)


_DIGESTS = []  # (array, digest) of recently hashed arrays, the arrays kept alive


def _arrayDigest(a):
    """Content digest of an array, remembered for the same (unchanged-identity) array object
    so that following steps on one mesh do not hash it again."""
    import hashlib

    for arr, d in _DIGESTS:
        if arr is a:
            return d
    c = np.ascontiguousarray(a)
    h = hashlib.blake2b(digest_size=16)
    h.update((str(c.shape) + str(c.dtype)).encode())
    h.update(c.tobytes())
    d = h.hexdigest()
    _DIGESTS.insert(0, (a, d))
    del _DIGESTS[16:]
    return d


def _fingerprint(model):
    """Identity of what a device system is built from: coordinates, connectivity, element
    type, materials and properties."""
    names = "|".join(np.atleast_1d(model.material["name"]).astype(str).tolist())
    return (_arrayDigest(model.coordinates), _arrayDigest(model.connectivity),
            _arrayDigest(np.asarray(model.material["properties"])), model.elementType, names)


class NISTGPU:
    identification = "NISTGPUSolver"

    NISTOptions = {
        "defaultMaxIter": 10,
        "defaultCriticalIter": 5,
        "defaultMaxGrowingIter": 10,
        "extrapolation": "linear",
        "linsolver": "pcg",
        "linsolverConfigFile": "",
        "preconditioner": "auto",
        "gmresRestart": 50,
        "precision": "fp64",
        "mixedOperator": "auto",
        "krylovAtolFactor": 1e-2,
        "krylovRtol": 1e-13,
        "krylovRtolMixed": 1e-4,
        "krylovRtolMixedFirst": 1e-7,
        "krylovMaxIter": 50000,
        "krylovCheckEvery": 32,
        "residualFirst": True,
    }

    def __init__(self, jobInfo=None, journal=None, **kwargs):
        self.journal = journal
        self.options = self.NISTOptions.copy()
        self._updateOptions(kwargs)
        self.system = None
        # per increment: the status of the reference's status file (step, inc, iters, converged,
        # time inc, time end, notes) plus stepProgress, Krylov iterations per solve and precision
        self.history = []
        # running host-side times per category (nonlinearimplicitstatic.py:166, config/timing.py)
        self.computationTimes = createTimingDict()
        self.krylovIterations = 0
        self.linearSolves = 0
        self.elementPasses = 0     # element passes that formed tangents
        self.residualPasses = 0    # residual-only passes (residualFirst)
        self.newtonChecks = 0      # element evaluations of the reference: one per Newton iteration

    def _updateOptions(self, updated):
        """updated: a dict, or the reference's option step action (its `options` dict,
        stepactions/options.py; nonlinearimplicitstatic.py:168-171)."""
        updated = getattr(updated, "options", updated)
        for k, v in updated.items():
            if k not in self.NISTOptions:
                raise AttributeError("Invalid option {:} for {:}".format(k, self.identification))
            self.options[k] = _convertOption(self.NISTOptions[k], v)

    # ---- setup ----------------------------------------------------------------------------------------
    def _setupSystem(self, model):
        """The device system of `model`: reused while coordinates, connectivity, element type,
        materials and properties are the same (its state then continues when the model is the
        one it last solved, or is set from the model's QP state, or starts at zero), rebuilt
        otherwise."""
        mat = model.material
        if mat is None:
            raise ValueError("model has no material")
        names = np.atleast_1d(mat["name"]).astype(str)
        if (names == "").any():
            raise ValueError("elements without a material (no section covers them)")
        key = _fingerprint(model)
        if self.system is not None and getattr(self, "_systemKey", None) == "attached":
            self._systemKey = key  # attached without a model: the system is this model's
        if self.system is None or getattr(self, "_systemKey", None) != key:
            if self.system is not None:
                self.system.close()
            self.system = DeviceSystem(model.coordinates, model.connectivity, model.elementType, mat["name"],
                                       mat["properties"])
            self._systemKey = key
            if model._qpState is not None:
                self.system.set_qp_state(model._qpState)
        elif model.deviceSystem is self.system:
            pass  # a following step of the same model: the state is already on the device
        elif model._qpState is not None:
            self.system.set_qp_state(model._qpState)
        else:
            self.system.reset_state()
        return self.system

    def attach(self, system, model=None):
        """Use an already created DeviceSystem (mesh, pattern and buffers resident on the GPU);
        with `model`, the system is the one for that model's mesh and materials."""
        self.system = system
        self._systemKey = _fingerprint(model) if model is not None else "attached"

    @property
    def _mixed(self):
        return self.options["precision"].lower() == "mixed"

    # ---- public interface (nonlinearimplicitstatic.py:110) --------------------------------------------------
    def solveStep(self, step, model, fieldOutputController=None, outputmanagers=()):
        opts = step.actions.get("options", {}).get("NISTSolver", None)
        if opts:
            self._updateOptions(opts)
        sysm = self._setupSystem(model)
        dev = sysm.device
        D = sysm.n_dof
        dirichlets = list(step.actions["dirichlet"].values())
        self._dirIdx = [d.indices(model) for d in dirichlets]
        cat = np.concatenate(self._dirIdx) if dirichlets else np.zeros(0, dtype=np.int64)
        # R[idx] = delta is applied Dirichlet by Dirichlet in the reference (:616-618): on a DOF
        # named twice the last one wins, so keep the last occurrence of every DOF
        rev = cat[::-1]
        _, first_in_rev = np.unique(rev, return_index=True)
        self._dirPick = np.sort(cat.size - 1 - first_in_rev)
        allDir = cat[self._dirPick]
        sysm.set_dirichlet(allDir)
        self._dirAll = torch.as_tensor(allDir, dtype=torch.int64, device=dev)
        self._nAccFlux = model.nElements * (8 if model.elementType == "C3D8" else 20) * 3

        U = torch.as_tensor(model.U, dtype=torch.float64, device=dev).clone()
        dU = torch.zeros(D, dtype=torch.float64, device=dev)
        P = torch.zeros(D, dtype=torch.float64, device=dev)
        prevTimeStep = None
        maxIter, criticalIter = step.maxIter, step.criticalIter
        maxGrowingIter, cutbackFactor = step.maxGrowIter, step.cutbackFactor
        # results go to the model after every increment only when someone reads them there
        perIncrement = bool(outputmanagers) or fieldOutputController is not None
        try:
            for timeStep in step.getTimeStep():
                record = {"step": step.number, "inc": timeStep.number, "iters": None, "converged": False,
                          "time inc": timeStep.timeIncrement, "time end": timeStep.totalTime, "notes": "",
                          "stepProgress": timeStep.stepProgress, "krylov": [], "precision": self._precisionCode()}
                self.history.append(record)
                try:
                    U_new, dU, P, iters = self.solveIncrement(U, dU, P, step, model, timeStep, prevTimeStep,
                                                              self.options["extrapolation"], maxIter,
                                                              maxGrowingIter, record)
                except CutbackRequest as e:
                    step.discardAndChangeIncrement(max(e.cutbackSize, cutbackFactor))
                    prevTimeStep = None
                    record["iters"] = np.inf
                    record["notes"] = str(e)
                    self._finalizeOutputs(outputmanagers, record, failed=True)
                except (ReachedMaxIterations, DivergingSolution) as e:
                    step.discardAndChangeIncrement(cutbackFactor)
                    prevTimeStep = None
                    record["iters"] = np.inf
                    record["notes"] = str(e)
                    self._finalizeOutputs(outputmanagers, record, failed=True)
                else:
                    prevTimeStep = timeStep
                    if iters >= criticalIter:
                        step.preventIncrementIncrease()
                    U = U_new
                    sysm.accept_state()
                    model.time = timeStep.totalTime
                    if perIncrement:
                        self._writeResults(model, U, P, dU)
                    record["iters"] = iters
                    record["converged"] = True
                    if fieldOutputController is not None:
                        fieldOutputController.finalizeIncrement()
                    self._finalizeOutputs(outputmanagers, record, failed=False)
        except (ReachedMaxIncrements, ReachedMinIncrementSize):
            raise StepFailed()
        except ConditionalStop:
            pass  # a successful end of the step (nonlinearimplicitstatic.py:323-325)
        if dev.type == "cuda":
            torch.cuda.synchronize(dev)
        self._writeResults(model, U, P, dU)
        model._qpState = None
        model.deviceSystem = sysm
        for group in ("dirichlet", "nodeforces", "distributedload", "bodyforce", "geostatic"):
            for action in step.actions.get(group, {}).values():
                action.applyAtStepEnd(model)
        return True, model

    @staticmethod
    def _writeResults(model, U, P, dU):
        model.U = U.cpu().numpy()
        model.P = P.cpu().numpy()
        model.dU = dU.cpu().numpy()

    _STATUS_KEYS = ("step", "inc", "iters", "converged", "time inc", "time end", "notes")

    def _finalizeOutputs(self, outputmanagers, record, failed):
        """finalize(Failed)Increment of every output manager with the reference's statusInfoDict
        and the running computationTimes (nonlinearimplicitstatic.py:255-315); the time spent
        there is charged to "output".  A ConditionalStop raised by a manager propagates."""
        if not outputmanagers:
            return
        status = {k: record[k] for k in self._STATUS_KEYS}
        tic = _time.perf_counter()
        try:
            for man in outputmanagers:
                if failed:
                    man.finalizeFailedIncrement(statusInfoDict=status, currentComputingTimes=self.computationTimes)
                else:
                    man.finalizeIncrement(statusInfoDict=status, currentComputingTimes=self.computationTimes)
        finally:
            self.computationTimes["output"] += _time.perf_counter() - tic

    # ---- one increment (nonlinearimplicitstatic.py:334-458) ---------------------------------------------------
    def solveIncrement(self, U_n, dU, P, step, model, timeStep, prevTimeStep, extrapolation, maxIter, maxGrowingIter,
                       record=None):
        sysm = self.system
        dev = sysm.device
        D = sysm.n_dof
        dirichlets = list(step.actions["dirichlet"].values())
        prec = record["precision"] if record is not None else self._precisionCode()
        iterationCounter = 0
        history = (0.0, 0)
        ddU = None
        F = torch.zeros(D, dtype=torch.float64, device=dev)
        R = torch.zeros(D, dtype=torch.float64, device=dev)
        dirVals = None

        if dirichlets:
            delta = np.concatenate([d.getDelta(model, timeStep) for d in dirichlets])[self._dirPick]
            delta = torch.as_tensor(delta, dtype=torch.float64).to(dev, non_blocking=True)
        if extrapolation == "linear" and prevTimeStep is not None and prevTimeStep.timeIncrement:
            dU = dU * (timeStep.stepProgressIncrement / prevTimeStep.stepProgressIncrement)
            if dirichlets:
                dU[self._dirAll] = delta
            extrapolated = True
        else:
            extrapolated = False
            dU = torch.zeros(D, dtype=torch.float64, device=dev)
        PExt = self._externalLoads(step, model, timeStep)

        geostatics = list(step.actions.get("geostatic", {}).values())
        while True:
            for g in geostatics:  # nonlinearimplicitstatic.py:413-414
                g.applyAtIterationStart(sysm, model)
            first = iterationCounter == 0 and not extrapolated and bool(dirichlets)
            if first:
                dirVals = torch.zeros(D, dtype=torch.float64, device=dev)
                dirVals[self._dirAll] = delta
                corr = torch.empty(D, dtype=torch.float64, device=dev)
            else:
                corr = None
            # Newton iteration 0 without a prescribed-value solve checks convergence first:
            # a residual-only pass, and the tangent pass only if a solve follows
            residualOnly = self.options["residualFirst"] and iterationCounter == 0 and not first
            self.newtonChecks += 1
            tic = _time.perf_counter()
            if residualOnly:
                st, pnew = sysm.compute_residual(dU, P, F, precision=prec, time=(timeStep.stepTime, timeStep.totalTime),
                                                 dT=timeStep.timeIncrement)
                self.residualPasses += 1
            else:
                st, pnew = sysm.compute_elements(dU, P, F, dir_values=dirVals if first else None,
                                                 dir_correction=corr, precision=prec,
                                                 time=(timeStep.stepTime, timeStep.totalTime),
                                                 dT=timeStep.timeIncrement)
                self.elementPasses += 1
            self.computationTimes["elements"] += _time.perf_counter() - tic
            if st == _lib.EFE_CUTBACK:
                raise CutbackRequest(_cutbackMessage(model), pnew)
            tic = _time.perf_counter()
            torch.add(P, PExt, out=R)
            if first:
                R[self._dirAll] = delta
            elif self._dirAll.numel():
                R[self._dirAll] = 0.0
            self.computationTimes["dirichlet R"] += _time.perf_counter() - tic
            tic = _time.perf_counter()
            norms = self.system.norms(R, ddU, F)
            self.computationTimes["convergence check"] += _time.perf_counter() - tic
            fluxThreshold = self._fluxThreshold(norms, iterationCounter)
            if not first:
                converged, history = self._checkConvergence(norms, ddU, iterationCounter, history)
                if converged:
                    break
                if history[1] > maxGrowingIter:
                    raise DivergingSolution("Residual grew {:} times, cutting back".format(maxGrowingIter))
                if iterationCounter == maxIter:
                    raise ReachedMaxIterations("Reached max. iterations in current increment, cutting back")
            if residualOnly:  # not converged: the solve needs the tangent of this same state
                tic = _time.perf_counter()
                st, pnew = sysm.compute_elements(dU, P, F, precision=prec,
                                                 time=(timeStep.stepTime, timeStep.totalTime),
                                                 dT=timeStep.timeIncrement)  # P and F as before
                self.computationTimes["elements"] += _time.perf_counter() - tic
                self.elementPasses += 1
                if st == _lib.EFE_CUTBACK:
                    raise CutbackRequest(_cutbackMessage(model), pnew)
            b = R - corr if first else R
            ddU = torch.empty(D, dtype=torch.float64, device=dev)
            tic = _time.perf_counter()
            info = self._linearSolve(b, ddU, prec, fluxThreshold, first=first)
            self.computationTimes["linear solve"] += _time.perf_counter() - tic
            if record is not None:
                record["krylov"].append(int(info.iterations))
            if info.status == _lib.EFE_NAN:
                raise DivergingSolution("Obtained NaN in linear solve")
            if info.status == _lib.EFE_NOT_CONVERGED:
                raise DivergingSolution("Krylov solver did not converge")
            dU = dU + ddU
            iterationCounter += 1
        return U_n + dU, dU, P, iterationCounter

    def _precisionCode(self):
        if not self._mixed:
            return _lib.EFE_FP64
        op = self.options["mixedOperator"].lower()
        if op not in ("matfree", "assembled", "auto"):
            raise AttributeError("invalid mixedOperator {:}".format(self.options["mixedOperator"]))
        if self.options["linsolver"].lower() != "pcg":
            return _lib.EFE_FP32
        if op == "auto":
            from .. import autotune

            costs = autotune.mixed_operator_costs(self.system)
            ratio = self.krylovIterations / self.elementPasses if self.elementPasses else 100.0
            op = autotune.choose_mixed_operator(costs, ratio)
        return _lib.EFE_FP32_MATFREE if op == "matfree" else _lib.EFE_FP32

    def _linearSolve(self, b, x, prec, fluxThreshold, first=False):
        o = self.options
        name = o["linsolver"].lower()
        if prec == _lib.EFE_FP64:
            rtol = o["krylovRtol"]
        elif first and name == "pcg":
            rtol = o["krylovRtolMixedFirst"]
        else:
            rtol = o["krylovRtolMixed"]
        if name not in ("pcg", "gmres"):
            raise AttributeError("invalid linear solver {:} requested".format(o["linsolver"]))
        solver = _lib.EFE_SOLVER_CG if name == "pcg" else _lib.EFE_SOLVER_GMRES
        pc = o["preconditioner"].lower()
        if pc == "auto":
            pc = "bjacobi" if name == "pcg" else "ilu0"
        precond = {"bjacobi": _lib.EFE_PRECOND_BLOCK_JACOBI, "ilu0": _lib.EFE_PRECOND_ILU0}[pc]
        info = self.system.linsolve(b, x, solver=solver, precond=precond, precision=prec, rtol=rtol,
                                    atol=o["krylovAtolFactor"] * fluxThreshold, max_iter=o["krylovMaxIter"],
                                    restart=o["gmresRestart"], check_every=o["krylovCheckEvery"])
        self.krylovIterations += info.iterations
        self.linearSolves += 1
        return info

    def _externalLoads(self, step, model, timeStep):
        """assembleLoads (nonlinearimplicitstatic.py:903-945): node forces, distributed loads
        (surface tractions, computeDistributedLoads :460-508) and body forces."""
        sysm = self.system
        PExt = torch.zeros(sysm.n_dof, dtype=torch.float64, device=sysm.device)
        for nf in step.actions.get("nodeforces", {}).values():
            nodes = model.nodeSets[nf.nSet]
            idx = torch.as_tensor(model.dofsOfNodeSet(nf.nSet), device=sysm.device)
            cur = np.asarray(nf.getCurrentLoad(timeStep), dtype=float)
            if cur.ndim == 1:  # one vector for every node of the set
                cur = np.tile(cur, (len(nodes), 1))
            load = torch.as_tensor(cur.ravel(), dtype=torch.float64, device=sysm.device)
            PExt[idx] += load
        tic = _time.perf_counter()
        for dl in step.actions.get("distributedload", {}).values():
            sysm.surface_traction(model.surfaces[dl.surface], dl.getCurrentLoad(timeStep), PExt)
        toc = _time.perf_counter()
        self.computationTimes["distributed loads"] += toc - tic
        for bf in step.actions.get("bodyforce", {}).values():  # computeBodyForces (:510-557)
            els = model.elementSets[bf.elSet]
            sysm.body_force(bf.getCurrentLoad(timeStep), PExt, None if len(els) == model.nElements else els)
        self.computationTimes["body forces"] += _time.perf_counter() - toc
        return PExt

    def _fluxThreshold(self, norms, iterationCounter):
        """max(tol * avgFlux, 1e-7) with avgFlux = max(1e-10, ||F||_1 / nAccumulatedFluxes)
        (nonlinearimplicitstatic.py:666-682, :786-790)."""
        avgFlux = max(1e-10, norms[3] / self._nAccFlux)
        tolFlux = FLUX_RESIDUAL_TOL if iterationCounter < 15 else FLUX_RESIDUAL_TOL_ALT
        return max(tolFlux * avgFlux, 1e-7)

    def _checkConvergence(self, n, ddU, iterationCounter, history):
        """checkConvergence for the displacement field (nonlinearimplicitstatic.py:625-725)."""
        fluxResidual, correction, nan = n[0], n[2], n[4]
        if nan:
            raise DivergingSolution("Obtained NaN in the residual")
        if ddU is None:
            correction = 0.0
        convergedCorrection = correction < FIELD_CORRECTION_TOL
        convergedFlux = fluxResidual <= self._fluxThreshold(n, iterationCounter)
        prev, nGrew = history
        if fluxResidual > prev:
            nGrew += 1
        return convergedCorrection and convergedFlux, (fluxResidual, nGrew)

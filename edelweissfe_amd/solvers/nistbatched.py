"""NISTBatched: many independent specimen models solved together on one GPU.

Each specimen is an ordinary model with its own step (time stepper, Dirichlet and node
force step actions) and is solved with the semantics of the reference NIST solver
(edelweissfe/solvers/nonlinearimplicitstatic.py:110-458): its own increments, Newton
iteration counter, convergence test, divergence count, cutbacks and state acceptance.
What is shared is the device work: all specimens form one block-diagonal system, every
Newton iteration is one element/assembly launch over all specimens, the convergence
quantities come from one per-specimen reduction, and the linear solves of all specimens
still iterating are one batched PCG with per-specimen scalars and tolerances.

Increments advance in lockstep rounds: in every round each unfinished specimen takes the
next increment of its own time stepper.  A specimen that has converged idles (its
correction is zero; the element passes skip it and keep its last state) until the slowest
specimen of the round has converged or failed; a failed specimen cuts back by its step's
cutback factor and retries in the next round.

A material cutback request (CutbackRequest, vonmises.py:230-231) is reported per specimen
(efe_get_specimen_status): that specimen's increment fails and is cut back by
max(0.5, cutbackFactor), as nonlinearimplicitstatic.py:247-249 does for one model.  A
specimen whose time stepper gives up (StepFailed, :317-319) stops with the state and U of
its last accepted increment and is listed in failedSpecimens; the others go on.
"""

import numpy as np
import torch

from .. import _lib
from ..device import DeviceSystem
from ..exceptions import ReachedMaxIncrements, ReachedMinIncrementSize

FIELD_CORRECTION_TOL = 1e-8      # phenomena.py fieldCorrectionTolerance["displacement"]
FLUX_RESIDUAL_TOL = 1e-8         # phenomena.py fluxResidualTolerance["displacement"]
FLUX_RESIDUAL_TOL_ALT = 5e-3     # phenomena.py fluxResidualToleranceAlternative["displacement"]

ITERATING, CONVERGED, FAILED, IDLE = 0, 1, 2, 3


class NISTBatched:
    identification = "NISTBatchedSolver"

    # krylovRtolMixed 1e-5: at 1e-4 the inexact FP32 solves of the 512-specimen batch leave the
    # next correction above the 1e-8 correction tolerance in the plastic increments of 507 of
    # 512 specimens (4230 specimen solves against FP64's 3455, stress error 1.5e-8); at 1e-5,
    # 150 specimens take one extra Newton iteration (3605 solves), U and stress stay within
    # 1e-8 of FP64 (3.8e-10, 2.0e-9), and the batch takes 2.5 % more time than at 1e-4; at 1e-6
    # every specimen takes FP64's Newton counts, for 8 % more time
    options_default = {"extrapolation": "linear", "precision": "fp64", "krylovAtolFactor": 1e-2,
                       "krylovRtol": 1e-13, "krylovRtolMixed": 1e-5, "krylovRtolMixedFirst": 1e-7,
                       "krylovMaxIter": 50000,
                       "krylovCheckEvery": 16, "autotune": True, "newtonOnDevice": True}

    def __init__(self, **kwargs):
        self.options = dict(self.options_default)
        for k, v in kwargs.items():
            if k not in self.options:
                raise AttributeError("Invalid option {:} for {:}".format(k, self.identification))
            self.options[k] = type(self.options_default[k])(v)
        self.system = None
        self.tuning = None
        self.linearSolves = 0
        self.elementPasses = 0
        self.krylovIterations = 0
        self.rounds = 0
        self.specimenPasses = self.specimenSolves = self.specimenKrylov = None

    @property
    def _mixed(self):
        return self.options["precision"].lower() == "mixed"

    # ---- assembling the batch --------------------------------------------------------------------------
    def build(self, models):
        """Creates the block-diagonal device system of `models` (one element type, one material)."""
        et = models[0].elementType
        name = models[0].material["name"]
        if any(m.elementType != et or m.material["name"].lower() != name.lower() for m in models):
            raise ValueError("a batch holds one element type and one material")
        coords, conn, props, spec = [], [], [], []
        self.nodeOffset = [0]
        self.elementOffset = [0]
        for s, m in enumerate(models):
            coords.append(m.coordinates)
            conn.append(m.connectivity + self.nodeOffset[-1])
            p = np.asarray(m.material["properties"], dtype=np.float64)
            props.append(np.broadcast_to(p, (m.nElements, p.shape[-1])))
            spec.append(np.full(m.nNodes, s, dtype=np.int32))
            self.nodeOffset.append(self.nodeOffset[-1] + m.nNodes)
            self.elementOffset.append(self.elementOffset[-1] + m.nElements)
        self.system = DeviceSystem(np.vstack(coords), np.vstack(conn), et, name, np.vstack(props),
                                   specimen_of_node=np.concatenate(spec), n_specimens=len(models))
        self.dofOffset = [3 * o for o in self.nodeOffset]
        if self.options["autotune"]:
            from .. import autotune

            self.tuning = autotune.tune(self.system, _lib.EFE_FP32_MATFREE if self._mixed else _lib.EFE_FP64)
        return self.system

    def attach(self, other):
        """Reuse the device system (and batch layout) another NISTBatched built for the same models."""
        self.system = other.system
        self.nodeOffset, self.elementOffset, self.dofOffset = other.nodeOffset, other.elementOffset, other.dofOffset
        self.tuning = other.tuning

    # ---- public interface ------------------------------------------------------------------------------
    def solveSteps(self, models, steps):
        """Solves steps[s] on models[s] for every specimen s; model.U and model.time are
        updated.  A specimen whose increment control gives up (ReachedMaxIncrements /
        ReachedMinIncrementSize, i.e. StepFailed of nonlinearimplicitstatic.py:330-331 for
        that model) stops there while the others go on; its index is listed in
        self.failedSpecimens.  Returns (True if no specimen failed, models)."""
        S = len(models)
        if len(steps) != S:
            raise ValueError("one step per model")
        sysm = self.system if self.system is not None else self.build(models)
        if sysm.n_specimens != S:
            raise ValueError("the batch was built for {} specimens".format(sysm.n_specimens))
        dev = sysm.device
        D = sysm.n_dof
        prec = _lib.EFE_FP32_MATFREE if self._mixed else _lib.EFE_FP64
        rtol = self.options["krylovRtolMixed"] if self._mixed else self.options["krylovRtol"]
        nn = 8 if models[0].elementType == "C3D8" else 20
        nAcc = np.array([m.nElements * nn * 3 for m in models], dtype=np.float64)

        # Dirichlet DOFs in the batch numbering; R[idx] = delta is applied action by action in
        # the reference (:616-618), so a DOF named twice takes the last action's value
        owners = []
        hasDir = np.zeros(S, dtype=bool)
        for s, (m, st) in enumerate(zip(models, steps)):
            if st.actions.get("bodyforce") or st.actions.get("distributedload"):
                raise NotImplementedError("body forces and distributed loads in a batch")
            for d in st.actions["dirichlet"].values():
                owners.append((s, d, d.indices(m) + self.dofOffset[s]))
                hasDir[s] = True
        allDir = np.unique(np.concatenate([ix for _, _, ix in owners])) if owners else np.zeros(0, np.int64)
        sysm.set_dirichlet(allDir)
        dirT = torch.as_tensor(allDir, dtype=torch.int64, device=dev)
        # the prescribed increments as a device table: DOF, value over the step and owning
        # action of every prescribed DOF, the last action naming a DOF winning; per round only
        # the actions' amplitude factors travel to the device
        if owners:
            ix = np.concatenate([o[2] for o in owners])
            vals = np.concatenate([np.tile(d.values, len(i) // len(d.values)) for _, d, i in owners])
            own = np.concatenate([np.full(len(o[2]), k, dtype=np.int64) for k, o in enumerate(owners)])
            _, firstInRev = np.unique(ix[::-1], return_index=True)
            pick = np.sort(ix.size - 1 - firstInRev)
            self._dirTable = (torch.as_tensor(ix[pick], device=dev), torch.as_tensor(vals[pick], device=dev),
                              torch.as_tensor(own[pick], device=dev))
        else:
            self._dirTable = None
        isDir = torch.zeros(D, dtype=torch.bool, device=dev)
        isDir[dirT] = True
        specOfDof = torch.as_tensor(np.repeat(np.arange(S), np.diff(self.dofOffset)), device=dev)

        def perDof(a):
            return torch.as_tensor(np.asarray(a), device=dev)[specOfDof]

        gens = [st.getTimeStep() for st in steps]
        finished = np.zeros(S, dtype=bool)
        self.failedSpecimens = []
        prevTS = [None] * S
        U = torch.as_tensor(np.concatenate([m.U for m in models]), dtype=torch.float64, device=dev)
        dU = torch.zeros(D, dtype=torch.float64, device=dev)
        P = torch.zeros(D, dtype=torch.float64, device=dev)
        F = torch.zeros(D, dtype=torch.float64, device=dev)
        self.history = [[] for _ in range(S)]
        # per specimen, why each failed increment failed: "cutback" (material request), "nan",
        # "iterations" (maxIter or maxGrowIter) or "krylov" (linear solve diverged)
        self.failures = [[] for _ in range(S)]
        # per-specimen work counts: element passes, linear solves and Krylov iterations the
        # specimen would need on its own
        self.specimenPasses = np.zeros(S, dtype=np.int64)
        self.specimenSolves = np.zeros(S, dtype=np.int64)
        self.specimenKrylov = np.zeros(S, dtype=np.int64)
        try:
            while True:
                ts = [None] * S
                for s in range(S):
                    if not finished[s]:
                        try:
                            ts[s] = next(gens[s], None)
                        except (ReachedMaxIncrements, ReachedMinIncrementSize):
                            ts[s] = None
                            self.failedSpecimens.append(s)
                        finished[s] = ts[s] is None
                live = ~finished
                if not live.any():
                    break
                self.rounds += 1
                dU, PExt, deltaT, extrap = self._startIncrements(models, steps, ts, prevTS, live, owners, dU, isDir,
                                                                 perDof)
                self._cutbackSize = np.zeros(S)
                if self.options["newtonOnDevice"]:
                    state, it = self._newtonRoundDevice(sysm, steps, live, extrap, hasDir, dU, PExt, deltaT, nAcc,
                                                        prec, rtol)
                else:
                    state, it = self._newtonRound(sysm, steps, live, extrap, hasDir, dU, P, F, PExt, deltaT, dirT,
                                                  perDof, nAcc, prec, rtol)
                dU = self._lastDU
                ok = state == CONVERGED
                sysm.accept_state_batched(ok)
                U = torch.where(perDof(ok), U + dU, U)
                for s in np.flatnonzero(live):
                    if ok[s]:
                        prevTS[s] = ts[s]
                        if it[s] >= steps[s].criticalIter:
                            steps[s].preventIncrementIncrease()
                        models[s].time = ts[s].totalTime
                        self.history[s].append(int(it[s]))
                    else:
                        prevTS[s] = None
                        try:
                            steps[s].discardAndChangeIncrement(max(self._cutbackSize[s], steps[s].cutbackFactor))
                        except (ReachedMaxIncrements, ReachedMinIncrementSize):
                            finished[s] = True
                            self.failedSpecimens.append(s)
                            continue
                        self.history[s].append(np.inf)
                dU = torch.where(perDof(~ok), torch.zeros_like(dU), dU)
        finally:
            sysm.set_active_specimens(None)  # later passes (autotuner, probes) see every specimen
        for s, st in enumerate(steps):
            for group in ("dirichlet", "nodeforces"):
                for action in st.actions.get(group, {}).values():
                    action.applyAtStepEnd(models[s])
        if dev.type == "cuda":
            torch.cuda.synchronize(dev)
        Uh = U.cpu().numpy()
        for s, m in enumerate(models):
            m.U = Uh[self.dofOffset[s]:self.dofOffset[s + 1]].copy()
        return not self.failedSpecimens, models

    # ---- one lockstep round ------------------------------------------------------------------------------
    def _startIncrements(self, models, steps, ts, prevTS, live, owners, dU, isDir, perDof):
        """Linear extrapolation (nonlinearimplicitstatic.py:947-986), increments of the
        prescribed values and the external node forces (:903-945) of every live specimen."""
        S = len(models)
        D = dU.numel()
        dev = dU.device
        scale = np.zeros(S)
        extrap = np.zeros(S, dtype=bool)
        for s in np.flatnonzero(live):
            p = prevTS[s]
            if self.options["extrapolation"] == "linear" and p is not None and p.timeIncrement:
                scale[s] = ts[s].stepProgressIncrement / p.stepProgressIncrement
                extrap[s] = True
        deltaT = torch.zeros(D, dtype=torch.float64, device=dev)
        if self._dirTable is not None:
            factor = np.array([d.incrementFactor(ts[s]) if live[s] else 0.0 for s, d, _ in owners])
            ix, vals, own = self._dirTable
            deltaT[ix] = vals * torch.as_tensor(factor, device=dev)[own]
        dU = dU * perDof(scale)
        dU = torch.where(perDof(extrap) & isDir, deltaT, dU)
        PExt = None
        for s in np.flatnonzero(live):
            for nf in steps[s].actions.get("nodeforces", {}).values():
                if PExt is None:
                    PExt = np.zeros(D)
                nodes = np.asarray(models[s].nodeSets[nf.nSet]) + self.nodeOffset[s]
                PExt[(3 * nodes[:, None] + np.arange(3)[None, :]).ravel()] += np.tile(nf.getCurrentLoad(ts[s]),
                                                                                        len(nodes))
        if PExt is not None:
            PExt = torch.as_tensor(PExt, device=dev)
        return dU, PExt, deltaT, extrap

    def _newtonRound(self, sysm, steps, live, extrap, hasDir, dU, P, F, PExt, deltaT, dirT, perDof, nAcc, prec, rtol):
        """Newton iterations of all live specimens until each has converged or failed
        (nonlinearimplicitstatic.py:334-458 per specimen)."""
        S = live.size
        D = dU.numel()
        dev = dU.device
        o = self.options
        it = np.zeros(S, dtype=int)
        grow = [(0.0, 0) for _ in range(S)]
        state = np.where(live, ITERATING, IDLE)
        ddU = None
        while True:
            # a specimen failed in its first pass keeps it == 0 but takes no further pass: only
            # iterating specimens are first, so the passes after the first form no correction
            first = live & (it == 0) & ~extrap & hasDir & (state == ITERATING)
            if first.any():
                dirVals = torch.where(perDof(first), deltaT, torch.zeros_like(deltaT))
                corr = torch.empty(D, dtype=torch.float64, device=dev)
            else:
                dirVals = corr = None
            # only the specimens still iterating are computed; the others keep their last pass
            sysm.set_active_specimens(state == ITERATING)
            st, _ = sysm.compute_elements(dU, P, F, dir_values=dirVals, dir_correction=corr, precision=prec)
            self.elementPasses += 1
            self.specimenPasses += state == ITERATING
            if st == _lib.EFE_CUTBACK:
                # only the specimens still iterating count: the others keep results of a finished pass
                asked = sysm.specimen_status() & (state == ITERATING)
                state[asked] = FAILED
                for s in np.flatnonzero(asked):
                    self.failures[s].append("cutback")
                self._cutbackSize[asked] = 0.5
            R = P + PExt if PExt is not None else P.clone()
            if dirT.numel():
                R[dirT] = dirVals[dirT] if dirVals is not None else 0.0
            nrm = sysm.norms_batched(R, ddU, F)
            avg = np.maximum(1e-10, nrm[:, 2] / nAcc)
            tolF = np.where(it < 15, FLUX_RESIDUAL_TOL, FLUX_RESIDUAL_TOL_ALT)
            fluxThreshold = np.maximum(tolF * avg, 1e-7)
            for s in np.flatnonzero((state == ITERATING) & ~first):
                flux = nrm[s, 0]
                correction = nrm[s, 1] if ddU is not None else 0.0
                prev, nGrew = grow[s]
                grow[s] = (flux, nGrew + (1 if flux > prev else 0))
                if nrm[s, 3]:
                    state[s] = FAILED
                    self.failures[s].append("nan")
                elif correction < FIELD_CORRECTION_TOL and flux <= fluxThreshold[s]:
                    state[s] = CONVERGED
                elif grow[s][1] > steps[s].maxGrowIter or it[s] == steps[s].maxIter:
                    state[s] = FAILED
                    self.failures[s].append("iterations")
            go = state == ITERATING
            if not go.any():
                break
            b = R - corr if corr is not None else R
            ddU = torch.empty(D, dtype=torch.float64, device=dev)
            # mixed form: a specimen's FP32 solve of its prescribed increment runs to
            # krylovRtolMixedFirst, as in NISTGPU (its next correction then passes 1e-8)
            specRtol = np.where(first, o["krylovRtolMixedFirst"], rtol) if self._mixed else None
            info, its = sysm.linsolve_batched(b, ddU, o["krylovAtolFactor"] * fluxThreshold, go.astype(np.uint8),
                                              spec_rtol=specRtol, precision=prec, rtol=rtol,
                                              max_iter=o["krylovMaxIter"],
                                              check_every=o["krylovCheckEvery"])
            self.linearSolves += int(go.sum())
            self.krylovIterations += int(its[go].sum())
            self.specimenSolves += go
            self.specimenKrylov += np.where(go, its, 0)
            # a Krylov solve that did not converge or produced NaN fails its specimen
            # (DivergingSolution, :748-749); its correction is discarded below
            bad = go & (its >= o["krylovMaxIter"])
            if info.status == _lib.EFE_NAN:
                bad |= go & ~np.isfinite(sysm.norms_batched(ddU, None, ddU)[:, 0])
            state[bad] = FAILED
            for s in np.flatnonzero(bad):
                self.failures[s].append("krylov")
            dU = dU + ddU
            it[go] += 1
        self._lastDU = dU
        return state, it

    _REASONS = {_lib.EFE_FAIL_CUTBACK: "cutback", _lib.EFE_FAIL_NAN: "nan", _lib.EFE_FAIL_ITERATIONS: "iterations",
                _lib.EFE_FAIL_KRYLOV: "krylov"}

    def _newtonRoundDevice(self, sysm, steps, live, extrap, hasDir, dU, PExt, deltaT, nAcc, prec, rtol):
        """_newtonRound with every decision on the device (efe_newton_round_batched): one
        replayed hipGraph per Newton pass, the same per-specimen arithmetic and outcomes."""
        o = self.options
        opts = _lib.NewtonOpts(precision=prec, krylov_max_iter=o["krylovMaxIter"], check_every=o["krylovCheckEvery"],
                               mixed_first=1 if self._mixed else 0, alt_after=15, rtol=rtol,
                               rtol_first=o["krylovRtolMixedFirst"], atol_factor=o["krylovAtolFactor"],
                               flux_tol=FLUX_RESIDUAL_TOL, flux_tol_alt=FLUX_RESIDUAL_TOL_ALT,
                               correction_tol=FIELD_CORRECTION_TOL, flux_floor=1e-7, avg_floor=1e-10)
        maxIter = [st.maxIter for st in steps]
        maxGrow = [st.maxGrowIter for st in steps]
        state, it, reason, work = sysm.newton_round_batched(dU, PExt, deltaT, live, extrap, hasDir, maxIter, maxGrow,
                                                            nAcc, opts)
        self.elementPasses += int(work[:, 0].max()) if live.any() else 0
        self.linearSolves += int(work[:, 1].sum())
        self.krylovIterations += int(work[:, 2].sum())
        self.specimenPasses += work[:, 0]
        self.specimenSolves += work[:, 1]
        self.specimenKrylov += work[:, 2]
        for s in np.flatnonzero(state == FAILED):
            self.failures[s].append(self._REASONS[int(reason[s])])
        self._cutbackSize[reason == _lib.EFE_FAIL_CUTBACK] = 0.5
        self._lastDU = dU
        return state, it.astype(int)

    def qpStates(self, models, which=0):
        """Reference (0) or temporary (1) QP state of every specimen, as a list of
        [n_elements][n_qp][n_state] arrays."""
        st = self.system.qp_state(which)
        return [st[self.elementOffset[s]:self.elementOffset[s + 1]] for s in range(len(models))]


class NISTBatchedGroups:
    """Specimens of several element and material types, grouped by (element type, material):
    every group is one NISTBatched, i.e. one block-diagonal context with one element launch
    per Newton iteration and one batched PCG, and the groups run one after the other on the
    GPU.  Each specimen is solved exactly as in a batch of its own group; per-specimen results
    (history, failures, work counts, failedSpecimens) come back in the order of the models."""

    identification = "NISTBatchedGroupsSolver"

    def __init__(self, **kwargs):
        NISTBatched(**kwargs)  # rejects unknown options now rather than at the first group
        self.kwargs = kwargs
        self.groups = []  # (key, specimen indices, NISTBatched)

    @staticmethod
    def groupKey(model):
        return model.elementType.upper(), model.material["name"].lower()

    def solveSteps(self, models, steps):
        """As NISTBatched.solveSteps, for specimens of any element types and materials."""
        S = len(models)
        if len(steps) != S:
            raise ValueError("one step per model")
        keys = [self.groupKey(m) for m in models]
        self.groups = []
        self.history, self.failures = [None] * S, [None] * S
        self.specimenPasses = np.zeros(S, dtype=np.int64)
        self.specimenSolves = np.zeros(S, dtype=np.int64)
        self.specimenKrylov = np.zeros(S, dtype=np.int64)
        self.failedSpecimens = []
        allOk = True
        for key in dict.fromkeys(keys):  # groups in order of first appearance
            idx = [i for i, k in enumerate(keys) if k == key]
            solver = NISTBatched(**self.kwargs)
            ok, _ = solver.solveSteps([models[i] for i in idx], [steps[i] for i in idx])
            allOk = allOk and ok
            for j, i in enumerate(idx):
                self.history[i] = solver.history[j]
                self.failures[i] = solver.failures[j]
                self.specimenPasses[i] = solver.specimenPasses[j]
                self.specimenSolves[i] = solver.specimenSolves[j]
                self.specimenKrylov[i] = solver.specimenKrylov[j]
            self.failedSpecimens += [idx[j] for j in solver.failedSpecimens]
            self.groups.append((key, idx, solver))
        self.failedSpecimens.sort()
        return allOk, models

    def qpStates(self, models, which=0):
        """QP state of every specimen, in the order of the models."""
        out = [None] * len(models)
        for _, idx, solver in self.groups:
            for j, st in enumerate(solver.qpStates([models[i] for i in idx], which)):
                out[idx[j]] = st
        return out

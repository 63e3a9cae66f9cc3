"""Benchmark of the implicit static Newton step on one MI355X.

Workload (the FP64 GPU configuration of the headline metric): 64 x 64 x 64 C3D8
unit cube (262,144 elements, 823,875 DOFs), interior nodes jittered +-10 % of h
(seed 0), J2 von Mises with linear isotropic hardening (E 210000, nu 0.3,
sigma_y 250, H 1000), uniaxial displacement-controlled compression to 1 %
strain along y in 20 increments (maxInc 0.05), symmetry BCs on x = 0 and z = 0,
block-Jacobi PCG.  One "step" is the whole load history (zero increment + 20
increments).  The same history is run in the mixed-precision form (FP32
stiffness contractions and FP32 PCG inside the FP64 Newton refinement) and its
error against the FP64 run is reported per field.

Prints one JSON line.  `value` is FP64 Newton iterations (linear solves) per
second over the timed load histories; inputs (mesh, pattern, state) are
resident in HBM before the timed region starts.  Unless --no-others is given,
the default run then measures the three other workloads (96^3 lognormal J2, the
512-specimen batch, the 48^3 hexa20R footing), one child process each after
the headline's own timing, and carries their lines under "other_workloads".
"""

import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Newton iters/s + DOF-updates/s, mixed vs FP64 GPU at 1e-8 rel err"
HBM_PEAK_GBS = 8000.0
J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def lognormal_props(n_elements, seed=1, median=250.0, sigma_log=0.1):
    """J2 properties per element with sigma_y lognormal (median 250, sigma_log 0.1, seed 1)."""
    from edelweissfe_amd.utils import rng

    props = np.tile(np.asarray(J2), (n_elements, 1))
    props[:, 2] = rng.lognormal(seed, n_elements, median, sigma_log)
    return props


def build_case(n, increments, strain, lognormal=False):
    from edelweissfe_amd.generators.boxgen import generateBox
    from edelweissfe_amd.steps import AdaptiveStep, Dirichlet

    model = generateBox(n, n, n, 1.0, 1.0, 1.0, elType="C3D8", jitter=0.1, seed=0)
    model.assignMaterial("vonmises", lognormal_props(model.nElements) if lognormal else J2)

    def step_for(m):
        acts = {"dirichlet": {
            "top": Dirichlet("top", "gen_top", {1: -strain}),
            "bottom": Dirichlet("bottom", "gen_bottom", {1: 0.0}),
            "left": Dirichlet("left", "gen_left", {0: 0.0}),
            "back": Dirichlet("back", "gen_back", {2: 0.0})}}
        return AdaptiveStep(1, m, acts, maxInc=1.0 / increments, minInc=1e-6, maxNumInc=1000, maxIter=15)

    return model, step_for


def fresh_model(template):
    from edelweissfe_amd.model import FEModel

    m = FEModel(template.coordinates, template.elementType, template.connectivity, template.nodeSets,
                template.elementSets)
    m.material = template.material
    return m


def run_history(system, template, step_for, precision, steps, warmup):
    import torch

    from edelweissfe_amd.solvers import NISTGPU

    times, solvers, models = [], [], []
    for k in range(warmup + steps):
        m = fresh_model(template)
        s = NISTGPU(precision=precision)
        s.attach(system)
        st = step_for(m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.solveStep(st, m)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        iters = [r["iters"] for r in s.history]
        log("[bench] %s %s history: %.3f s, Newton %s, solves %d, Krylov its %d" % (
            precision, "warmup" if k < warmup else "timed", dt, iters, s.linearSolves, s.krylovIterations))
        if k >= warmup:
            times.append(dt)
            solvers.append(s)
            models.append(m)
    return times, solvers, models


def spmv_roofline(system, n_launch=50):
    """Average device duration of the FP64 sliced-ELL SpMV (the dominant kernel of the FP64 PCG)
    measured with HIP events on the stream it is launched on."""
    import torch

    x = torch.randn(system.n_dof, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        system.spmv(x, y)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(n_launch):
        system.spmv(x, y)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / n_launch
    nb = system.nnz // 9
    # algorithmic bytes: 3x3 block values (8 B each) + one int32 column index per block + x read once + y written once
    alg = 8 * system.nnz + 4 * nb + 8 * system.n_dof + 8 * system.n_dof
    return ms, alg


def csr_model_bytes(system):
    """Bytes of the same product in scalar CSR (8 B value + 4 B column index per nonzero, the
    row pointers, x and y once): the per-unit model of a CSR SpMV, which the 3x3 block layout
    undercuts by storing one column index per block."""
    return 12 * system.nnz + 4 * (system.n_dof + 1) + 16 * system.n_dof


def profile_spmv(system, n_launch=20, copy_mb=512):
    """Counter-collection mode: one FP64 element pass and assembly of the operator, then
    n_launch FP64 SpMV launches, n_launch plain device copies of copy_mb MiB (the calibration
    kernel whose FETCH_SIZE / WRITE_SIZE per launch are known: copy_mb MiB each), then one
    matrix-free FP32 element pass and a block-Jacobi PCG of n_launch iterations on the
    matrix-free operator (the mixed form's inner solve)."""
    import torch

    from edelweissfe_amd import _lib

    dU = torch.zeros(system.n_dof, dtype=torch.float64, device="cuda")
    P, F = system.zeros(), system.zeros()
    system.compute_elements(dU, P, F)
    x = torch.randn(system.n_dof, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    a = torch.empty(copy_mb * 2 ** 20 // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(n_launch):
        system.spmv(x, y)
    for _ in range(n_launch):
        b.copy_(a)
    system.compute_elements(dU, P, F, precision=_lib.EFE_FP32_MATFREE)
    system.linsolve(x, y, precision=_lib.EFE_FP32_MATFREE, rtol=1e-30, max_iter=n_launch, check_every=n_launch)
    torch.cuda.synchronize()
    nb = system.nnz // 9
    alg = 8 * system.nnz + 4 * nb + 16 * system.n_dof
    log("[profile] %d spmv launches (algorithmic %d bytes each), %d copies of %d MiB, %d matrix-free PCG "
        "iterations" % (n_launch, alg, n_launch, copy_mb, n_launch))


def profile_elements(system, n_pass=3, n_spmv=5):
    """Counter-collection mode of the footing workload: n_pass FP64 and n_pass assembled FP32
    element passes (with assembly) of the C3D20R Drucker-Prager operator at a random increment
    of +-1e-3 per DOF (seed 3: plastic QPs on the cone and at the apex), then n_spmv FP64 SpMVs."""
    import torch

    from edelweissfe_amd import _lib

    g = torch.Generator(device="cpu").manual_seed(3)
    dU = (torch.rand(system.n_dof, generator=g, dtype=torch.float64) * 2e-3 - 1e-3).cuda()
    P, F = system.zeros(), system.zeros()
    for prec in (_lib.EFE_FP64, _lib.EFE_FP32):
        for _ in range(n_pass):
            system.compute_elements(dU, P, F, precision=prec)
    x = torch.randn(system.n_dof, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    system.compute_elements(dU, P, F)
    for _ in range(n_spmv):
        system.spmv(x, y)
    torch.cuda.synchronize()
    log("[profile] %d FP64 and %d FP32 element passes with assembly, %d FP64 SpMVs" % (n_pass, n_pass, n_spmv))


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def reset_batch(solver, count, strain, increments):
    from edelweissfe_amd.generators.specimens import specimenBatch

    models, steps, par = specimenBatch(count, seed=2, strain=strain, increments=increments)
    if solver.system is not None:
        solver.system.reset_state()
    return models, steps, par


def run_batched(solver, count, strain, increments, steps, warmup):
    import torch

    times, models = [], None
    for k in range(warmup + steps):
        models, st, par = reset_batch(solver, count, strain, increments)
        if solver.system is None:
            t0 = time.perf_counter()
            solver.build(models)
            torch.cuda.synchronize()
            log("[bench] batch system: %d specimens, %d DOFs, %d nnz, tuned %s (%.1f s)" % (
                count, solver.system.n_dof, solver.system.nnz, solver.tuning and {
                    k2: v for k2, v in solver.tuning.items() if k2 != "timings"}, time.perf_counter() - t0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        solver.solveSteps(models, st)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log("[bench] batched %s %s: %.3f s, %d rounds, %d element passes, %d specimen solves, %d Krylov its" % (
            solver.options["precision"], "warmup" if k < warmup else "timed", dt, solver.rounds,
            solver.elementPasses, solver.linearSolves, solver.krylovIterations))
        solver.lastRun = {"rounds": solver.rounds, "element_passes": solver.elementPasses,
                          "krylov_iterations": solver.krylovIterations}
        solver.rounds = solver.elementPasses = solver.linearSolves = solver.krylovIterations = 0
        if k >= warmup:
            times.append(dt)
    return times, models, par


def bench_batched(args):
    """Batched form: 512 independent J2 cube specimens n^3, n uniform in {8..24}, E uniform in
    [150000, 250000], sigma_y uniform in [150, 350] (seed 2), 1 % uniaxial compression in 10
    increments each, FP64 and mixed."""
    import torch

    from edelweissfe_amd import autotune, _lib
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    count, inc = args.specimens, args.batch_increments
    s64 = NISTBatched(precision="fp64")
    t64, m64, par = run_batched(s64, count, args.strain, inc, args.steps, args.warmup)
    sysm = s64.system
    D = np.diff(s64.dofOffset)
    solves64 = s64.specimenSolves.copy()
    hist64 = [list(h) for h in s64.history]
    counts64 = (s64.specimenPasses.copy(), s64.specimenSolves.copy(), s64.specimenKrylov.copy())
    q64 = sysm.qp_state(0)
    spmv_ms, spmv_alg = spmv_roofline(sysm)
    log("[bench] batch SpMV %.4f ms, %.1f GB/s algorithmic" % (spmv_ms, spmv_alg / spmv_ms / 1e6))
    wall64 = sum(t64) / len(t64)
    result = {
        "metric": METRIC,
        "value": float(solves64.sum()) / wall64,
        "unit": "Newton iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d independent hexa8 C3D8 J2 unit-cube specimens n^3, n uniform in {8..24}, "
                               "E uniform in [150000, 250000], sigma_y uniform in [150, 350] (seed 2), 10%% jitter, "
                               "uniaxial compression to %g strain in %d increments each, batched block-Jacobi PCG"
                               % (count, args.strain, inc),
                   "specimens": count, "elements": int(sysm.n_elements), "dofs": int(sysm.n_dof),
                   "csr_nnz": int(sysm.nnz)},
        "note": "value counts specimen Newton iterations (linear solves) per second over the whole batch",
        "dof_updates_per_s": float((solves64 * D).sum()) / wall64,
        "batch_element_passes": s64.lastRun["element_passes"],
        "lockstep_rounds": s64.lastRun["rounds"],
        "batch_krylov_iterations": s64.lastRun["krylov_iterations"],
        "specimen_newton_iterations": int(solves64.sum()),
        "autotune": {k: v for k, v in (s64.tuning or {}).items() if k != "timings"},
    }
    # the same FP64 batch with the Newton decisions on the host (one readback and ~15 launches
    # per pass) against the device-resident rounds timed above
    sh = NISTBatched(precision="fp64", newtonOnDevice=False)
    sh.attach(s64)
    th, mh, _ = run_batched(sh, count, args.strain, inc, args.steps, args.warmup)
    wallh = sum(th) / len(th)
    result["host_newton_rounds"] = {
        "ms_per_batch_history": 1000.0 * wallh,
        "device_round_speedup": wallh / wall64,
        "same_results_as_device_rounds": bool([list(h) for h in sh.history] == hist64 and
                                              all(np.array_equal(a.U, b.U) for a, b in zip(mh, m64))),
    }
    log("[bench] batched fp64, host Newton rounds: %.3f s (device rounds %.3f s)" % (wallh, wall64))
    if not args.no_mixed:
        s32 = NISTBatched(precision="mixed")
        s32.attach(s64)
        s32.tuning = autotune.tune(sysm, _lib.EFE_FP32_MATFREE)
        t32, m32, _ = run_batched(s32, count, args.strain, inc, args.steps, args.warmup)
        wall32 = sum(t32) / len(t32)
        q32 = sysm.qp_state(0)
        uerr = [float(np.abs(a.U - b.U).max() / np.abs(b.U).max()) for a, b in zip(m32, m64)]
        result["mixed"] = {
            "newton_iterations_per_s": float(s32.specimenSolves.sum()) / wall32,
            "dof_updates_per_s": float((s32.specimenSolves * D).sum()) / wall32,
            "ms_per_batch_history": 1000.0 * wall32,
            "history_speedup_vs_fp64": wall64 / wall32,
            "same_newton_counts_as_fp64": [list(h) for h in s32.history] == hist64,
            "autotune": {k: v for k, v in (s32.tuning or {}).items() if k != "timings"},
            "error_vs_fp64": {
                "U_rel_inf_max_over_specimens": max(uerr),
                "stress_rel_inf": float(np.abs(q32[..., 0:6] - q64[..., 0:6]).max() / np.abs(q64[..., 0:6]).max()),
            },
        }
        log("[bench] batched mixed: %.3f s (%.2fx), U rel err %.2e" % (wall32, wall64 / wall32, max(uerr)))
    result["roofline"] = {
        "kernel": "spmv_kernel<double> (sliced-ELL 3x3 blocks) on the block-diagonal batch operator",
        "bound": "hbm",
        "achieved": spmv_alg / (spmv_ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": spmv_alg / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": None,
        "algorithmic_bytes": spmv_alg,
        "avg_launch_ms": spmv_ms,
        "csr_model_bytes": csr_model_bytes(sysm),
        "csr_model_achieved": csr_model_bytes(sysm) / (spmv_ms * 1e-3) / 1e9,
    }
    if not args.no_cpu:
        try:
            from oracle import cpu_baseline

            log("[bench] CPU baseline (bounded sample)")
            result["cpu_baseline"] = cpu_baseline.measure_batched(par["n"], *counts64, J2, args.strain, inc,
                                                                  budget_s=args.cpu_seconds, log=log)
        except Exception as e:
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


DP = [30000.0, 0.2, 30.0, 10.0, 10.0, 0.0]
GEO = dict(p1=0.0, h1=1.0, p2=-10.0, h2=0.0, xLateral=0.5, zLateral=0.5)


def footing_case(n, increments, settlement):
    """n^3 C3D20R unit cube (10 % jitter, seed 0), non-associative Drucker-Prager (friction 30,
    dilation 10, cohesion 10, E 30000, nu 0.2), rollers on the sides and the bottom.  Step 1:
    geostatic prestress sigma_yy = -10 (1 - y), K0 = 0.5, with the matching gravity
    b_y = -10 (displacements reset at its end).  Step 2: rigid footing, u_y = -settlement on
    the top-face nodes with |x - 1/2| <= 1/4 and |z - 1/2| <= 1/4, in `increments` increments."""
    from edelweissfe_amd.generators.boxgen import generateBox
    from edelweissfe_amd.steps import AdaptiveStep, BodyForce, Dirichlet, Geostatic

    model = generateBox(n, n, n, 1.0, 1.0, 1.0, elType="C3D20R", jitter=0.1, seed=0)
    model.assignMaterial("druckerprager", DP)
    top = model.nodeSets["gen_top"]
    x, z = model.coordinates[top, 0], model.coordinates[top, 2]
    model.nodeSets["footing"] = top[(np.abs(x - 0.5) <= 0.25 + 1e-12) & (np.abs(z - 0.5) <= 0.25 + 1e-12)]
    rollers = [("bottom", "gen_bottom", {1: 0.0}), ("left", "gen_left", {0: 0.0}), ("right", "gen_right", {0: 0.0}),
               ("back", "gen_back", {2: 0.0}), ("front", "gen_front", {2: 0.0})]
    slope = (GEO["p2"] - GEO["p1"]) / (GEO["h2"] - GEO["h1"])

    def steps_for(m):
        grav = BodyForce("gravity", "all", (0.0, -slope, 0.0), amplitude=lambda t: 1.0)
        a1 = {"dirichlet": {nm: Dirichlet(nm, ns, v) for nm, ns, v in rollers}, "bodyforce": {"gravity": grav},
              "geostatic": {"geo": Geostatic("geo", **GEO)}}
        a2 = {"dirichlet": {nm: Dirichlet(nm, ns, v) for nm, ns, v in
                            rollers + [("footing", "footing", {1: -settlement})]},
              "bodyforce": {"gravity": grav}}
        kw = dict(minInc=1e-6, maxNumInc=1000, maxIter=25)
        return [lambda: AdaptiveStep(1, m, a1, maxInc=1.0, **kw),
                lambda: AdaptiveStep(2, m, a2, maxInc=1.0 / increments, **kw)]

    return model, steps_for


def phase_avg(before, after):
    """History averages of the device phase times (ms per call) between two phase_totals
    snapshots: element kernel and assembly per tangent pass, residual-only pass, Krylov
    iteration (CG or GMRES), ILU(0) factorisation per solve, norms per call; and the counts."""
    d = {k: after[k] - before[k] for k in after if k != "avg_ms"}

    def avg(a, n):
        return d[a] / d[n] if d[n] else 0.0

    return {"elements_per_tangent_pass": avg("elements_ms", "element_passes"),
            "assembly_per_tangent_pass": avg("assembly_ms", "element_passes"),
            "residual_only_pass": avg("residual_ms", "residual_passes"),
            "krylov_iteration": avg("krylov_ms", "krylov_iterations"),
            "ilu_factor_per_solve": avg("ilu_factor_ms", "solves"),
            "norms": avg("norms_ms", "norm_calls"),
            "tangent_passes": int(d["element_passes"]), "residual_only_passes": int(d["residual_passes"]),
            "solves": int(d["solves"]), "krylov_iterations": int(d["krylov_iterations"])}


def run_footing(system, template, steps_for, precision, steps, warmup):
    import torch

    from edelweissfe_amd.solvers import NISTGPU

    times, solvers, models = [], [], []
    for k in range(warmup + steps):
        m = fresh_model(template)
        s = NISTGPU(precision=precision, linsolver="gmres", preconditioner="ilu0", gmresRestart=50)
        s.attach(system)
        makers = steps_for(m)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for make in makers:
            s.solveStep(make(), m)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        log("[bench] footing %s %s: %.3f s, Newton %s, solves %d, GMRES its %d" % (
            precision, "warmup" if k < warmup else "timed", dt, [r["iters"] for r in s.history], s.linearSolves,
            s.krylovIterations))
        if k >= warmup:
            times.append(dt)
            solvers.append(s)
            models.append(m)
    return times, solvers, models


def bench_footing(args):
    import torch

    from edelweissfe_amd.device import DeviceSystem

    n, inc, settle = args.n, args.increments, args.settlement
    log("[bench] building %d^3 C3D20R Drucker-Prager footing model" % n)
    template, steps_for = footing_case(n, inc, settle)
    t0 = time.perf_counter()
    system = DeviceSystem(template.coordinates, template.connectivity, "C3D20R", "druckerprager", DP)
    torch.cuda.synchronize()
    log("[bench] device system ready: %d DOFs, %d nnz (%.1f s)" % (system.n_dof, system.nnz, time.perf_counter() - t0))
    if args.profile_spmv:
        profile_elements(system)
        return
    pt0 = system.phase_totals()
    t64, s64, m64 = run_footing(system, template, steps_for, "fp64", args.steps, args.warmup)
    pt1 = system.phase_totals()
    ph64 = phase_avg(pt0, pt1)
    q64 = m64[-1].qpState
    spmv_ms, spmv_alg = spmv_roofline(system)
    D = system.n_dof
    wall64 = sum(t64) / len(t64)
    nis64 = s64[-1].linearSolves / wall64
    result = {
        "metric": METRIC,
        "value": nis64,
        "unit": "Newton iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall64,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d^3 hexa20 C3D20R (2x2x2 Gauss) unit cube, 10%% jitter, non-associative "
                               "Drucker-Prager (phi 30, psi 10, c 10, E 30000, nu 0.2); geostatic prestress step "
                               "with gravity, then rigid footing u_y = -%g on the top-face centre quarter in %d "
                               "increments; GMRES(50) + multicolor block ILU(0)" % (n, settle, inc),
                   "elements": int(template.nElements), "dofs": int(D), "csr_nnz": int(system.nnz)},
        "dof_updates_per_s": nis64 * D,
        "newton_iterations_per_history": s64[-1].linearSolves,
        "newton_iterations_per_increment": [r["iters"] for r in s64[-1].history],
        "gmres_iterations_per_history": s64[-1].krylovIterations,
        "phase_ms": ph64,
    }
    if not args.no_mixed:
        t32, s32, m32 = run_footing(system, template, steps_for, "mixed", args.steps, args.warmup)
        wall32 = sum(t32) / len(t32)
        u64, u32 = m64[-1].U, m32[-1].U
        q32 = m32[-1].qpState
        result["mixed"] = {
            "newton_iterations_per_s": s32[-1].linearSolves / wall32,
            "dof_updates_per_s": s32[-1].linearSolves / wall32 * D,
            "ms_per_history": 1000.0 * wall32,
            "history_speedup_vs_fp64": wall64 / wall32,
            "newton_iterations_per_history": s32[-1].linearSolves,
            "newton_iterations_per_increment": [r["iters"] for r in s32[-1].history],
            "gmres_iterations_per_history": s32[-1].krylovIterations,
            "error_vs_fp64": {
                "U_rel_inf": float(np.abs(u32 - u64).max() / np.abs(u64).max()),
                "stress_rel_inf": float(np.abs(q32[..., 0:6] - q64[..., 0:6]).max() / np.abs(q64[..., 0:6]).max()),
                "kappa_rel_inf": float(np.abs(q32[..., 12] - q64[..., 12]).max()
                                       / max(np.abs(q64[..., 12]).max(), 1e-300)),
            },
            "phase_ms": phase_avg(pt1, system.phase_totals()),
        }
        log("[bench] footing mixed: %.3f s (%.2fx), U rel err %.2e" % (
            wall32, wall64 / wall32, result["mixed"]["error_vs_fp64"]["U_rel_inf"]))
    result["roofline"] = {
        "kernel": "spmv_kernel<double> (sliced-ELL 3x3 blocks)",
        "bound": "hbm",
        "achieved": spmv_alg / (spmv_ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": spmv_alg / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": None,
        "algorithmic_bytes": spmv_alg,
        "avg_launch_ms": spmv_ms,
        "csr_model_bytes": csr_model_bytes(system),
        "csr_model_achieved": csr_model_bytes(system) / (spmv_ms * 1e-3) / 1e9,
    }
    if not args.no_cpu:
        try:
            from oracle import cpu_baseline

            log("[bench] CPU baseline (bounded sample)")
            P, F = system.zeros(), system.zeros()
            system.compute_elements(system.zeros(), P, F)  # FP64 operator with the footing step's constraints
            data = system.export_csr_values().cpu().numpy()
            indptr, indices = system.csr_pattern()
            # the reference evaluates elements with tangents at every Newton iteration
            counts = {"element_passes": s64[-1].newtonChecks, "solves": s64[-1].linearSolves,
                      "krylov_iterations": s64[-1].krylovIterations}
            result["cpu_baseline"] = cpu_baseline.measure_footing(template.coordinates, template.connectivity, DP,
                                                                  indptr, indices, data, counts,
                                                                  budget_s=args.cpu_seconds, log=log)
        except Exception as e:
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


# The other workloads of the default run: a single history each (the batch with one warm-up
# history; the 96^3 and footing histories take 20-30 s each and run without one, so their
# figures include first-use costs: graph captures and the mixed-operator probe).
OTHER_WORKLOADS = (
    ("96cube_lognormal_j2", ["--workload", "lognormal", "--steps", "1", "--warmup", "0"]),
    ("batched_512_specimens", ["--workload", "batched", "--steps", "1", "--warmup", "1"]),
    ("footing_48cube_hexa20r_dp", ["--workload", "footing", "--steps", "1", "--warmup", "0"]),
)
_DROP = ("metric", "n_gpus", "higher_is_better", "scaling", "vs_baseline", "data")


def run_other_workloads(timeout_s=400):
    """Each of OTHER_WORKLOADS in a child process (its progress lines pass through on
    stderr); its JSON line without the keys it shares with the headline line, or an error
    entry when the child fails, so that one workload's failure does not lose the others."""
    out = {}
    for name, extra in OTHER_WORKLOADS:
        cmd = [sys.executable, "-u", os.path.abspath(__file__), "--no-cpu", "--no-others"] + extra
        log("[bench] other workload %s: %s" % (name, " ".join(extra)))
        try:
            proc = subprocess.run(cmd, stdout=subprocess.PIPE, timeout=timeout_s, text=True)
            lines = [ln for ln in proc.stdout.splitlines() if ln.startswith("{")]
            if proc.returncode != 0 or not lines:
                out[name] = {"error": "exit status %d" % proc.returncode}
                continue
            d = json.loads(lines[-1])
            out[name] = {k: v for k, v in d.items() if k not in _DROP}
        except Exception as e:  # reported in the line, never dropped silently
            out[name] = {"error": repr(e)}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--increments", type=int, default=20)
    ap.add_argument("--strain", type=float, default=0.01)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile-spmv", action="store_true",
                    help="counter-collection mode (no JSON line): cube, the operators of the load history; "
                         "footing, the C3D20R element passes and SpMV")
    ap.add_argument("--settlement", type=float, default=0.005)
    ap.add_argument("--workload", choices=("cube", "lognormal", "batched", "footing"), default="cube",
                    help="cube: 64^3 J2 load history (default); lognormal: 96^3 J2 with per-element lognormal "
                         "yield stress, 30 increments to 2 %% strain; batched: 512 uneven J2 specimens; "
                         "footing: 48^3 C3D20R Drucker-Prager geostatic + rigid footing, GMRES + ILU(0)")
    ap.add_argument("--specimens", type=int, default=512)
    ap.add_argument("--batch-increments", type=int, default=10)
    ap.add_argument("--no-others", action="store_true",
                    help="cube workload: skip the other workloads' child runs after the headline")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library runs the Newton step on one GPU (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        bench_batched(args)
        return
    if args.workload == "footing":
        if args.n == 64:
            args.n = 48
        if args.increments == 20:
            args.increments = 25
        bench_footing(args)
        return
    lognormal = args.workload == "lognormal"
    if lognormal:
        args.n, args.increments, args.strain = 96, 30, 0.02

    import torch

    from edelweissfe_amd.device import DeviceSystem

    log("[bench] building %d^3 hexa8 J2 model" % args.n)
    template, step_for = build_case(args.n, args.increments, args.strain, lognormal)
    t0 = time.perf_counter()
    system = DeviceSystem(template.coordinates, template.connectivity, "C3D8", "vonmises",
                          template.material["properties"])
    torch.cuda.synchronize()
    log("[bench] device system ready: %d DOFs, %d nnz (%.1f s)" % (system.n_dof, system.nnz, time.perf_counter() - t0))
    if args.profile_spmv:
        profile_spmv(system)
        return

    pt0 = system.phase_totals()
    t64, s64, m64 = run_history(system, template, step_for, "fp64", args.steps, args.warmup)
    pt1 = system.phase_totals()
    ph64 = phase_avg(pt0, pt1)
    q64 = m64[-1].qpState  # download now: the mixed run reuses the device state buffers
    spmv_ms, spmv_alg = spmv_roofline(system)
    log("[bench] fp64 SpMV %.4f ms, %.1f GB/s algorithmic" % (spmv_ms, spmv_alg / spmv_ms / 1e6))
    solves64 = s64[-1].linearSolves
    D = system.n_dof
    wall64 = sum(t64)
    nis64 = sum(s.linearSolves for s in s64) / wall64

    result = {
        "metric": METRIC,
        "value": nis64,
        "unit": "Newton iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall64 / len(t64),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "%d^3 hexa8 C3D8 unit cube, 10%% jitter, J2 linear hardening%s, uniaxial "
                               "compression to %g strain in %d increments, block-Jacobi PCG" % (
                                   args.n, " with per-element sigma_y lognormal (median 250, sigma_log 0.1, seed 1)"
                                   if lognormal else "", args.strain, args.increments),
                   "elements": int(template.nElements), "dofs": int(D), "csr_nnz": int(system.nnz)},
        "dof_updates_per_s": nis64 * D,
        "newton_iterations_per_history": solves64,
        "newton_iterations_per_increment": [r["iters"] for r in s64[-1].history],
        "krylov_iterations_per_history": s64[-1].krylovIterations,
        "phase_ms": ph64,
    }

    if not args.no_mixed:
        t32, s32, m32 = run_history(system, template, step_for, "mixed", args.steps, args.warmup)
        ph32 = phase_avg(pt1, system.phase_totals())
        wall32 = sum(t32)
        nis32 = sum(s.linearSolves for s in s32) / wall32
        u64, u32 = m64[-1].U, m32[-1].U
        q32 = m32[-1].qpState
        err = {
            "U_rel_inf": float(np.abs(u32 - u64).max() / np.abs(u64).max()),
            "U_rel_2": float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64)),
            "stress_rel_inf": float(np.abs(q32[..., 0:6] - q64[..., 0:6]).max() / np.abs(q64[..., 0:6]).max()),
            "kappa_rel_inf": float(np.abs(q32[..., 12] - q64[..., 12]).max() / max(np.abs(q64[..., 12]).max(), 1e-300)),
        }
        result["mixed"] = {
            "newton_iterations_per_s": nis32,
            "dof_updates_per_s": nis32 * D,
            "ms_per_history": 1000.0 * wall32 / len(t32),
            "history_speedup_vs_fp64": (wall64 / len(t64)) / (wall32 / len(t32)),
            "newton_iterations_per_history": s32[-1].linearSolves,
            "newton_iterations_per_increment": [r["iters"] for r in s32[-1].history],
            "krylov_iterations_per_history": s32[-1].krylovIterations,
            "operator_per_increment": [{1: "assembled", 2: "matfree"}.get(r.get("precision"), "fp64")
                                       for r in s32[-1].history],
            "error_vs_fp64": err,
            "phase_ms": ph32,
        }
        log("[bench] mixed: %.3f s/history (%.2fx), U rel err %.2e" % (
            wall32 / len(t32), result["mixed"]["history_speedup_vs_fp64"], err["U_rel_inf"]))

    traffic = load_traffic("spmv_fp64_64cube_traffic.json") if args.n == 64 else None
    result["roofline"] = {
        "kernel": "spmv_kernel<double> (sliced-ELL 3x3 blocks)",
        "bound": "hbm",
        "achieved": spmv_alg / (spmv_ms * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": spmv_alg / (spmv_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "traffic": traffic["bytes_per_launch"] if traffic else None,
        "algorithmic_bytes": spmv_alg,
        "avg_launch_ms": spmv_ms,
        "csr_model_bytes": csr_model_bytes(system),
        "csr_model_achieved": csr_model_bytes(system) / (spmv_ms * 1e-3) / 1e9,
    }

    if not args.no_cpu:
        try:
            from oracle import cpu_baseline

            log("[bench] CPU baseline (bounded sample)")
            # the reference evaluates elements with tangents at every Newton iteration
            counts = {"element_passes": s64[-1].newtonChecks, "solves": s64[-1].linearSolves,
                      "krylov_iterations": s64[-1].krylovIterations}
            result["cpu_baseline"] = cpu_baseline.measure(args.n, J2, args.strain, args.increments,
                                                          budget_s=args.cpu_seconds, counts=counts, log=log)
        except Exception as e:  # reported, never silently replaced
            result["cpu_baseline"] = {"error": repr(e)}
    if not args.no_others and not lognormal and args.n == 64:
        system.close()  # the children build their own device systems
        torch.cuda.empty_cache()
        result["other_workloads"] = run_other_workloads()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

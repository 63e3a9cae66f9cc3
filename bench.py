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
resident in HBM before the timed region starts.
"""

import argparse
import json
import os
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


def build_case(n, increments, strain):
    from edelweissfe_amd.generators.boxgen import generateBox
    from edelweissfe_amd.steps import AdaptiveStep, Dirichlet

    model = generateBox(n, n, n, 1.0, 1.0, 1.0, elType="C3D8", jitter=0.1, seed=0)
    model.assignMaterial("vonmises", J2)

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


def profile_spmv(system, n_launch=20, copy_mb=512):
    """Counter-collection mode: one FP64 assembly of the 64^3 operator, then n_launch SpMV
    launches and n_launch plain device copies of copy_mb MiB (the calibration kernel whose
    FETCH_SIZE / WRITE_SIZE per launch are known: copy_mb MiB each)."""
    import torch

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
    torch.cuda.synchronize()
    nb = system.nnz // 9
    alg = 8 * system.nnz + 4 * nb + 16 * system.n_dof
    log("[profile] %d spmv launches (algorithmic %d bytes each), %d copies of %d MiB" % (n_launch, alg, n_launch,
                                                                                      copy_mb))


def load_traffic(name):
    p = os.path.join(ROOT, "profiles", name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--increments", type=int, default=20)
    ap.add_argument("--strain", type=float, default=0.01)
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--profile-spmv", action="store_true", help="counter-collection mode (no JSON line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from edelweissfe_amd.device import DeviceSystem

    log("[bench] building %d^3 hexa8 J2 model" % args.n)
    template, step_for = build_case(args.n, args.increments, args.strain)
    t0 = time.perf_counter()
    system = DeviceSystem(template.coordinates, template.connectivity, "C3D8", "vonmises", J2)
    torch.cuda.synchronize()
    log("[bench] device system ready: %d DOFs, %d nnz (%.1f s)" % (system.n_dof, system.nnz, time.perf_counter() - t0))
    if args.profile_spmv:
        profile_spmv(system)
        return

    t64, s64, m64 = run_history(system, template, step_for, "fp64", args.steps, args.warmup)
    ph64 = system.phase_times()
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
        "config": {"workload": "%d^3 hexa8 C3D8 unit cube, 10%% jitter, J2 linear hardening, uniaxial compression "
                               "to %g strain in %d increments, block-Jacobi PCG" % (args.n, args.strain,
                                                                                   args.increments),
                   "elements": int(template.nElements), "dofs": int(D), "csr_nnz": int(system.nnz)},
        "dof_updates_per_s": nis64 * D,
        "newton_iterations_per_history": solves64,
        "newton_iterations_per_increment": [r["iters"] for r in s64[-1].history],
        "krylov_iterations_per_history": s64[-1].krylovIterations,
        "phase_ms": ph64,
    }

    if not args.no_mixed:
        t32, s32, m32 = run_history(system, template, step_for, "mixed", args.steps, args.warmup)
        ph32 = system.phase_times()
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
            "error_vs_fp64": err,
            "phase_ms": ph32,
        }
        log("[bench] mixed: %.3f s/history (%.2fx), U rel err %.2e" % (
            wall32 / len(t32), result["mixed"]["history_speedup_vs_fp64"], err["U_rel_inf"]))

    traffic = load_traffic("spmv_fp64_64cube_traffic.json")
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
    }

    if not args.no_cpu:
        try:
            from oracle import cpu_baseline

            log("[bench] CPU baseline (bounded sample)")
            counts = {"element_passes": s64[-1].elementPasses, "solves": s64[-1].linearSolves,
                      "krylov_iterations": s64[-1].krylovIterations}
            result["cpu_baseline"] = cpu_baseline.measure(args.n, J2, args.strain, args.increments,
                                                          budget_s=args.cpu_seconds, counts=counts, log=log)
        except Exception as e:  # reported, never silently replaced
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

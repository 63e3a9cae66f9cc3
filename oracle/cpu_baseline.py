"""Timed CPU baseline of the benchmark workload (the reference algorithm restated in C).

The reference itself never runs on the GPU host, so the baseline is the C
restatement in oracle/csrc/oracle_nist.c, on all host threads OpenMP gives it:
element loop over elements (nonlinearimplicitstaticparallelmk2.pyx:157-185),
serial COO->CSR scatter (csrgenerator.pyx:110-113), serial P/F accumulation,
symmetric Dirichlet elimination, and a 3x3 block-Jacobi CPU PCG with the same
stopping rule as the device solve (a direct solve, the reference's choice, is
infeasible at this size; this is labelled "not the reference solver").

Bounded sample: the first Newton iteration of increment 1 of the load history
(one element pass, one scatter, one solve), timed phase by phase; the PCG is
cut off after a time budget and its per-iteration time is used.  The sample is
scaled to the whole history with the element-pass, solve and Krylov-iteration
counts of the GPU run of the same history (same algorithm, same counts).
"""

import os
import time

import numpy as np
import scipy.sparse as sp

from . import clib
from . import mesh as omesh


def measure(n, props, strain, increments, budget_s=20.0, counts=None, log=print):
    t_setup = time.perf_counter()
    m = omesh.box_mesh(n, n, n, 1.0, 1.0, 1.0, nodes_per_element=8, jitter=0.1, seed=0)
    coords, conn = m["coords"], m["conn"].astype(np.int32)
    N = coords.shape[0]
    D = 3 * N
    E = conn.shape[0]
    dofmap = (3 * conn[:, :, None] + np.arange(3, dtype=np.int32)[None, None, :]).reshape(E, 24).astype(np.int32)
    indptr, indices = clib.csr_pattern(D, dofmap)
    x = clib.x_map(dofmap, indptr, indices)
    ns = m["nodesets"]
    bcs = [("gen_top", 1, -strain), ("gen_bottom", 1, 0.0), ("gen_left", 0, 0.0), ("gen_back", 2, 0.0)]
    dir_idx = np.concatenate([3 * ns[s] + c for s, c, _ in bcs])
    p = 1.0 / increments
    dir_val = np.concatenate([np.full(ns[s].size, v * (p - (p - p))) for s, _, v in bcs])
    state = np.zeros((E, 8, 13))
    log("[cpu] setup %d^3: %d DOFs, %d nnz, %.1f s" % (n, D, indices.size, time.perf_counter() - t_setup))

    # one Newton iteration (iteration 0 of increment 1: dU = 0, R[dir] = delta)
    t0 = time.perf_counter()
    V, Pe, _, fails = clib.elements_hexa8(coords, conn, np.asarray(props, dtype=np.float64), 1, np.zeros(D), state)
    t_el = time.perf_counter() - t0
    t0 = time.perf_counter()
    data = clib.update_csr(V, x, indices.size)
    t_csr = time.perf_counter() - t0
    t0 = time.perf_counter()
    P, F = clib.gather_pf(dofmap, Pe, D)
    t_pf = time.perf_counter() - t0
    del V
    t0 = time.perf_counter()
    K = sp.csr_matrix((data, indices, indptr), shape=(D, D))
    delta = np.zeros(D)
    delta[dir_idx] = dir_val
    R = P.copy()
    R[dir_idx] = dir_val
    b = R - K @ delta
    b[dir_idx] = dir_val
    mask = np.zeros(D, dtype=bool)
    mask[dir_idx] = True
    rows = np.repeat(np.arange(D), np.diff(indptr))
    kill = mask[rows] | mask[indices]
    K.data[kill] = 0.0
    K.data[kill & (rows == indices)] = 1.0
    t_dir = time.perf_counter() - t0
    avg_flux = max(1e-10, np.abs(F).sum() / (E * 24))
    atol = 1e-2 * max(1e-8 * avg_flux, 1e-7)
    _, its, res, t_pcg = clib.pcg(indptr, indices, K.data, b, 1e-13, atol, 100000, budget_s)
    t_it = t_pcg / max(its, 1)
    threads = clib.threads()
    log("[cpu] elements %.2f s, scatter %.2f s, P/F %.2f s, Dirichlet %.2f s, PCG %d its in %.2f s (%.2f ms/it), "
        "%d threads" % (t_el, t_csr, t_pf, t_dir, its, t_pcg, 1e3 * t_it, threads))

    out = {"unit": "Newton iterations/s", "cores": threads, "kind": "port"}
    t_pass = t_el + t_csr + t_pf
    if counts:
        hist = (counts["element_passes"] * t_pass + counts["solves"] * t_dir + counts["krylov_iterations"] * t_it)
        out["value"] = counts["solves"] / hist
        out["history_seconds_estimate"] = hist
        out["sample"] = ("%d^3 J2 history, increment 1 Newton iteration 0 timed phase by phase on %d OpenMP threads "
                         "(C restatement: element loop %.2f s, serial COO->CSR scatter %.2f s, P/F %.2f s, "
                         "Dirichlet %.2f s, block-Jacobi PCG %d its at %.2f ms/it, not the reference direct "
                         "solver), scaled to the history with the GPU run's counts: %d element passes, %d solves, "
                         "%d Krylov iterations" % (n, threads, t_el, t_csr, t_pf, t_dir, its, 1e3 * t_it,
                                                   counts["element_passes"], counts["solves"],
                                                   counts["krylov_iterations"]))
    else:
        one = t_pass + t_dir + its * t_it
        out["value"] = 1.0 / one
        out["sample"] = "%d^3 J2 history, one Newton iteration (%d PCG its), %d threads" % (n, its, threads)
    out["phase_seconds"] = {"elements": t_el, "csr_scatter": t_csr, "pf_gather": t_pf, "dirichlet": t_dir,
                            "pcg_per_iteration": t_it, "pcg_iterations_sampled": its, "material_failures": int(fails)}
    out["omp_num_threads_env"] = os.environ.get("OMP_NUM_THREADS")
    return out

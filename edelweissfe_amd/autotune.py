"""Launch-shape autotuner: elements per workgroup of the element kernel and threads per
workgroup of the Krylov kernels, chosen by timing on the device the system lives on.

Results are cached per (element type, material, element count, node count, device name,
library build) in a JSON file (EFE_AUTOTUNE_CACHE, default edelweissfe_amd/lib/autotune.json):
a rebuilt library (another size or modification time of libefe_gpu.so) is tuned afresh, so a
changed kernel never runs with the launch shapes or operator costs measured for the old one.
"""

import json
import os

import numpy as np
import torch

from . import _lib

EPW_CANDIDATES = {"C3D8": (1, 2, 4, 8), "C3D20": (1, 2), "C3D20R": (2,)}
KRYLOV_CANDIDATES = (64, 128, 256, 512)


def _cache_path():
    return os.environ.get("EFE_AUTOTUNE_CACHE",
                          os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "autotune.json"))


def _build_id():
    """Identity of the loaded library build: size and modification time of libefe_gpu.so."""
    try:
        st = os.stat(_lib.LIB_PATH)
        return "%d-%d" % (st.st_size, st.st_mtime_ns)
    except OSError:
        return "unknown"


def _key(system):
    dev = torch.cuda.get_device_name(system.device)
    return "%s|%s|%d|%d|%s|%s" % (system.el_type, system.material, system.n_elements, system.n_nodes, dev,
                                  _build_id())


def _load():
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(cache):
    try:
        os.makedirs(os.path.dirname(_cache_path()), exist_ok=True)
        with open(_cache_path(), "w") as f:
            json.dump(cache, f, indent=1)
    except OSError:
        pass


def tune(system, precision=_lib.EFE_FP64, reps=3, use_cache=True):
    """Times every candidate and leaves the fastest configured on `system`.
    Returns {"elements_per_workgroup", "krylov_block", "timings"}."""
    cache = _load() if use_cache else {}
    key = _key(system) + "|%d" % precision
    if key in cache:
        best = cache[key]
        system.set_launch(best["elements_per_workgroup"], best["krylov_block"])
        return best
    state = system.qp_state(0)
    dU = torch.zeros(system.n_dof, dtype=torch.float64, device=system.device)
    P, F = system.zeros(), system.zeros()
    timings = {"elements": {}, "krylov": {}}
    for epw in EPW_CANDIDATES[system.el_type]:
        system.set_launch(epw, 0)
        ts = []
        for _ in range(reps):
            system.compute_elements(dU, P, F, precision=precision)
            ts.append(system.phase_times()["elements"])
        timings["elements"][epw] = float(np.min(ts))
    best_epw = min(timings["elements"], key=timings["elements"].get)
    system.set_launch(best_epw, 0)
    system.compute_elements(dU, P, F, precision=precision)
    b = torch.randn(system.n_dof, dtype=torch.float64, device=system.device)
    x = torch.empty_like(b)
    for kb in KRYLOV_CANDIDATES:
        system.set_launch(0, kb)
        ts = []
        for _ in range(reps):
            system.linsolve(b, x, precision=precision, rtol=1e-30, max_iter=64, check_every=32)
            ts.append(system.phase_times()["krylov_iteration"])
        timings["krylov"][kb] = float(np.min(ts))
    best_kb = min(timings["krylov"], key=timings["krylov"].get)
    system.set_launch(best_epw, best_kb)
    system.set_qp_state(state)
    best = {"elements_per_workgroup": int(best_epw), "krylov_block": int(best_kb),
            "timings": {k: {str(a): b for a, b in v.items()} for k, v in timings.items()}}
    cache[key] = best
    _store(cache)
    return best


def mixed_operator_costs(system, reps=2, use_cache=True):
    """Device time (ms) of one element pass (elements + assembly) and of one PCG iteration
    with each FP32 operator of the mixed form: "matfree" (compact per-QP tangents) and
    "assembled" (sliced-ELL FP32 operator).  Which one is faster for a whole load history
    depends on the Krylov iterations per element pass; NISTGPU weighs the two with the ratio
    it observes.  Cached like the launch shapes."""
    cached = getattr(system, "_mixed_operator_costs", None)
    if cached is not None:
        return cached
    cache = _load() if use_cache else {}
    key = _key(system) + "|mixed-operators"
    if key in cache:
        system._mixed_operator_costs = cache[key]
        return cache[key]
    dU = torch.zeros(system.n_dof, dtype=torch.float64, device=system.device)
    P, F = system.zeros(), system.zeros()
    b = torch.randn(system.n_dof, dtype=torch.float64, device=system.device)
    x = torch.empty_like(b)
    out = {}
    for name, prec in (("matfree", _lib.EFE_FP32_MATFREE), ("assembled", _lib.EFE_FP32)):
        tp, ti = [], []
        for _ in range(reps):
            system.compute_elements(dU, P, F, precision=prec)
            ph = system.phase_times()
            tp.append(ph["elements"] + ph["assembly"])
        for _ in range(reps):
            system.linsolve(b, x, precision=prec, rtol=1e-30, max_iter=64, check_every=32)
            ti.append(system.phase_times()["krylov_iteration"])
        out[name] = {"pass_ms": float(np.min(tp)), "iteration_ms": float(np.min(ti))}
    system._mixed_operator_costs = out
    cache[key] = out
    _store(cache)
    return out


def choose_mixed_operator(costs, iterations_per_pass):
    """The FP32 operator with the smaller pass + iterations_per_pass x iteration time."""
    return min(costs, key=lambda k: costs[k]["pass_ms"] + iterations_per_pass * costs[k]["iteration_ms"])

"""Batched specimens through the C ABI: each specimen of a block-diagonal batch follows its
own oracle NIST run (increments, cutbacks, Newton counts, U and QP stress)."""

import numpy as np
import pytest

from helpers import (CUTBACK_SPECIMENS, cutback_oracle_run, device_model, device_step, oracle_model, oracle_step,
                     rel_err, uniaxial_bcs)
from oracle.nist import NIST

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

# mesh, E, yield stress, applied strain, maxInc, maxIter; the second specimen needs three
# Newton iterations in its first plastic increment but may take two, so it cuts back
SPECIMENS = [
    ((2, 2, 2), 210000.0, 250.0, 0.010, 0.25, 15),
    ((3, 3, 3), 180000.0, 300.0, 0.020, 1.0, 2),
    ((4, 3, 2), 150000.0, 200.0, 0.015, 0.2, 15),
    ((5, 4, 4), 250000.0, 350.0, 0.010, 0.1, 15),
]


def _props(E, fy):
    return [E, 0.3, fy, 1000.0, 0.0, 0.0]


def _batch(precision="fp64", autotune=False):
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    models, steps = [], []
    for s, (n, E, fy, strain, maxInc, maxIter) in enumerate(SPECIMENS):
        pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", _props(E, fy), jitter=0.1, seed=s)
        models.append(pm)
        steps.append(device_step(pm, uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=100,
                                 maxIter=maxIter))
    solver = NISTBatched(precision=precision, autotune=autotune)
    solver.solveSteps(models, steps)
    return solver, models


def _oracle(s):
    n, E, fy, strain, maxInc, maxIter = SPECIMENS[s]
    om = oracle_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", _props(E, fy), jitter=0.1, seed=s)
    so = NIST()
    so.solve_step(om, oracle_step(uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=100, maxIter=maxIter))
    return so, om


def test_batched_specimens_follow_their_oracle_runs(tmp_path, monkeypatch):
    monkeypatch.setenv("EFE_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    solver, models = _batch(autotune=True)
    assert solver.tuning["elements_per_workgroup"] in (1, 2, 4, 8)
    states = solver.qpStates(models)
    for s, pm in enumerate(models):
        so, om = _oracle(s)
        want = [np.inf if h[2] is None else h[2] for h in so.history]
        print(s, solver.history[s], want)
        assert solver.history[s] == want
        assert rel_err(pm.U, om.U) < 1e-9
        assert rel_err(states[s][..., 0:6], om.state[..., 0:6]) < 1e-9
        assert pm.time == pytest.approx(1.0)
    assert np.inf in solver.history[1]


def test_batched_mixed_precision_error_vs_fp64():
    _, m64 = _batch("fp64")
    _, m32 = _batch("mixed")
    for a, b in zip(m32, m64):
        assert rel_err(a.U, b.U) < 1e-8


def _two_specimen_system():
    from edelweissfe_amd.device import DeviceSystem

    a = device_model((3, 2, 2), (1.0, 1.0, 1.0), "C3D8", "vonmises", _props(210000.0, 250.0), jitter=0.1, seed=3)
    b = device_model((2, 3, 4), (1.0, 1.0, 1.0), "C3D8", "vonmises", _props(150000.0, 180.0), jitter=0.1, seed=4)
    coords = np.vstack([a.coordinates, b.coordinates])
    conn = np.vstack([a.connectivity, b.connectivity + a.nNodes])
    props = np.vstack([np.tile(a.material["properties"], (a.nElements, 1)),
                       np.tile(b.material["properties"], (b.nElements, 1))])
    spec = np.r_[np.zeros(a.nNodes, np.int32), np.ones(b.nNodes, np.int32)]
    return DeviceSystem(coords, conn, "C3D8", "vonmises", props, specimen_of_node=spec, n_specimens=2), a, b


def test_batched_norms_solve_and_selective_accept():
    sysm, a, b = _two_specimen_system()
    D, da = sysm.n_dof, 3 * a.nNodes
    g = torch.Generator(device="cpu").manual_seed(7)
    dU = (torch.rand(D, generator=g, dtype=torch.float64) * 4e-3 - 2e-3).cuda()
    P, F = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, P, F)
    # per-specimen convergence quantities
    ddU = (torch.rand(D, generator=g, dtype=torch.float64) - 0.5).cuda()
    out = sysm.norms_batched(P, ddU, F)
    Ph, Fh, dh = P.cpu().numpy(), F.cpu().numpy(), ddU.cpu().numpy()
    for s, sl in enumerate((slice(0, da), slice(da, D))):
        np.testing.assert_allclose(out[s, 0], np.abs(Ph[sl]).max(), rtol=0, atol=0)
        np.testing.assert_allclose(out[s, 1], np.abs(dh[sl]).max(), rtol=0, atol=0)
        np.testing.assert_allclose(out[s, 2], np.abs(Fh[sl]).sum(), rtol=1e-13)
        assert out[s, 3] == 0.0
    # batched PCG against a direct solve of each specimen's block
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    dirs = np.r_[a.dofsOfNodeSet("gen_bottom"), da + b.dofsOfNodeSet("gen_bottom")]
    sysm.set_dirichlet(dirs)
    sysm.compute_elements(dU, P, F)
    K = sp.csr_matrix((sysm.export_csr_values().cpu().numpy(), sysm.csr_pattern()[1], sysm.csr_pattern()[0]),
                      shape=(D, D))
    rhs = P.clone()
    rhs[torch.as_tensor(dirs, device=rhs.device)] = 0.0
    x = torch.empty_like(rhs)
    info, its = sysm.linsolve_batched(rhs, x, np.array([1e-11, 1e-11]), rtol=1e-14, max_iter=5000)
    assert info.status == 0 and (its > 0).all()
    ref = spla.spsolve(K.tocsc(), rhs.cpu().numpy())
    assert rel_err(x.cpu().numpy(), ref) < 1e-9
    # an excluded specimen gets x = 0 and no iterations
    info, its = sysm.linsolve_batched(rhs, x, np.array([1e-9, 1e-9]), spec_active=[1, 0], rtol=1e-14, max_iter=5000)
    assert its[1] == 0 and not x[da:].abs().max().item()
    # accept specimen 1 only: its reference state becomes the temporary state, specimen 0 keeps zeros
    sysm.accept_state_batched([0, 1])
    ref_state, tmp_state = sysm.qp_state(0), sysm.qp_state(1)
    ea = a.nElements
    assert not ref_state[:ea].any()
    assert np.array_equal(ref_state[ea:], tmp_state[ea:])
    assert np.abs(tmp_state[ea:, :, 0:6]).max() > 0


def test_per_specimen_relative_tolerances_match_solves_alone():
    """efe_linsolve_batched with spec_rtol: each specimen stops by its own relative tolerance;
    its iteration count and solution equal, bit for bit, those of the specimen solved as the
    only one of its own system with that tolerance."""
    from edelweissfe_amd.device import DeviceSystem

    sysm, a, b = _two_specimen_system()
    D, da = sysm.n_dof, 3 * a.nNodes
    g = torch.Generator(device="cpu").manual_seed(11)
    dU = (torch.rand(D, generator=g, dtype=torch.float64) * 4e-3 - 2e-3).cuda()
    dirs = np.r_[a.dofsOfNodeSet("gen_bottom"), da + b.dofsOfNodeSet("gen_bottom")]
    sysm.set_dirichlet(dirs)
    P, F = sysm.zeros(), sysm.zeros()
    sysm.compute_elements(dU, P, F)
    rhs = P.clone()
    rhs[torch.as_tensor(dirs, device=rhs.device)] = 0.0
    rtols = [1e-7, 1e-4]
    x = torch.empty_like(rhs)
    info, its = sysm.linsolve_batched(rhs, x, np.array([1e-300, 1e-300]), spec_rtol=rtols, rtol=0.5, max_iter=5000,
                                      check_every=1)
    assert info.status == 0 and (its > 0).all()
    # the tolerances swapped: each specimen's count follows its own tolerance
    _, its_sw = sysm.linsolve_batched(rhs, torch.empty_like(rhs), np.array([1e-300, 1e-300]), spec_rtol=rtols[::-1],
                                      rtol=0.5, max_iter=5000, check_every=1)
    assert its_sw[0] < its[0] and its_sw[1] > its[1]
    for s, (m, sl, d) in enumerate(((a, slice(0, da), a.dofsOfNodeSet("gen_bottom")),
                                    (b, slice(da, D), b.dofsOfNodeSet("gen_bottom")))):
        one = DeviceSystem(m.coordinates, m.connectivity, "C3D8", "vonmises",
                           np.tile(m.material["properties"], (m.nElements, 1)),
                           specimen_of_node=np.zeros(m.nNodes, np.int32), n_specimens=1)
        one.set_dirichlet(d)
        P1, F1 = one.zeros(), one.zeros()
        one.compute_elements(dU[sl].contiguous(), P1, F1)
        assert torch.equal(P1, P[sl])
        r1 = P1.clone()
        r1[torch.as_tensor(d, device=r1.device)] = 0.0
        x1 = torch.empty_like(r1)
        info1, its1 = one.linsolve_batched(r1, x1, np.array([1e-300]), rtol=rtols[s], max_iter=5000, check_every=1)
        print(s, rtols[s], its[s], its1[0], flush=True)
        assert its1[0] == its[s]
        assert torch.equal(x1, x[sl])
        one.close()
    sysm.close()


def test_material_cutbacks_and_step_failure_stay_per_specimen():
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    models, steps = [], []
    for n, props, strain, maxInc, maxNumInc, seed in CUTBACK_SPECIMENS:
        pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", props, jitter=0.1, seed=seed)
        models.append(pm)
        steps.append(device_step(pm, uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=maxNumInc,
                                 maxIter=15))
    solver = NISTBatched(precision="fp64", autotune=False, krylovAtolFactor=1e-8)
    ok, _ = solver.solveSteps(models, steps)
    states = solver.qpStates(models)
    for s, pm in enumerate(models):
        so, om, failed = cutback_oracle_run(*CUTBACK_SPECIMENS[s])
        want = [np.inf if h[2] is None else h[2] for h in so.history]
        print(s, solver.history[s], want, solver.failures[s], flush=True)
        assert failed == (s == 2)
        assert solver.history[s] == want
        # the state and time of the last accepted increment
        assert rel_err(states[s][..., 0:6], om.state[..., 0:6]) < 1e-9
        assert pm.time == pytest.approx(om.time)
        if not failed:
            assert rel_err(pm.U, om.U) < 1e-9
    assert not ok and solver.failedSpecimens == [2]
    assert np.inf in solver.history[0] and solver.history[0][-1] != np.inf


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_specimen_alone_and_in_a_batch_are_bitwise_equal(precision):
    """The batched PCG sums each specimen's dot products over 256-node chunks that start at
    the specimen's first node: a specimen of several chunks (8^3, 729 nodes), placed behind
    others so that its nodes do not start on a chunk boundary of the batch, takes the same
    Krylov iterations and ends with the same bits as when it is the only specimen."""
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    cases = [((2, 2, 2), 210000.0, 250.0, 0), ((8, 8, 8), 200000.0, 260.0, 7), ((3, 3, 3), 180000.0, 300.0, 2)]

    def run(which):
        models, steps = [], []
        for n, E, fy, seed in [cases[i] for i in which]:
            pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", _props(E, fy), jitter=0.1, seed=seed)
            models.append(pm)
            steps.append(device_step(pm, uniaxial_bcs(0.01), maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15))
        solver = NISTBatched(precision=precision, autotune=False)
        solver.solveSteps(models, steps)
        return solver, models

    alone, ma = run([1])
    batch, mb = run([0, 1, 2])
    print(alone.history[0], batch.history[1], alone.specimenKrylov[0], batch.specimenKrylov[1])
    assert alone.history[0] == batch.history[1]
    assert alone.specimenKrylov[0] == batch.specimenKrylov[1] > 0
    assert np.array_equal(ma[0].U, mb[1].U)
    assert np.array_equal(alone.qpStates(ma)[0], batch.qpStates(mb)[1])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_groups_by_element_and_material_on_the_device(precision):
    """NISTBatchedGroups: five interleaved specimens of three (element type, material)
    groups (hexa8 J2, hexa8 linear elastic, hexa20R J2), one batched context per group;
    every specimen follows its own oracle NIST run (FP64: Newton counts equal, U and stress
    to 1e-9; mixed, which runs the hexa20R group on the per-slot matrix-free operator with
    frozen specimens skipped: U to 1e-8)."""
    from edelweissfe_amd.solvers import NISTBatchedGroups
    from test_batched_control import GROUPED

    models, steps = [], []
    for et, mat, props, n, strain, seed in GROUPED:
        pm = device_model(n, (1.0, 1.0, 1.0), et, mat, props, jitter=0.1, seed=seed)
        models.append(pm)
        steps.append(device_step(pm, uniaxial_bcs(strain), maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15))
    solver = NISTBatchedGroups(precision=precision, autotune=False)
    ok, _ = solver.solveSteps(models, steps)
    assert ok and [g[1] for g in solver.groups] == [[0, 3], [1, 4], [2]]
    states = solver.qpStates(models)
    for s, (et, mat, props, n, strain, seed) in enumerate(GROUPED):
        om = oracle_model(n, (1.0, 1.0, 1.0), et, mat, props, jitter=0.1, seed=seed)
        so = NIST()
        so.solve_step(om, oracle_step(uniaxial_bcs(strain), maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15))
        want = [np.inf if h[2] is None else h[2] for h in so.history]
        print(s, et, mat, solver.history[s], want)
        if precision == "fp64":
            assert solver.history[s] == want
            assert rel_err(models[s].U, om.U) < 1e-9
            assert rel_err(states[s][..., 0:6], om.state[..., 0:6]) < 1e-9
        else:
            assert rel_err(models[s].U, om.U) < 1e-8


def _run_specimens(specs, precision, newton_on_device, **kw):
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    models, steps = [], []
    for n, props, strain, maxInc, maxNumInc, maxIter, seed in specs:
        pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", props, jitter=0.1, seed=seed)
        models.append(pm)
        steps.append(device_step(pm, uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=maxNumInc,
                                 maxIter=maxIter))
    solver = NISTBatched(precision=precision, autotune=False, newtonOnDevice=newton_on_device, **kw)
    ok, _ = solver.solveSteps(models, steps)
    return solver, models, ok


@pytest.mark.parametrize("case", ["plain-fp64", "plain-mixed", "cutbacks-fp64", "cutbacks-mixed"])
def test_newton_round_on_device_equals_host_round_bitwise(case):
    """efe_newton_round_batched (every convergence, divergence, iteration-limit, cutback and
    Krylov decision on the device, one replayed hipGraph per Newton pass) against the host
    round it replaces: the same increments, Newton counts, failure reasons, work counts,
    failed specimens, and the same U and QP state bit for bit.  The cases take in a specimen
    that cuts back by its iteration limit (maxIter 2), material cutbacks and a step failure."""
    kind, precision = case.split("-")
    if kind == "plain":
        specs = [(n, _props(E, fy), strain, maxInc, 100, maxIter, s)
                 for s, (n, E, fy, strain, maxInc, maxIter) in enumerate(SPECIMENS)]
        kw = {}
    else:
        specs = [(n, props, strain, maxInc, maxNumInc, 15, seed)
                 for n, props, strain, maxInc, maxNumInc, seed in CUTBACK_SPECIMENS]
        kw = {"krylovAtolFactor": 1e-8}
    host, mh, okh = _run_specimens(specs, precision, False, **kw)
    dev, md, okd = _run_specimens(specs, precision, True, **kw)
    print(case, host.history, dev.history, host.failures, dev.failures, flush=True)
    assert okh == okd
    assert host.history == dev.history
    assert host.failures == dev.failures
    assert host.failedSpecimens == dev.failedSpecimens
    assert np.array_equal(host.specimenPasses, dev.specimenPasses)
    assert np.array_equal(host.specimenSolves, dev.specimenSolves)
    assert np.array_equal(host.specimenKrylov, dev.specimenKrylov)
    assert (host.elementPasses, host.linearSolves, host.krylovIterations) == (dev.elementPasses, dev.linearSolves,
                                                                            dev.krylovIterations)
    for a, b in zip(mh, md):
        assert np.array_equal(a.U, b.U)
    for a, b in zip(host.qpStates(mh), dev.qpStates(md)):
        assert np.array_equal(a, b)
    if precision == "fp64":  # the cases take the paths they are meant to
        if kind == "plain":
            assert np.inf in dev.history[1]
        else:
            assert "cutback" in dev.failures[0] and dev.failedSpecimens == [2]

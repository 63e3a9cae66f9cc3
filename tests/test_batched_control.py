"""NISTBatched control flow on the CPU: the driver runs against an oracle-backed stand-in
for the device system (tests/batch_mock.py), so every specimen of a batch must follow its
own oracle NIST run exactly (increments, cutbacks, Newton counts, step failure, U and
state).  This isolates the lockstep round logic from device numerics."""

import numpy as np
import pytest

from batch_mock import attach_mock
from helpers import (CUTBACK_SPECIMENS, cutback_oracle_run, device_model, device_step, oracle_model, rel_err,
                     uniaxial_bcs)


def _run(specimens):
    from edelweissfe_amd.solvers.nistbatched import NISTBatched

    models, steps, oms = [], [], []
    for n, props, strain, maxInc, maxNumInc, seed in specimens:
        pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", props, jitter=0.1, seed=seed)
        models.append(pm)
        steps.append(device_step(pm, uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=maxNumInc,
                                 maxIter=15))
        oms.append(oracle_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", props, jitter=0.1, seed=seed))
    solver = NISTBatched(precision="fp64", autotune=False)
    mock = attach_mock(solver, models, oms)
    ok, _ = solver.solveSteps(models, steps)
    return solver, mock, models, ok


def test_cutbacks_and_step_failure_follow_each_oracle_run():
    solver, mock, models, ok = _run(CUTBACK_SPECIMENS)
    states = solver.qpStates(models)
    for s, pm in enumerate(models):
        so, om, failed = cutback_oracle_run(*CUTBACK_SPECIMENS[s])
        want = [np.inf if h[2] is None else h[2] for h in so.history]
        assert solver.history[s] == want, (s, solver.history[s], want, solver.failures[s])
        assert rel_err(states[s][..., 0:6], om.state[..., 0:6]) < 1e-12
        assert pm.time == pytest.approx(om.time)
        if not failed:
            assert rel_err(pm.U, om.U) < 1e-12
    assert not ok and solver.failedSpecimens == [2]
    assert set(solver.failures[0]) == {"cutback"}


def test_specimen_alone_matches_specimen_in_batch():
    # the same specimen solved alone and next to others takes the same increments
    alone, _, m1, _ = _run(CUTBACK_SPECIMENS[:1])
    batch, _, m3, _ = _run(CUTBACK_SPECIMENS)
    assert alone.history[0] == batch.history[0]
    np.testing.assert_array_equal(m1[0].U, m3[0].U)

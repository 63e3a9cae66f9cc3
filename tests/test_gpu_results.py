"""Device result collection (efe_collect_results / ElementResultCollector) against the QP
state of a J2 load history, in the reference collector's shapes
(elementresultcollector.pyx:37-132)."""

import numpy as np
import pytest

from helpers import device_model, device_step, uniaxial_bcs

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

J2 = [210000.0, 0.3, 250.0, 1000.0, 0.0, 0.0]


def test_collector_matches_qp_state_after_a_j2_history():
    from edelweissfe_amd.solvers import NISTGPU
    from edelweissfe_amd.utils.elementresultcollector import ElementResultCollector

    pm = device_model((4, 3, 5), (1.0, 1.0, 1.0), "C3D8", "vonmises", J2, jitter=0.1, seed=2)
    st = device_step(pm, uniaxial_bcs(0.01), maxInc=0.25, minInc=1e-6, maxNumInc=100, maxIter=15)
    s = NISTGPU(precision="fp64")
    s.solveStep(st, pm)
    sysm = s.system
    q = pm.qpState  # [element][QP][state]
    els = np.array([7, 0, 59, 33, 12])
    c = ElementResultCollector(sysm, els, range(0, 8), "stress")
    r = c.getCurrentResults().cpu().numpy()
    assert r.shape == (5, 8, 6)
    np.testing.assert_array_equal(r, q[els][:, :, 0:6])
    # the field output "mean over QPs 0:8" of the reference's test files
    np.testing.assert_allclose(r.mean(axis=1), q[els, :, 0:6].mean(axis=1), rtol=0, atol=0)
    r = ElementResultCollector(sysm, None, [3], "STRAIN").getCurrentResults().cpu().numpy()
    assert r.shape == (pm.nElements, 6)
    np.testing.assert_array_equal(r, q[:, 3, 6:12])
    r = ElementResultCollector(sysm, els, [1, 6], "kappa").getCurrentResults().cpu().numpy()
    np.testing.assert_array_equal(r[..., 0], q[els][:, [1, 6], 12])
    assert (q[..., 12] > 0).any()
    with pytest.raises(Exception):
        ElementResultCollector(sysm, els, [0], "porePressure")


def test_status_file_of_a_cutback_run_matches_the_oracle(tmp_path):
    """The status file (statusfile.py) of a device run with material cutbacks against the
    oracle run's increments: number, Newton iterations (inf when failed), converged flag,
    time increment, end time and failure message; the compute-time monitor reports every
    increment."""
    from edelweissfe_amd.outputmanagers import ComputeTimeMonitor, StatusFile
    from edelweissfe_amd.outputmanagers.statusfile import read_status_file
    from edelweissfe_amd.solvers import NISTGPU
    from helpers import CUTBACK_SPECIMENS, cutback_oracle_run

    n, props, strain, maxInc, maxNumInc, seed = CUTBACK_SPECIMENS[0]
    pm = device_model(n, (1.0, 1.0, 1.0), "C3D8", "vonmises", props, jitter=0.1, seed=seed)
    sf, ct = StatusFile(), ComputeTimeMonitor()
    sf.updateDefinition(filename=str(tmp_path / "job.sta"))
    ct.initializeStep(None)
    s = NISTGPU(precision="fp64", krylovAtolFactor=1e-8)
    s.solveStep(device_step(pm, uniaxial_bcs(strain), maxInc=maxInc, minInc=1e-6, maxNumInc=maxNumInc, maxIter=15),
                pm, outputmanagers=[sf, ct])
    so, om, failed = cutback_oracle_run(*CUTBACK_SPECIMENS[0])
    rows = read_status_file(tmp_path / "job.sta")
    want = so.increments
    print([(r["inc"], r["iters"]) for r in rows], [(w["inc"], w["iters"]) for w in want])
    assert not failed and any(not w["converged"] for w in want)
    assert len(rows) == len(want)
    for r, w in zip(rows, want):
        assert (r["inc"], r["iters"], r["converged"], r["notes"]) == (w["inc"], float(w["iters"]), w["converged"],
                                                                      w["notes"])
        assert r["time inc"] == pytest.approx(w["time inc"], rel=5e-4)
        assert r["time end"] == pytest.approx(w["time end"], rel=5e-4)
    assert len(ct.increments) == len(want) and ct.increments[-1][2] > 0

"""Material edge cases on the device, against the oracle's behaviour."""

import numpy as np
import pytest

from helpers import device_model

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


def test_drucker_prager_apex_without_dilatancy_requests_a_cutback():
    """psi = 0 and a trial state in hydrostatic tension beyond the apex: no return exists, so
    the element pass returns EFE_CUTBACK (pNewDT = 0.5) with finite stresses, as the oracle
    raises CutbackRequest(.., 0.5); with psi > 0 the same state returns to the apex."""
    from edelweissfe_amd import _lib
    from edelweissfe_amd.device import DeviceSystem

    for psi, want in ((0.0, _lib.EFE_CUTBACK), (10.0, _lib.EFE_OK)):
        pm = device_model((2, 2, 2), (1.0, 1.0, 1.0), "C3D8", "druckerprager", [30000.0, 0.2, 30.0, psi, 10.0, 0.0])
        sysm = DeviceSystem(pm.coordinates, pm.connectivity, "C3D8", "druckerprager", pm.material["properties"])
        dU = torch.as_tensor((1e-2 * pm.coordinates).ravel(), device="cuda")  # eps = 1e-2 I
        P, F = sysm.zeros(), sysm.zeros()
        rc, pnew = sysm.compute_elements(dU, P, F)
        assert rc == want
        assert pnew == (0.5 if want == _lib.EFE_CUTBACK else 1.0)
        st = sysm.qp_state(1)
        assert np.isfinite(st).all() and torch.isfinite(P).all()
        sysm.close()

"""ElementResultCollector on the device: the interface of
edelweissfe/utils/elementresultcollector.pyx:37-132, with the gather done by one kernel
(efe_collect_results) into a device array instead of a pointer loop on the host.

Results follow DisplacementElement.getResultArray (element.py:386-409): "stress" and
"strain" (6 values each), "materialstate" (the material's state variables) and the
material's own result "kappa" (vonmises.py:281-299; also the Drucker-Prager hardening
variable here).  Names are case insensitive, as the element's CaseInsensitiveDict.  The
values are those of the accepted (reference) state, as the element's views are.
"""

import ctypes

import numpy as np
import torch

from .._lib import check

_RESULTS = {"stress": (0, 6), "strain": (6, 6)}


class ElementResultCollector:
    def __init__(self, system, elements, quadraturePoints, result):
        """system: the DeviceSystem holding the elements' state; elements: element indices
        (None: all, in order); quadraturePoints: an iterable of QP numbers; result: name."""
        self.system = system
        self.elements = None if elements is None else np.ascontiguousarray(elements, dtype=np.int32)
        self.quadraturePoints = np.ascontiguousarray(list(quadraturePoints), dtype=np.int32)
        name = result.lower()
        nmat = system.n_state - 12
        if name in _RESULTS:
            self.first, self.size = _RESULTS[name]
        elif name == "materialstate" and nmat > 0:
            self.first, self.size = 12, nmat
        elif name == "kappa" and system.material in ("vonmises", "druckerprager"):
            self.first, self.size = 12, 1
        else:
            raise Exception("This result doesn't exist for the current material.")
        nEls = system.n_elements if self.elements is None else self.elements.size
        self.nGauss = self.quadraturePoints.size
        self.resultsTable = torch.empty((nEls, self.nGauss, self.size), dtype=torch.float64, device=system.device)

    def update(self):
        """Update all results (one device gather)."""
        s = self.system
        el = self.elements
        check(s.lib.efe_collect_results(s.ctx, 0, el.ctypes.data if el is not None else None,
                                        int(el.size) if el is not None else 0, self.quadraturePoints.ctypes.data,
                                        int(self.nGauss), int(self.first), int(self.size),
                                        ctypes.c_void_p(self.resultsTable.data_ptr())), "efe_collect_results")

    def getCurrentResults(self):
        """Update and get current results: [elements, quadrature points, result] on the device,
        [elements, result] for a single quadrature point (the reference's shapes).  The
        caller copies it if persistent results are needed."""
        self.update()
        if self.nGauss == 1:
            return self.resultsTable.reshape(self.resultsTable.shape[0], -1)
        return self.resultsTable

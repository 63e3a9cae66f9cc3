"""GPUElementLoop: the element loop of NISTParallelForMarmotElements on the GPU.

The reference's parallel solver for native elements (solvers/nonlinearimplicitstaticparallel.pyx:65-167)
keeps the DofManager, the VIJ system matrix, the CSR conversion and the linear solve on the host
and calls MarmotElement::computeYourself per element (marmotelement/element.pxd:97-103) in an
OpenMP loop.  This class is that loop's GPU twin behind the same hook,
``computeElements(elements, Un1, dU, P, K, F, timeStep) -> (P, K, F)``
(nonlinearimplicitstatic.py:794-849): one efe_element_compute call evaluates every element on
the device with the native ABI's layouts (DOF increments per element, caller-owned QP state,
Pe, Ke column-major); then, as the reference's serial tail does, Pe is summed into P and |Pe|
into F in element order, and Ke is written into the element's VIJ slice.

The element state is owned here (device tensors [E][nQP][nState], the reference and the
temporary set of MarmotElementWrapper.setMaterial, element.pyx:188-198): every call starts from
the reference set, acceptLastState() makes the temporary set the reference one
(models/femodel.py:245-258 -> element.acceptLastState).
"""

import numpy as np
import torch

from ..exceptions import CutbackRequest


class GPUElementLoop:
    def __init__(self, system, dofMap, vijOffsets=None):
        """system: the DeviceSystem of the elements (same element order); dofMap [E][nDof]: the
        DOF indices of every element, node-major (DofManager.idcsOfElementsInDofVector,
        dofmanager.py:455-471); vijOffsets [E]: first index of each element's nDof x nDof slice
        of the VIJ values (default: slices back to back in element order, dofmanager.py:522-557)."""
        self.system = system
        self.dofMap = np.ascontiguousarray(dofMap, dtype=np.int64)
        E, nd = self.dofMap.shape
        self.nElements, self.nDof = E, nd
        self.vijOffsets = (np.arange(E, dtype=np.int64) * nd * nd if vijOffsets is None
                           else np.asarray(vijOffsets, dtype=np.int64))
        dev = system.device
        self._map = torch.as_tensor(self.dofMap, device=dev)
        self._elements = np.arange(E, dtype=np.int32)
        shape = (E, system.n_qp, system.n_state)
        self.stateRef = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.stateTemp = torch.zeros(shape, dtype=torch.float64, device=dev)
        self.Pe = torch.empty((E, nd), dtype=torch.float64, device=dev)
        self.Ke = torch.empty((E, nd, nd), dtype=torch.float64, device=dev)
        # VIJ destination of every Ke entry in column-major order
        self._vijIdx = (self.vijOffsets[:, None] + np.arange(nd * nd)[None, :]).ravel()

    def computeElements(self, elements, Un1, dU, P, K, F, timeStep):
        """The hook of nonlinearimplicitstatic.py:794-849 on host DofVectors (numpy [nDof]) and
        the VIJ value vector K; raises CutbackRequest when a material asks for a smaller increment
        (as nonlinearimplicitstaticparallel.pyx:156-158)."""
        dev = self.system.device
        dUd = torch.as_tensor(np.asarray(dU, dtype=np.float64), device=dev)
        dQ = dUd[self._map].contiguous()
        Ped, Ked = self.compute(dQ, timeStep)
        Pe = Ped.cpu().numpy()
        P[:] = 0.0
        F[:] = 0.0
        np.add.at(P, self.dofMap.ravel(), Pe.ravel())
        np.add.at(F, self.dofMap.ravel(), np.abs(Pe).ravel())
        K[self._vijIdx] = Ked.cpu().numpy().ravel()  # [e][col][row]: column-major per element
        return P, K, F

    def compute(self, dQ, timeStep=None, formK=True):
        """efe_element_compute on device arrays: dQ [E][nDof]; returns (Pe, Ke) device tensors."""
        t = (0.0, 0.0) if timeStep is None else (timeStep.stepTime, timeStep.totalTime)
        dT = 0.0 if timeStep is None else timeStep.timeIncrement
        pnew = self.system.element_compute(self._elements, dQ, self.stateRef, self.stateTemp, self.Pe,
                                           self.Ke if formK else None, time=t, dT=dT)
        if pnew < 1.0:
            raise CutbackRequest("An element requests for a cutback", pnew)
        return self.Pe, self.Ke

    def acceptLastState(self):
        self.stateRef, self.stateTemp = self.stateTemp, self.stateRef


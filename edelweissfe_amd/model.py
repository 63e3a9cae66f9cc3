"""Model tree of one homogeneous hexahedral element set.

Plays the part of edelweissfe/models/femodel.py:45-298 for the hot path:
node coordinates, one element type with one material (optionally with
per-element properties, as sections/sectionbase.py:114-128 allow), node and
element sets, the model time and the converged node-field values U / P / dU
(nonlinearimplicitstatic.py:288-291).  DOF numbering follows
dofmanager.py:259-300 for a single displacement field: DOF 3 p + c for the node
at position p.
"""

import numpy as np


class FEModel:
    def __init__(self, coordinates, elementType, connectivity, nodeSets, elementSets, material=None, surfaces=None):
        self.coordinates = np.ascontiguousarray(coordinates, dtype=np.float64)
        self.elementType = elementType.upper()
        self.connectivity = np.ascontiguousarray(connectivity, dtype=np.int64)
        self.nodeSets = dict(nodeSets)
        self.elementSets = dict(elementSets)
        # element-based surfaces {name: {face ID (1..6): element indices}} (femodel.py surfaces)
        self.surfaces = dict(surfaces or {})
        self.material = material
        self.time = 0.0
        self.nDof = 3 * self.coordinates.shape[0]
        self.U = np.zeros(self.nDof)
        self.P = np.zeros(self.nDof)
        self.dU = np.zeros(self.nDof)
        self._qpState = None
        self.deviceSystem = None  # holds the QP state on the GPU after a device solve

    @property
    def qpState(self):
        """QP state [element][QP][stress 6, strain 6, material state] (element.py:222-236),
        copied from the device on first access after a solve."""
        if self._qpState is None and self.deviceSystem is not None:
            self._qpState = self.deviceSystem.qp_state(0)
        return self._qpState

    @qpState.setter
    def qpState(self, value):
        self._qpState = value
        self.deviceSystem = None

    @property
    def nNodes(self):
        return self.coordinates.shape[0]

    @property
    def nElements(self):
        return self.connectivity.shape[0]

    def assignMaterial(self, name, properties):
        """Material by name (linearelastic, vonmises, druckerprager) and properties:
        one row for all elements or one row per element."""
        self.material = {"name": name.lower(), "properties": np.asarray(properties, dtype=np.float64)}

    def dofsOfNodeSet(self, nSet, components=(0, 1, 2)):
        nodes = self.nodeSets[nSet]
        return (3 * np.asarray(nodes)[:, None] + np.asarray(components)[None, :]).ravel()

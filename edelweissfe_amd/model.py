"""Model tree of one hexahedral element set.

Plays the part of edelweissfe/models/femodel.py:45-298 for the hot path: node coordinates,
one element type with one material or several (sections, sections/base/sectionbase.py:86-128,
and per-element properties as materialParameterFromField gives them, :114-128), node and
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

    def assignMaterial(self, name, properties, elementSet=None):
        """Material by name (linearelastic, vonmises, druckerprager) and properties (one row, or
        one row per element it covers) on every element (`name` may also be one name per element),
        or on the elements of `elementSet` (a
        set name or element indices): a section (sections/base/sectionbase.py:86-128), a later
        assignment replacing an earlier one on the same elements.  With several materials,
        model.material holds one name and one row of 6 properties per element."""
        props = np.asarray(properties, dtype=np.float64)
        if elementSet is None and not isinstance(name, str):  # a name per element
            names = np.array([str(v).lower() for v in np.asarray(name).ravel()], dtype=object)
            if names.size != self.nElements:
                raise ValueError("one material name per element needed")
            rows = np.zeros((self.nElements, 6))
            p = np.atleast_2d(props)
            rows[:, :p.shape[1]] = p if p.shape[0] == self.nElements else np.repeat(p, self.nElements, axis=0)
            self.material = {"name": names, "properties": rows}
            return
        if elementSet is None:
            self.material = {"name": name.lower(), "properties": props}
            return
        E = self.nElements
        els = np.asarray(self.elementSets[elementSet] if isinstance(elementSet, str) else elementSet, dtype=np.int64)
        names, rows = self._perElementMaterial()
        r = np.atleast_2d(props)
        if r.shape[0] == 1:
            r = np.repeat(r, els.size, axis=0)
        if r.shape[0] != els.size:
            raise ValueError("one row of properties, or one per element of the set")
        names[els] = name.lower()
        rows[els] = 0.0
        rows[els, :r.shape[1]] = r
        if (names == names[0]).all() and names[0]:
            n = 2 if names[0] == "linearelastic" else 6
            self.material = {"name": str(names[0]), "properties": rows[:, :n].copy() if E else rows}
        else:
            self.material = {"name": names, "properties": rows}

    def _perElementMaterial(self):
        E = self.nElements
        names = np.full(E, "", dtype=object)
        rows = np.zeros((E, 6))
        if self.material is not None:
            names[:] = self.material["name"]
            p = np.atleast_2d(self.material["properties"])
            rows[:, :p.shape[1]] = p if p.shape[0] == E else np.repeat(p, E, axis=0)
        return names, rows

    def materialNames(self):
        """The material names present (one for a homogeneous model)."""
        if self.material is None:
            return []
        return sorted(set(np.atleast_1d(self.material["name"]).tolist()))

    def dofsOfNodeSet(self, nSet, components=(0, 1, 2)):
        nodes = self.nodeSets[nSet]
        return (3 * np.asarray(nodes)[:, None] + np.asarray(components)[None, :]).ravel()

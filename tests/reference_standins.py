"""Stand-ins shaped like the reference's own model tree and step actions, carrying the
attributes and methods the reference's NIST reads (for driving the drop-in solver in tests):

* Node                 points/node.py:30-60 (label, coordinates, fields)
* NodeField            fields/nodefield.py:83-120 (nodes, values by entry name [nNodes, dim])
* VonMisesMaterial,    materials/vonmises/vonmises.py:105-124,
  LinearElasticMaterial materials/linearelastic/linearelastic.py:59-61 (parameters as attributes)
* DisplacementElement  elements/displacementelement/element.py:140-240 (nNodes, nodes, _nInt,
                       nSpatialDimensions, material, _stateVarsRef)
* FEModel              models/femodel.py:45-80 (nodes, elements, nodeSets, elementSets,
                       nodeFields, surfaces, time)
* Dirichlet            stepactions/dirichlet.py:53-112 (nSet of nodes, components, delta,
                       getDelta(timeStep), active, applyAtStepEnd(model))
* NodeForces           stepactions/nodeforces.py:53-152 (nodeSet, getCurrentLoad per node)
* BodyForce            stepactions/bodyforce.py:49-105 (elementSet, getCurrentLoad)
* Geostatic            stepactions/geostatic.py:51-101 (geostaticElements, geostaticDefinition,
                       active, applyAtStepEnd resets the displacement field)
* Options              stepactions/options.py (options dict)
"""

import numpy as np

_LIB = {"C3D8": (8, 8), "C3D20": (20, 27), "C3D20R": (20, 8)}


class Node:
    def __init__(self, label, coordinates):
        self.label = label
        self.coordinates = np.asarray(coordinates, dtype=float)
        self.fields = {}


class NodeField:
    def __init__(self, name, dimension, nodes):
        self.name = name
        self.dimension = dimension
        self.nodes = list(nodes)
        self._values = {k: np.zeros((len(self.nodes), dimension)) for k in ("U", "P", "dU")}

    def __getitem__(self, key):
        return self._values[key]


class VonMisesMaterial:
    def __init__(self, p):
        self._E, self._v, self.yieldStress, self.HLin, self.deltaYieldStress, self.delta = (float(v) for v in p)

    def getNumberOfRequiredStateVars(self):
        return 1


class LinearElasticMaterial:
    def __init__(self, p):
        self._E, self._v = float(p[0]), float(p[1])

    def getNumberOfRequiredStateVars(self):
        return 0


class DisplacementElement:
    def __init__(self, elementType, elNumber):
        self.elNumber = elNumber
        self.nNodes, self._nInt = _LIB[elementType]
        self.nSpatialDimensions = 3
        self.nodes = None
        self.material = None

    def setNodes(self, nodes):
        self.nodes = nodes

    def setMaterial(self, material):
        self.material = material
        self._stateVarsRef = np.zeros((self._nInt, 12 + material.getNumberOfRequiredStateVars()))


class FEModel:
    def __init__(self):
        self.time = 0.0
        self.nodes, self.elements, self.nodeSets, self.elementSets = {}, {}, {}, {}
        self.nodeFields, self.surfaces = {}, {}
        self.domainSize = 3


class Dirichlet:
    def __init__(self, nSet, values, amplitude=None):
        self.nSet = nSet
        self._components = sorted(values)
        self.delta = np.tile([float(values[c]) for c in self._components], (len(nSet), 1))
        self.amplitude = amplitude or (lambda x: x)
        self.active = True

    @property
    def components(self):
        return self._components

    def getDelta(self, timeStep):
        if self.active:
            return self.delta * (self.amplitude(timeStep.stepProgress)
                                 - self.amplitude(timeStep.stepProgress - timeStep.stepProgressIncrement))
        return self.delta * 0.0

    def applyAtStepEnd(self, model):
        self.active = False


class NodeForces:
    def __init__(self, nSet, vector):
        self._nSet = nSet
        self.nodeForcesStepStart = np.zeros((len(nSet), 3))
        self.nodeForcesDelta = np.tile(np.asarray(vector, dtype=float), (len(nSet), 1))
        self.amplitude = lambda x: x
        self._idle = False

    @property
    def nodeSet(self):
        return self._nSet

    def getCurrentLoad(self, timeStep):
        if self._idle:
            return self.nodeForcesStepStart
        return self.nodeForcesStepStart + self.nodeForcesDelta * self.amplitude(timeStep.stepProgress)

    def applyAtStepEnd(self, model, stepMagnitude=None):
        if not self._idle:
            self.nodeForcesStepStart += self.nodeForcesDelta * self.amplitude(1.0)
            self.nodeForcesDelta[:] = 0
            self._idle = True


class BodyForce:
    def __init__(self, elSet, vector):
        self._elSet = elSet
        self._forceAtStepStart = 0.0
        self._delta = np.asarray(vector, dtype=float)
        self._amplitude = lambda x: x
        self._idle = False

    @property
    def elementSet(self):
        return self._elSet

    def getCurrentLoad(self, timeStep):
        t = 1.0 if self._idle else timeStep.stepProgress
        return self._forceAtStepStart + self._delta * self._amplitude(t)

    def applyAtStepEnd(self, model, stepMagnitude=None):
        if not self._idle:
            self._forceAtStepStart += self._delta * self._amplitude(1.0)
            self._delta = 0
            self._idle = True


class Geostatic:
    def __init__(self, elSet, p1, h1=1.0, p2=None, h2=-1.0, xLateral=1.0, zLateral=1.0):
        self.geostaticElements = elSet
        self.geostaticDefinition = np.array([p1, h1, p1 if p2 is None else p2, h2, xLateral, zLateral], dtype=float)
        self.active = True

    def applyAtStepEnd(self, model, stepMagnitude=None):
        if not self.active:
            return
        model.nodeFields["displacement"]["U"][:] = 0
        self.active = False


class Options:
    def __init__(self, **options):
        self.options = {k: str(v) for k, v in options.items()}  # as read from an input file


def reference_model(coords, conn, el_type, material_of_element, node_order=None, node_sets=None, element_sets=None):
    """A reference-shaped model of a mesh: nodes labelled 1.., the displacement node field in
    `node_order` (the DOF order of the reference's DofManager), elements 1.. in connectivity
    order with the material objects given per element."""
    m = FEModel()
    nodes = [Node(i + 1, c) for i, c in enumerate(coords)]
    for n in nodes:
        m.nodes[n.label] = n
    order = np.arange(len(nodes)) if node_order is None else np.asarray(node_order)
    m.nodeFields["displacement"] = NodeField("displacement", 3, [nodes[i] for i in order])
    for e, row in enumerate(conn):
        el = DisplacementElement(el_type, e + 1)
        el.setNodes([nodes[i] for i in row])
        el.setMaterial(material_of_element[e])
        m.elements[e + 1] = el
    els = list(m.elements.values())
    for k, v in (node_sets or {}).items():
        m.nodeSets[k] = [nodes[i] for i in v]
    for k, v in (element_sets or {}).items():
        m.elementSets[k] = [els[i] for i in v]
    return m

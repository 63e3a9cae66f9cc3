"""Step actions of the hot path, same meaning as the reference's:

* Dirichlet   edelweissfe/stepactions/dirichlet.py:53-158: prescribed values on
              components of a node set; the increment of an increment is
              value * (amp(p) - amp(p - dp)); inactive (zero) after the step.
* NodeForces  edelweissfe/stepactions/nodeforces.py:53-152: the current load is
              start + delta * amp(p).
* BodyForce   edelweissfe/stepactions/bodyforce.py: force per volume on an element
              set, start + delta * amp(p).
* DistributedLoad edelweissfe/stepactions/distributedload.py:52-110: a load of a given
              type ("surface traction": force per area, a vector) on a surface
              {face ID: element set}; the current load is start + delta * amp(p).
* Geostatic   edelweissfe/stepactions/geostatic.py:51-101: geostatic initial stress
              [p1, h1, p2, h2, xLateral, zLateral] on an element set, set into the
              reference state at every iteration start; displacements are reset at the
              end of the step, after which the action is inactive.
amp(t) = t unless an amplitude callable is given (the reference's f(t)).
"""

import numpy as np


def _identity(t):
    return t


class Dirichlet:
    def __init__(self, name, nSet, values, amplitude=None):
        """values: {component index (0..2): prescribed value over the step}."""
        self.name = name
        self.nSet = nSet
        self.components = sorted(values)
        self.values = np.array([float(values[c]) for c in self.components])
        self.amplitude = amplitude or _identity
        self.active = True

    def indices(self, model):
        nodes = np.asarray(model.nodeSets[self.nSet])
        return (3 * nodes[:, None] + np.asarray(self.components)[None, :]).ravel()

    def incrementFactor(self, timeStep):
        """amp(p) - amp(p - dp) of this increment, 0 once the action is inactive: the
        increment of every prescribed value is value * incrementFactor."""
        if not self.active:
            return 0.0
        a = self.amplitude
        return a(timeStep.stepProgress) - a(timeStep.stepProgress - timeStep.stepProgressIncrement)

    def getDelta(self, model, timeStep):
        n = len(model.nodeSets[self.nSet])
        return (np.tile(self.values, (n, 1)) * self.incrementFactor(timeStep)).ravel()

    def applyAtStepEnd(self, model):
        self.active = False


class NodeForces:
    def __init__(self, name, nSet, vector, amplitude=None):
        self.name = name
        self.nSet = nSet
        self.delta = np.asarray(vector, dtype=float)
        self.start = np.zeros(3)
        self.amplitude = amplitude or _identity
        self.idle = False

    def getCurrentLoad(self, timeStep):
        if self.idle:
            return self.start
        return self.start + self.delta * self.amplitude(timeStep.stepProgress)

    def applyAtStepEnd(self, model):
        if not self.idle:
            self.start = self.start + self.delta * self.amplitude(1.0)
            self.delta = np.zeros(3)
            self.idle = True


class BodyForce:
    def __init__(self, name, elSet, vector, amplitude=None):
        self.name = name
        self.elSet = elSet
        self.delta = np.asarray(vector, dtype=float)
        self.start = np.zeros(3)
        self.amplitude = amplitude or _identity
        self.idle = False

    def getCurrentLoad(self, timeStep):
        t = 1.0 if self.idle else timeStep.stepProgress
        return self.start + self.delta * self.amplitude(t)

    def applyAtStepEnd(self, model):
        if not self.idle:
            self.start = self.start + self.delta * self.amplitude(1.0)
            self.delta = np.zeros(3)
            self.idle = True


class DistributedLoad:
    def __init__(self, name, surface, magnitude, loadType="surface traction", amplitude=None):
        """surface: the model's surface name; magnitude: the load vector over the step."""
        if loadType.replace(" ", "").lower() != "surfacetraction":
            raise NotImplementedError("distributed load type {:}".format(loadType))
        self.name = name
        self.surface = surface
        self.loadType = loadType
        self.delta = np.asarray(magnitude, dtype=float)
        self.start = np.zeros(3)
        self.amplitude = amplitude or _identity
        self.idle = False

    def getCurrentLoad(self, timeStep):
        t = 1.0 if self.idle else timeStep.stepProgress
        return self.start + self.delta * self.amplitude(t)

    def applyAtStepEnd(self, model):
        if not self.idle:
            self.start = self.start + self.delta * self.amplitude(1.0)
            self.delta = np.zeros(3)
            self.idle = True


class Geostatic:
    def __init__(self, name, p1, h1=1.0, p2=None, h2=-1.0, xLateral=1.0, zLateral=1.0, elSet="all"):
        self.name = name
        self.elSet = elSet
        self.definition = np.array([p1, h1, p1 if p2 is None else p2, h2, xLateral, zLateral], dtype=float)
        self.active = True

    def applyAtIterationStart(self, system, model):
        if not self.active:
            return
        els = model.elementSets.get(self.elSet) if self.elSet != "all" else None
        system.set_geostatic(self.definition, els)

    def applyAtStepEnd(self, model):
        if not self.active:
            return
        model.U[:] = 0.0
        self.active = False

"""AdaptiveStep with the option names and defaults of
edelweissfe/steps/adaptivestep.py:56-113 and its step-action dictionary
(actions["dirichlet"], ["nodeforces"], ["distributedload"], ["bodyforce"], ["geostatic"],
["options"]["NISTSolver"])."""

from .timestepper import AdaptiveTimeStepper


class AdaptiveStep:
    defaultStartInc = 1.0
    defaultMaxInc = 1.0
    defaultMinInc = 1e-4
    defaultMaxNumInc = 1000
    defaultMaxIter = 10
    defaultCriticalIter = 5
    defaultMaxGrowingIter = 10
    defaultCutbackFactor = 0.25

    def __init__(self, number, model, actions=None, **kwargs):
        self.number = number
        self.model = model
        self.length = kwargs.get("stepLength", 1.0)
        self.startIncrementSize = kwargs.get("startInc", self.defaultStartInc)
        self.maxIncrementSize = kwargs.get("maxInc", self.defaultMaxInc)
        self.minIncrementSize = kwargs.get("minInc", self.defaultMinInc)
        self.maxNumberIncrements = kwargs.get("maxNumInc", self.defaultMaxNumInc)
        self.maxIter = kwargs.get("maxIter", self.defaultMaxIter)
        self.criticalIter = kwargs.get("criticalIter", self.defaultCriticalIter)
        self.maxGrowIter = kwargs.get("maxGrowIter", self.defaultMaxGrowingIter)
        self.cutbackFactor = kwargs.get("cutbackFactor", self.defaultCutbackFactor)
        self.actions = {"dirichlet": {}, "nodeforces": {}, "distributedload": {}, "bodyforce": {}, "geostatic": {},
                        "options": {}}
        for k, v in (actions or {}).items():
            self.actions[k] = v
        self.incrementGenerator = AdaptiveTimeStepper(model.time, self.length, self.startIncrementSize,
                                                      self.maxIncrementSize, self.minIncrementSize,
                                                      self.maxNumberIncrements)

    def getTimeStep(self):
        return self.incrementGenerator.generateTimeStep()

    def discardAndChangeIncrement(self, cutbackFactor):
        return self.incrementGenerator.discardAndChangeIncrement(cutbackFactor)

    def preventIncrementIncrease(self):
        return self.incrementGenerator.preventIncrementIncrease()

"""Increment generation with the semantics of
edelweissfe/timesteppers/adaptivetimestepper.py:73-183: a zero-size first
increment, clamping to the remainder, growth by 1.1 after three good increments
(capped at maxInc), cutbacks, and the same failure exceptions.  TimeStep carries
the fields of edelweissfe/timesteppers/timestep.py:29-66."""

from dataclasses import dataclass

from ..exceptions import ReachedMaxIncrements, ReachedMinIncrementSize


@dataclass
class TimeStep:
    number: int
    stepProgressIncrement: float
    stepProgress: float
    timeIncrement: float
    stepTime: float
    totalTime: float


class AdaptiveTimeStepper:
    def __init__(self, currentTime, stepLength, startIncrement, maxIncrement, minIncrement, maxNumberIncrements,
                 increaseFactor=1.1):
        self.currentTime = currentTime
        self.stepLength = stepLength
        self.maxIncrement = maxIncrement
        self.minIncrement = minIncrement
        self.maxNumberIncrements = maxNumberIncrements
        self.increaseFactor = increaseFactor
        self.increment = min(startIncrement, maxIncrement)
        self.finishedStepProgress = 0.0
        self.incrementCounter = 0
        self.nPassedGoodIncrements = 0
        self.allowedToIncreasedNext = True

    def generateTimeStep(self):
        while self.finishedStepProgress < 1.0 - 1e-15:
            remainder = 1.0 - self.finishedStepProgress
            if remainder < self.increment:
                self.increment = remainder
            inc = self.increment if self.incrementCounter > 0 else 0.0
            self.finishedStepProgress += inc
            inStep = self.stepLength * self.finishedStepProgress
            yield TimeStep(self.incrementCounter, inc, self.finishedStepProgress, self.stepLength * inc, inStep,
                           self.currentTime + inStep)
            if self.incrementCounter > self.maxNumberIncrements:
                raise ReachedMaxIncrements()
            if self.nPassedGoodIncrements >= 3 and self.allowedToIncreasedNext:
                self.increment = min(self.increment * self.increaseFactor, self.maxIncrement)
            self.allowedToIncreasedNext = True
            self.incrementCounter += 1
            self.nPassedGoodIncrements += 1

    def preventIncrementIncrease(self):
        self.allowedToIncreasedNext = False

    def reduceNextIncrement(self, scaleFactor):
        if self.increment == self.minIncrement:
            raise ReachedMinIncrementSize()
        self.increment = min(max(self.increment * scaleFactor, self.minIncrement), self.maxIncrement)

    def discardAndChangeIncrement(self, scaleFactor):
        if self.incrementCounter == 0:
            raise ReachedMinIncrementSize()
        self.finishedStepProgress -= self.increment
        self.incrementCounter -= 1
        self.nPassedGoodIncrements = 0
        self.reduceNextIncrement(scaleFactor)

"""Output-manager interface (edelweissfe/outputmanagers/base/outputmanagerbase.py:39-137):
constructed with (name, model, fieldOutputController, journal, plotter), configured by
updateDefinition(**kwargs), then called at job, step and increment boundaries."""

from abc import ABC, abstractmethod


class OutputManagerBase(ABC):
    identification = "OutputManagerBase"

    @abstractmethod
    def __init__(self, name, model, fieldOutputController, journal, plotter):
        pass

    @abstractmethod
    def updateDefinition(self, **kwargs):
        pass

    def initializeJob(self):
        pass

    def initializeStep(self, step):
        pass

    @abstractmethod
    def finalizeIncrement(self, **kwargs):
        pass

    @abstractmethod
    def finalizeFailedIncrement(self, **kwargs):
        pass

    def finalizeStep(self):
        pass

    def finalizeJob(self):
        pass

"""Conditional stop (conditionalstop.py:49-90): after a converged increment, raise
ConditionalStop when any stop condition holds; the solver then ends the step as a success
(nonlinearimplicitstatic.py:323-325).  A condition is a callable of the model, the form the
reference's ``createModelAccessibleFunction`` compiles its expression strings into."""

from ..exceptions import ConditionalStop
from .base import OutputManagerBase


class OutputManager(OutputManagerBase):
    identification = "ConditionalStop"

    def __init__(self, name=None, model=None, fieldOutputController=None, journal=None, plotter=None):
        self.model = model
        self.journal = journal
        self.monitorJobs = []

    def updateDefinition(self, **kwargs):
        stop = kwargs["stop"]
        if not callable(stop):
            raise TypeError("stop must be a callable of the model")
        self.monitorJobs.append({"stop": stop})

    def finalizeIncrement(self, **kwargs):
        for job in self.monitorJobs:
            if job["stop"](self.model):
                raise ConditionalStop()

    def finalizeFailedIncrement(self, **kwargs):
        pass

"""Per-increment outputs of the Newton step: the reference's output-manager plugin interface
(edelweissfe/outputmanagers/base/outputmanagerbase.py) with the managers that log the step
itself (status file, compute times, conditional stop).  Solvers call finalizeIncrement /
finalizeFailedIncrement after every increment with the reference's keyword arguments
``statusInfoDict`` and ``currentComputingTimes`` (nonlinearimplicitstatic.py:207-315)."""

from .base import OutputManagerBase
from .computetimemonitor import OutputManager as ComputeTimeMonitor
from .conditionalstop import OutputManager as ConditionalStopManager
from .statusfile import OutputManager as StatusFile
from .timing import createTimingDict, timingTypes

__all__ = ["OutputManagerBase", "StatusFile", "ComputeTimeMonitor", "ConditionalStopManager", "createTimingDict",
           "timingTypes"]

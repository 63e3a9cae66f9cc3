"""Error contract of the reference (edelweissfe/utils/exceptions.py:36-73), same names and meaning."""


class StepFailed(Exception):
    """The computation of a step failed."""


class CutbackRequest(Exception):
    """The residual evaluation failed because the increment is too large."""

    def __init__(self, message, cutbackSize):
        super().__init__(message)
        self.cutbackSize = float(cutbackSize)


class ReachedMaxIterations(Exception):
    """The maximum number of nonlinear iterations was reached."""


class ReachedMaxIncrements(Exception):
    """The maximum number of increments within a step was reached."""


class ReachedMinIncrementSize(Exception):
    """The minimum increment size within a step was reached."""


class DivergingSolution(Exception):
    """The nonlinear iteration seems to diverge."""


class ConditionalStop(Exception):
    """The simulation stops successfully (a stop condition of an output manager holds)."""

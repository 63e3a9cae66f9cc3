"""Error contract of the reference (edelweissfe/utils/exceptions.py:36-73)."""


class StepFailed(Exception):
    pass


class CutbackRequest(Exception):
    def __init__(self, message, cutbackSize):
        super().__init__(message)
        self.cutbackSize = float(cutbackSize)


class ReachedMaxIterations(Exception):
    pass


class ReachedMaxIncrements(Exception):
    pass


class ReachedMinIncrementSize(Exception):
    pass


class DivergingSolution(Exception):
    pass

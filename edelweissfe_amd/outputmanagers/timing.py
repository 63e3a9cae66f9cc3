"""Timing categories of the Newton step, in the reference's order and spelling
(edelweissfe/config/timing.py:8-23): the keys of a solver's computationTimes, which the
compute-time monitor turns into per-increment columns."""

timingTypes = [
    "elements",
    "constraints",
    "distributed loads",
    "body forces",
    "dirichlet R",
    "CSR generation",
    "dirichlet K",
    "linear solve",
    "convergence check",
    "output",
]


def createTimingDict():
    return {k: 0.0 for k in timingTypes}

"""Solver registry beside the reference's (edelweissfe/config/solvers.py:48-76):
the same lookup by name, with the device solver registered as "NISTGPU"."""

import importlib

solverLibrary = {"NISTGPU": "nistgpu"}


def getSolverByName(name):
    module = importlib.import_module("edelweissfe_amd.solvers." + solverLibrary[name])
    return getattr(module, name)

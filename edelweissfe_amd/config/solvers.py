"""Solver registry beside the reference's (edelweissfe/config/solvers.py:48-76): the same
lookup by name.  "NISTGPU" is the device solver on this package's model tree;
"NISTGPUDropIn" takes the reference's own FEModel, step and step actions."""

import importlib

solverLibrary = {"NISTGPU": "nistgpu", "NISTGPUDropIn": "dropin"}


def getSolverByName(name):
    module = importlib.import_module("edelweissfe_amd.solvers." + solverLibrary[name])
    return getattr(module, name)

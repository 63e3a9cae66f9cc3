"""MI355X-native implementation of EdelweissFE's nonlinear implicit static Newton step.

Public pieces: ``generators.boxgen.generateBox`` (model), ``steps`` (AdaptiveStep,
Dirichlet, NodeForces, BodyForce), ``solvers.NISTGPU`` (the solver plugin),
``device.DeviceSystem`` (the C-ABI context) and ``build`` (compiles libefe_gpu.so).
"""

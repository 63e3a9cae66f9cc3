from .adaptivestep import AdaptiveStep  # noqa: F401
from .stepactions import BodyForce, Dirichlet, DistributedLoad, Geostatic, NodeForces  # noqa: F401
from .timestepper import AdaptiveTimeStepper, TimeStep  # noqa: F401

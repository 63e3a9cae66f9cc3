from .adaptivestep import AdaptiveStep  # noqa: F401
from .stepactions import BodyForce, Dirichlet, Geostatic, NodeForces  # noqa: F401
from .timestepper import AdaptiveTimeStepper, TimeStep  # noqa: F401

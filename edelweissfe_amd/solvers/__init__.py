from .dropin import NISTGPUDropIn  # noqa: F401
from .nistbatched import NISTBatched, NISTBatchedGroups  # noqa: F401
from .nistgpu import NISTGPU  # noqa: F401

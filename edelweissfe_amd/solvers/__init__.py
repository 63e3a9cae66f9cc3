from .nistgpu import NISTGPU  # noqa: F401

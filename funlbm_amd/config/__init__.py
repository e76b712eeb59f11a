from .base import (
    BaseConfig,
    Boundary,
    BoundaryCondition,
    BoundaryConfig,
    FileConfig,
    deep_get,
)

__all__ = [
    "BaseConfig",
    "Boundary",
    "BoundaryCondition",
    "BoundaryConfig",
    "FileConfig",
    "deep_get",
]

"""JSON configuration objects with the reference's schema and defaults.

Follows src/funlbm/config/base.py:9-195.  A value is read with
``deep_get(json, key) or default`` as the reference does, so falsy values
(0, 0.0, "") fall back to the default (src/funlbm/config/base.py:103).
"""

import json
import os
from enum import Enum
from typing import Any, Dict, Optional, Union


def deep_get(data: Any, key: str, default: Any = None) -> Any:
    """Dotted-path lookup in nested dicts (the reference's funutil.deep_get)."""
    if data is None:
        return default
    cur = data
    for part in str(key).split("."):
        if isinstance(cur, dict) and part in cur:
            cur = cur[part]
        else:
            return default
    return cur


class BoundaryCondition(Enum):
    """Boundary codes, values as src/funlbm/config/base.py:22-28."""

    PERIODICAL = 11000
    WALL = 1200
    WALL_WITH_SPEED = 1201
    FAR_FIELD = 1300
    NON_EQUILIBRIUM = 1400
    NON_EQUILIBRIUM_EXREAPOLATION = 1500
    FULL_DEVELOPMENT = 1600

    @classmethod
    def find(cls, code: Union[int, str]) -> "BoundaryCondition":
        """By value (int) or name (str); anything unknown is WALL
        (src/funlbm/config/base.py:30-45)."""
        try:
            if isinstance(code, int):
                return next(bc for bc in cls if bc.value == code)
            return cls[str(code)]
        except (StopIteration, KeyError):
            return cls.WALL


class BaseConfig:
    """src/funlbm/config/base.py:48-103."""

    def __init__(self, *args, **kwargs) -> None:
        self.expand: Dict[str, Any] = kwargs.copy()

    def _from_json(self, config_json: Dict[str, Any], **kwargs) -> None:
        pass

    def from_file(self, path: str) -> "BaseConfig":
        with open(path) as f:
            self.from_json(json.load(f))
        return self

    def from_json(self, config_json: Dict[str, Any], **kwargs) -> "BaseConfig":
        self.expand.update(kwargs)
        self.expand.update(config_json or {})
        self._from_json(config_json or {}, **kwargs)
        return self

    def get(self, key: str, default: Any = None) -> Any:
        return deep_get(self.expand, key) or default

    def to_json(self) -> Dict[str, Any]:
        return self.expand


class Boundary(BaseConfig):
    """One face (src/funlbm/config/base.py:106-130).  Extra keys such as the
    wall velocity ``uw`` of WALL_WITH_SPEED are read with ``get``."""

    def __init__(self, condition: BoundaryCondition = BoundaryCondition.WALL, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.condition: BoundaryCondition = condition
        self.poiseuille: Optional[Any] = None

    def is_condition(self, condition: BoundaryCondition) -> bool:
        return self.condition == condition

    def _from_json(self, config_json: Dict[str, Any], **kwargs) -> None:
        self.condition = BoundaryCondition.find(deep_get(config_json, "code") or "WALL")
        self.poiseuille = deep_get(config_json, "poiseuille")


FACE_NAMES = ("input", "output", "back", "forward", "bottom", "top")


class BoundaryConfig(BaseConfig):
    """The six faces, all WALL by default (src/funlbm/config/base.py:133-156)."""

    def __init__(self, *args, **kwargs) -> None:
        super().__init__(*args, **kwargs)
        self.input = Boundary(BoundaryCondition.WALL)
        self.output = Boundary(BoundaryCondition.WALL)
        self.back = Boundary(BoundaryCondition.WALL)
        self.forward = Boundary(BoundaryCondition.WALL)
        self.bottom = Boundary(BoundaryCondition.WALL)
        self.top = Boundary(BoundaryCondition.WALL)

    def _from_json(self, config_json: Dict[str, Any], **kwargs) -> None:
        for name in FACE_NAMES:
            getattr(self, name).from_json(deep_get(config_json, name) or {})

    def faces(self):
        return [getattr(self, name) for name in FACE_NAMES]


class FileConfig(BaseConfig):
    """Output directories (src/funlbm/config/base.py:159-195)."""

    def __init__(self, *args, **kwargs) -> None:
        super().__init__(*args, **kwargs)
        self.cache_dir: str = "./data"
        self.per_steps: int = 100

    def _path(self, name: str) -> str:
        path = os.path.join(self.cache_dir, name)
        os.makedirs(path, exist_ok=True)
        return path

    @property
    def vtk_path(self) -> str:
        return self._path("vtk")

    @property
    def tecplot_path(self) -> str:
        return self._path("tecplot")

    @property
    def checkpoint_path(self) -> str:
        return self._path("checkpoint")

    def _from_json(self, config_json: Dict[str, Any], **kwargs) -> None:
        self.cache_dir = deep_get(config_json, "cache_dir") or self.cache_dir
        self.per_steps = deep_get(config_json, "per_steps") or self.per_steps

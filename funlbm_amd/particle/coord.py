"""Particle coordinate frame (src/funlbm/particle/coord.py:13-122).

The reference keeps a float32 centre and rotation vector and rotates markers
with scipy on the host every step.  Here the frame lives in the particle's
double-precision state row on the GPU and the rotation runs in the pose
kernel; this class exposes the same attributes as views of that row.
"""

from typing import List, Optional

import numpy as np
import torch

from .. import _native as nat
from ..config.base import BaseConfig, deep_get


class CoordConfig(BaseConfig):
    """centre and rotation vector [alpha, beta, gamma], alpha = pi/2 by default
    (src/funlbm/particle/coord.py:13-43)."""

    def __init__(self, alpha: float = np.pi / 2, beta: float = 0, gamma: float = 0, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.center: List[float] = [0, 0, 0]
        self.alpha, self.beta, self.gamma = alpha, beta, gamma

    def _from_json(self, config_json: dict, *args, **kwargs) -> None:
        self.center = deep_get(config_json, "center") or self.center
        self.alpha = deep_get(config_json, "alpha") or self.alpha
        self.beta = deep_get(config_json, "beta") or self.beta
        self.gamma = deep_get(config_json, "gamma") or self.gamma


class Coordinate:
    """Frame of one particle, backed by its state row (set by bind())."""

    def __init__(self, config: Optional[CoordConfig] = None, device=None, *args, **kwargs):
        config = config or CoordConfig()
        self.device = device
        self.config = config
        # float32 initial values as the reference's tensors
        self.initial_center = np.array(config.center, dtype=np.float32).astype(np.float64)
        self.initial_angle = np.array([config.alpha, config.beta, config.gamma], dtype=np.float32).astype(
            np.float64
        )
        self._row = None

    def bind(self, row: torch.Tensor):
        self._row = row

    @property
    def center(self) -> torch.Tensor:
        return self._row[nat.P_CX : nat.P_CX + 3]

    @property
    def angle(self) -> torch.Tensor:
        return self._row[nat.P_ANGLE : nat.P_ANGLE + 3]

    @property
    def rotation(self) -> torch.Tensor:
        """Rotation matrix of `angle` (row-major 3x3, double)."""
        return self._row[nat.P_ROT : nat.P_ROT + 9].view(3, 3)

    def cul_point(self, points) -> torch.Tensor:
        """float32(R p) + float32(centre) (src/funlbm/particle/coord.py:73-94)."""
        p = torch.as_tensor(np.asarray(points, dtype=np.float64), device=self.device, dtype=torch.float64)
        r = (p.reshape(-1, 3) @ self.rotation.t()).to(torch.float32)
        out = r + self.center.to(torch.float32)
        return out.reshape(p.shape)

    def to_str(self):
        c = self.center.tolist()
        a = self.angle.tolist()
        return "center=" + ",".join(f"{i:.6f}" for i in c) + "\tangle=" + ",".join(f"{i:.6f}" for i in a)

    def to_json(self):
        return {"center": self.center.tolist(), "angle": self.angle.tolist()}

"""Sphere (src/funlbm/particle/base/sphere.py:9-53)."""

import math

import numpy as np

from ..markers import markers
from .base import Particle


class Sphere(Particle):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.r = self.config.get("r") or 10

    def _init(self, dx=1, *args, **kwargs):
        r = self.r
        self._lagrange = markers(self.config.markers, r, r, r, dx, device=self.device)
        self.mass = float(np.float32(4.0 / 3.0 * math.pi * r * r * r))
        self.area = float(np.float32(4.0 * math.pi * r * r))
        # float32(r^2 * float32(V) / 5) per axis, as the reference's numpy expression
        self.I = np.array([r * r] * 3, dtype=np.float64) * np.float32(self.mass) / 5.0
        self.I = self.I.astype(np.float32).astype(np.float64)

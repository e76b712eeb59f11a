"""Ellipsoid (src/funlbm/particle/base/ellipsoid.py:98-143)."""

import math

import numpy as np

from ..markers import markers
from .base import Particle


class Ellipsoid(Particle):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.ra, self.rb, self.rc = (
            self.config.get("a") or 10,
            self.config.get("b") or 10,
            self.config.get("c") or 10,
        )

    def _init(self, dx=1, *args, **kwargs):
        a, b, c = self.ra, self.rb, self.rc
        self._lagrange = markers(self.config.markers, a, b, c, dx, device=self.device)
        self.mass = float(np.float32(4.0 / 3.0 * math.pi * a * b * c))
        # the reference's area expression 4/3 pi (abc)^(2/3)
        self.area = float(np.float32(4.0 / 3.0 * math.pi * math.pow(a * b * c, 2.0 / 3.0)))
        self.I = np.array([b * c, a * c, a * b], dtype=np.float64) * np.float32(self.mass) / 5.0
        self.I = self.I.astype(np.float32).astype(np.float64)

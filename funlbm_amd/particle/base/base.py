"""Rigid particle (src/funlbm/particle/base/base.py:11-214).

Host side holds the configuration and the init-time constants (markers,
volume, area, inertia); the dynamic state (cx, cu, cw, cF, cT, the rotation
vector) is a double row of the swarm's state array on the GPU, and the marker
arrays (lx, lF, lu, lm, lrou) are views of the swarm's concatenated float
arrays.  The per-step arithmetic runs in HIP kernels over all particles.
"""

import numpy as np
import torch

from ... import _native as nat
from ...config.base import BaseConfig, deep_get
from ..coord import CoordConfig, Coordinate


class ParticleConfig(BaseConfig):
    """src/funlbm/particle/base/base.py:11-19; type defaults to "ellipsoid".
    Extension key: "markers", "cubes" (default: the reference's
    cube-intersection markers) or "fibonacci" (a golden-angle lattice at
    about dx spacing)."""

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.coord_config: CoordConfig = CoordConfig()
        self.type = "ellipsoid"
        self.markers = "cubes"

    def _from_json(self, config_json: dict, *args, **kwargs):
        self.coord_config.from_json(deep_get(config_json, "coord") or {})
        self.type = deep_get(config_json, "type") or self.type
        self.markers = deep_get(config_json, "markers") or self.markers


class Particle:
    def __init__(self, config: ParticleConfig = None, device=None, *args, **kwargs):
        self.config: ParticleConfig = config or ParticleConfig()
        self.device = device
        self.coord = Coordinate(config=self.config.coord_config, device=device)
        self.mass = None
        self.area = None
        self.I = None
        self.rou = None
        self._lagrange = None  # float32 [M, 3] host array
        self._row = None
        self.swarm = None
        self.index = None

    def _init(self, dx=1, *args, **kwargs):
        raise NotImplementedError()

    def init(self, *args, **kwargs):
        """Host constants (src/funlbm/particle/base/base.py:108-118)."""
        self.rou = float(self.config.get("rou", 1.0))
        self._init(*args, **kwargs)
        m = self._lagrange.shape[0]
        # lm = area / M in float32
        self.lm_value = np.float32(np.float32(self.area) / np.float32(m))

    @property
    def n_markers(self):
        return self._lagrange.shape[0]

    # views bound by the swarm -------------------------------------------
    def bind(self, swarm, index, row):
        self.swarm = swarm
        self.index = index
        self._row = row
        self.coord.bind(row)

    def _marker_view(self, arr):
        b, e = self.swarm.offsets_host[self.index], self.swarm.offsets_host[self.index + 1]
        return arr[b:e]

    @property
    def cx(self):
        return self._row[nat.P_CX : nat.P_CX + 3]

    @property
    def cu(self):
        return self._row[nat.P_CU : nat.P_CU + 3]

    @property
    def cw(self):
        return self._row[nat.P_CW : nat.P_CW + 3]

    @property
    def cF(self):
        return self._row[nat.P_CF : nat.P_CF + 3]

    @property
    def cT(self):
        return self._row[nat.P_CT : nat.P_CT + 3]

    @property
    def angle(self):
        return self._row[nat.P_ANGLE : nat.P_ANGLE + 3]

    @property
    def lx(self):
        return self._marker_view(self.swarm.lx)

    @property
    def lF(self):
        return self._marker_view(self.swarm.lF)

    @property
    def lu(self):
        return self._marker_view(self.swarm.lu)

    @property
    def lm(self):
        return self._marker_view(self.swarm.lm).unsqueeze(1)

    @property
    def lrou(self):
        return self._marker_view(self.swarm.lrou)

    # per-particle API -----------------------------------------------------
    def update_from_lar(self, dt: float, gl: float = 9.8, rouf: float = 1.0) -> None:
        """Force/torque and rigid-body update of this particle
        (src/funlbm/particle/base/base.py:120-151)."""
        if dt <= 0:
            raise ValueError("Time step must be positive")
        if rouf <= 0:
            raise ValueError("Fluid density must be positive")
        self.swarm.set_gravity_from_gl(self.index, gl)
        self.swarm.forces(dt, rouf, only=self.index)

    def update(self, dt=1.0):
        """Pose update (src/funlbm/particle/base/base.py:153-156)."""
        self.swarm.pose(only=self.index)

    def to_str(self, step=0, *args, **kwargs):
        return format_particle(self.to_json(step))

    def to_json(self, step=0, *args, **kwargs):
        lf = self.lF
        return {
            "m": float(self.mass),
            "cu": self.cu.tolist(),
            "cx": self.cx.tolist(),
            "cf": self.cF.tolist(),
            "lF": [float(lf.min()), float(lf.mean()), float(lf.max())],
            "cw": self.cw.tolist(),
            "coord": self.coord.to_json(),
        }


def format_particle(d: dict) -> str:
    """The reference's per-particle log text (src/funlbm/particle/base/base.py:161-171)
    from a to_json record."""

    def fmt(v):
        return ",".join(f"{i:.6f}" for i in v)

    res = f"m={d['m']:.2f}"
    res += "\tcu=" + fmt(d["cu"])
    res += "\tcx=" + fmt(d["cx"])
    res += "\tcf=" + fmt(d["cf"])
    res += "\tlF=" + fmt(d["lF"])
    res += "\tcw=" + fmt(d["cw"])
    c = d["coord"]
    res += "\tr=center=" + fmt(c["center"]) + "\tangle=" + fmt(c["angle"])
    return res

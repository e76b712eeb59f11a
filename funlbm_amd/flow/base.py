"""Lattice parameters, flow configuration and the Flow contract.

Mirrors src/funlbm/flow/base.py:15-217.  The Flow contract (init,
update_u_rou, cul_equ, cul_equ2, f_stream) and the tensors callers read
(f, u, rou, p, FOL, feq, tau, x) are kept; their storage lives on the GPU and
the arithmetic runs in HIP kernels (FlowD3 in flow/d3/base.py).
"""

import math
from typing import List

import numpy as np
import torch

from ..config.base import BaseConfig, BoundaryConfig, deep_get

# Direction tables in the reference's order (src/funlbm/flow/d3/d3q13.py:7-44,
# d3q15.py:7-48, d3q19.py:7-60, d3q27.py:7-77).  The order is part of the
# contract: it is the last axis of flow.f.
_Q19_E = [
    [0, 0, 0],
    [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
]
LATTICES = {
    "D3Q13": (
        [
            [0, 0, 0], [1, 1, 0], [1, -1, 0], [1, 0, 1], [1, 0, -1], [0, 1, 1], [0, 1, -1],
            [-1, -1, 0], [-1, 1, 0], [-1, 0, -1], [-1, 0, 1], [0, -1, -1], [0, -1, 1],
        ],
        [1.0 / 2] + [1.0 / 24] * 12,
    ),
    "D3Q15": (
        [
            [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
            [1, 1, 1], [-1, 1, 1], [1, -1, 1], [-1, -1, 1],
            [1, 1, -1], [-1, 1, -1], [1, -1, -1], [-1, -1, -1],
        ],
        [2.0 / 9] + [1.0 / 9] * 6 + [1.0 / 72] * 8,
    ),
    "D3Q19": (_Q19_E, [1.0 / 3] + [1.0 / 18] * 6 + [1.0 / 36] * 12),
    "D3Q27": (
        _Q19_E
        + [
            [1, 1, 1], [1, 1, -1], [1, -1, 1], [1, -1, -1],
            [-1, 1, 1], [-1, 1, -1], [-1, -1, 1], [-1, -1, -1],
        ],
        [8.0 / 27] + [2.0 / 27] * 6 + [1.0 / 54] * 12 + [1.0 / 216] * 8,
    ),
}


def reverse_map(e: List[List[int]]) -> List[int]:
    """Index of -e_k for every k (src/funlbm/flow/d3/d3q19.py:55-60)."""
    key = {tuple(v): i for i, v in enumerate(e)}
    return [key[tuple(-c for c in v)] for v in e]


class Param:
    """Lattice parameters (src/funlbm/flow/base.py:15-75): e [Q,3] and
    w [1,Q] as float32 tensors, cs = float32(sqrt(1/3)), vertex_reverse."""

    def __init__(self, e, w, vertex_reverse, device="cpu", *args, **kwargs):
        self.device = device
        self.vertex_reverse = list(vertex_reverse)
        self.e_list = [list(map(int, v)) for v in e]
        self.e = torch.tensor(e, device=device, dtype=torch.float32)
        self.w = torch.tensor([list(w)] if not isinstance(w[0], (list, tuple)) else w,
                              device=device, dtype=torch.float32)
        self.cs = torch.tensor(math.sqrt(1.0 / 3), device=device, dtype=torch.float32)
        self._vi = {}

    def eT(self) -> torch.Tensor:
        return self.e.t()

    def vertex_index(self, axis, value):
        """Directions with e[axis] == value (src/funlbm/flow/base.py:46-57)."""
        k = (int(axis), int(value))
        if k not in self._vi:
            self._vi[k] = [i for i, e in enumerate(self.e_list) if e[axis] == value]
        return self._vi[k]

    def index_reverse(self, index):
        if isinstance(index, int):
            return self.vertex_reverse[index]
        return [self.vertex_reverse[i] for i in index]

    @property
    def e_dim(self):
        return len(self.e_list)


class FlowConfig(BaseConfig):
    """Flow configuration (src/funlbm/flow/base.py:78-114).

    Extension keys of this package, all optional and absent from the
    reference's configs:
      storage      "f32" (default), "f16", "bf16" or "f64": population storage
                   of the fused timestep (f16/bf16 keep f - w rho0, scaled)
      dev_scale    scale of the f16/bf16 deviation storage (default 1024)
      rounding     "stochastic" (default for f16/bf16) or "nearest": rounding
                   of the f16/bf16 state at each step's store (stochastic:
                   unbiased, dithered by a hash of seed, step, node and
                   direction, so sub-ulp increments such as a weak body
                   force accumulate); float32/float64 states round to nearest
      seed         base seed of the stochastic rounding's dither (default 0)
      uzero_faces  list of face names on which u is zeroed after each moment
                   update (default: all six, the reference's behaviour)
      body_force   [gx, gy, gz] added to FOL everywhere (default none)
      gravity      [gx, gy, gz], the vector g of the particles' buoyancy force
                   (1 - rho_p/rho_f) V g; the reference's is [gl, 0, 0]
                   (src/funlbm/particle/base/base.py:138-142).  With this sign a
                   particle heavier than the fluid accelerates against g (the
                   reference's free-fall example sets gl < 0).
      ibm_delta    "cosine" (default, the reference's (1 + cos(pi r/2))/4,
                   src/funlbm/lbm/lbm3d.py:46-64) or "peskin4" (Peskin's
                   4-point kernel): the per-axis kernel of the
                   immersed-boundary interpolation and spreading weights
    """

    def __init__(self, *args, **kwargs):
        self.size: np.ndarray = np.zeros(3)
        self.param: dict = {}
        self.param_type: str = "D3Q19"
        self.boundary: BoundaryConfig = BoundaryConfig()
        self.gl: float = 0.0
        self.Re: float = 10
        self.mu: float = 10
        self.storage: str = "f32"
        self.dev_scale: float = 1024.0
        self.rounding = None
        self.seed: int = 0
        self.uzero_faces = None
        self.body_force = None
        self.gravity = None
        self.ibm_delta: str = "cosine"
        super().__init__(*args, **kwargs)

    def _from_json(self, config_json: dict, *args, **kwargs):
        self.size = np.array(deep_get(config_json, "size") or [100, 100, 100], dtype=int)
        self.param = deep_get(config_json, "param") or self.param
        self.boundary = BoundaryConfig().from_json(deep_get(config_json, "boundary") or {})
        self.param_type = deep_get(config_json, "param_type") or self.param_type
        self.Re = float(deep_get(self.param, "Re") or self.Re)
        self.mu = float(deep_get(self.param, "mu") or self.mu)
        self.gl = float(deep_get(self.param, "gl") or self.gl)
        self.storage = str(deep_get(config_json, "storage") or self.storage)
        self.dev_scale = float(deep_get(config_json, "dev_scale") or self.dev_scale)
        self.rounding = deep_get(config_json, "rounding") or self.rounding
        self.seed = int(deep_get(config_json, "seed") or self.seed)
        self.uzero_faces = deep_get(config_json, "uzero_faces")
        self.body_force = deep_get(config_json, "body_force")
        self.gravity = deep_get(self.param, "gravity") or deep_get(config_json, "gravity")
        self.ibm_delta = str(deep_get(config_json, "ibm_delta") or self.ibm_delta)


class Flow:
    """The flow contract (src/funlbm/flow/base.py:117-217)."""

    def __init__(self, param: Param, config: FlowConfig, device=None, *args, **kwargs):
        from ..base.work import device_detect

        self.device = device_detect(device)
        self.param: Param = param
        self.config: FlowConfig = config

    def init(self, *args, **kwargs) -> None:
        raise NotImplementedError("not implemented")

    def update_u_rou(self, *args, **kwargs) -> None:
        raise NotImplementedError("not implemented")

    def cul_equ(self, tau=None, *args, **kwargs) -> None:
        raise NotImplementedError("not implemented")

    def cul_equ2(self, *args, **kwargs) -> None:
        raise NotImplementedError("not implemented")

    def f_stream(self, *args, **kwargs) -> None:
        raise NotImplementedError("not implemented")

    def __repr__(self) -> str:
        return (
            f"{self.__class__.__name__}(size={tuple(self.config.size)}, "
            f"Re={self.config.Re}, mu={self.config.mu})"
        )

"""3-D flow on the GPU: the reference's FlowD3 contract backed by HIP kernels.

Mirrors src/funlbm/flow/d3/base.py:57-323.  Two representations of the same
state are kept:

explicit  the reference's tensors as they are between its method calls:
          f (pre-collision), u, rou, p, feq, FOL.  The methods cul_equ,
          f_stream, update_u_rou and cul_equ2 each run one HIP kernel on them.
fused     the post-collision populations f* that the next step streams, in the
          configured storage (f32, f16/bf16 deviation, f64), plus the FOL of
          the last two steps.  LBMD3.run_step advances this representation with
          one fused pull kernel per step.  Reading f, u, rou or p materialises
          the reference's tensors exactly (the same pull, forcing and moment
          arithmetic without the collision).

All device memory is allocated with torch; the kernels borrow raw pointers.
"""

import numpy as np
import torch

from ... import _native as nat
from ...config.base import FACE_NAMES, BoundaryCondition
from ..base import LATTICES, Flow, FlowConfig, Param, reverse_map

_F32 = torch.float32


def _uzero_mask(faces) -> int:
    if faces is None:
        return 0x3F
    m = 0
    for name in faces:
        m |= 1 << FACE_NAMES.index(name)
    return m


class FlowD3(Flow):
    """3-D lattice flow (src/funlbm/flow/d3/base.py:57-323)."""

    lattice_name = None

    def __init__(self, config: FlowConfig, device=None, *args, **kwargs):
        name = self.lattice_name
        if name is None:
            raise ValueError("FlowD3 needs a concrete lattice (FlowD3Q13/15/19/27)")
        e, w = LATTICES[name]
        param = Param(e=e, w=w, vertex_reverse=reverse_map(e), device="cpu")
        super().__init__(param=param, config=config, device=device)
        self.param.device = self.device
        self.param.e = self.param.e.to(self.device)
        self.param.w = self.param.w.to(self.device)
        self.param.cs = self.param.cs.to(self.device)
        self.q = len(e)
        self._initialised = False
        storage = getattr(config, "storage", "f32") or "f32"
        if storage not in nat.STORAGE_CODES:
            raise ValueError(f"unknown storage {storage!r}; expected one of {list(nat.STORAGE_CODES)}")
        self.storage = storage
        self.dev_scale = float(getattr(config, "dev_scale", 1024.0) or 1024.0)
        rounding = getattr(config, "rounding", None) or ("stochastic" if storage in ("f16", "bf16") else "nearest")
        if rounding not in nat.ROUNDING_CODES:
            raise ValueError(f"unknown rounding {rounding!r}; expected one of {list(nat.ROUNDING_CODES)}")
        if rounding == "stochastic" and storage not in ("f16", "bf16"):
            raise ValueError("stochastic rounding applies to f16/bf16 storage only")
        self.rounding = rounding
        self.seed = int(getattr(config, "seed", 0) or 0) & 0xFFFFFFFF
        # fused steps taken since init: the stochastic rounding's dither of a
        # step is drawn from step_seed(seed, steps_done)
        self.steps_done = 0
        # set by the run loop before a step whose statistics it will log: the
        # fused step then produces them in the same pass (stats())
        self.stats_next = False
        self._stats_valid = False
        self.uzero_mask = _uzero_mask(getattr(config, "uzero_faces", None))
        bf = getattr(config, "body_force", None)
        self.body_force = None if bf is None else [float(v) for v in bf]
        delta = getattr(config, "ibm_delta", None) or "cosine"
        if delta not in nat.DELTA_CODES:
            raise ValueError(f"unknown ibm_delta {delta!r}; expected one of {list(nat.DELTA_CODES)}")
        self.ibm_delta = delta

    # ------------------------------------------------------------------
    # lattice descriptors
    # ------------------------------------------------------------------
    def _make_lattice(self, storage_code: int) -> nat.Lattice:
        lat = nat.Lattice()
        lat.q = self.q
        lat.nx, lat.ny, lat.nz = self.nodes
        lat.storage = storage_code
        for i, b in enumerate(self.config.boundary.faces()):
            lat.bc[i] = b.condition.value
            if b.is_condition(BoundaryCondition.WALL_WITH_SPEED):
                uw = b.get("uw") or [0.0, 0.0, 0.0]
                uw = [float(v) for v in (uw.tolist() if hasattr(uw, "tolist") else uw)]
                for d in range(3):
                    lat.uw[i][d] = uw[d]
        lat.uzero_mask = self.uzero_mask
        lat.tau = float(np.float32(3 * self.config.mu + 0.5))
        lat.rho0 = 1.0
        lat.dev_scale = self.dev_scale
        if storage_code == nat.STORAGE_CODES[self.storage]:
            lat.rounding = nat.ROUNDING_CODES[self.rounding]
        lat.seed = self.seed
        lat.delta = nat.DELTA_CODES[self.ibm_delta]
        if self.body_force is not None:
            for d in range(3):
                lat.body_force[d] = self.body_force[d]
        return lat

    @property
    def has_neq(self) -> bool:
        """A NON_EQUILIBRIUM input/output face (explicit path only)."""
        return any(b.is_condition(BoundaryCondition.NON_EQUILIBRIUM) for b in self.config.boundary.faces())

    @property
    def has_moving_wall(self) -> bool:
        return any(b.is_condition(BoundaryCondition.WALL_WITH_SPEED) for b in self.config.boundary.faces())

    # ------------------------------------------------------------------
    # init (src/funlbm/flow/d3/base.py:72-101)
    # ------------------------------------------------------------------
    def init(self, *args, **kwargs):
        m, n, depth = self.config.size
        self.nodes = (int(m + 1), int(n + 1), int(depth + 1))
        self.n = self.nodes[0] * self.nodes[1] * self.nodes[2]
        self.lat32 = self._make_lattice(nat.STORE_F32)
        self.lat = self._make_lattice(nat.STORAGE_CODES[self.storage])
        dev = self.device
        q, N = self.q, self.n
        self.tau_value = float(self.lat32.tau)
        self._tau = torch.tensor(self.tau_value, device=dev, dtype=_F32)
        self._f = torch.empty((q, N), device=dev, dtype=_F32)
        self._ftmp = None
        self._feq = None
        self._u = torch.zeros((3, N), device=dev, dtype=_F32)
        self._rho = torch.ones((N,), device=dev, dtype=_F32)
        self._p = torch.zeros((N,), device=dev, dtype=_F32)
        self._fol = [torch.zeros((3, N), device=dev, dtype=_F32), None]
        self._fc = 0
        self._fol_live = [False, False]
        self._fol_tiled = [False, False]
        self._tiles = (None, None)
        self._state = [None, None]
        self._cur = 0
        self._rho_face = [None, None]
        self._mode = "explicit"
        self._mat_valid = True
        self._pristine = True
        self._x = None
        self._neq_profile = None
        self._neq_work = None
        if self.has_neq:
            from .neq import duct_profile

            if self.storage != "f32":
                raise NotImplementedError("NON_EQUILIBRIUM faces run on the explicit float32 path")
            self._neq_work = torch.zeros((1,), device=dev, dtype=_F32)
            self._neq_remu = float(np.float32(self.config.Re * self.config.mu))
            if self.config.boundary.input.is_condition(BoundaryCondition.NON_EQUILIBRIUM):
                prof = duct_profile(self.nodes[1], self.nodes[2])
                self._neq_profile = torch.from_numpy(prof.reshape(-1)).to(dev)
        # f = rho w with rho = 1 (exactly w_k in float32)
        nat.call("funlbm_equilibrium_state", self.lat32, None, None, nat.ptr(self._f), self._stream())
        self._initialised = True
        self.update_u_rou()
        self._pristine = True

    def _stream(self):
        return nat.stream_handle(self.device)

    # ------------------------------------------------------------------
    # explicit representation: the reference's methods
    # ------------------------------------------------------------------
    def _ensure_explicit(self):
        if self._mode == "fused":
            self._materialize()
            self._mode = "explicit"

    def _fol_ptr(self, idx):
        return nat.ptr(self._fol[idx]) if self._fol_live[idx] and self._fol[idx] is not None else None

    def update_u_rou(self, step=0, *args, **kwargs):
        """rho = sum f, p = rho/3, u = (f.e + FOL/2)/rho, u := 0 on the
        masked faces (src/funlbm/flow/d3/base.py:116-133)."""
        self._ensure_explicit()
        nat.call(
            "funlbm_moments", self.lat32, nat.ptr(self._f), self._fol_ptr(self._fc),
            nat.ptr(self._u), nat.ptr(self._rho), nat.ptr(self._p), self._stream(),
        )
        self.update_u_rou_boundary()
        self._pristine = False

    def update_u_rou_boundary(self, *args, **kwargs):
        """The face zeroing is part of the moments kernel; the NON_EQUILIBRIUM
        overrides (inlet duct profile scaled by Re mu / max(rho) / min(Y, Z),
        rho copied from the next plane in; outlet rho and u copied) run here
        (src/funlbm/flow/d3/base.py:135-148)."""
        if self.has_neq:
            nat.call(
                "funlbm_neq_moments", self.lat32, nat.ptr(self._u), nat.ptr(self._rho),
                nat.ptr(self._neq_profile), self._neq_remu, nat.ptr(self._neq_work), self._stream(),
            )

    def cul_equ(self, step=0, *args, **kwargs):
        """BGK collision, keeps feq (src/funlbm/flow/d3/base.py:150-159)."""
        self._ensure_explicit()
        if self._feq is None:
            self._feq = torch.empty_like(self._f)
        nat.call(
            "funlbm_collide", self.lat32, nat.ptr(self._f), nat.ptr(self._u), nat.ptr(self._rho),
            nat.ptr(self._feq), self._stream(),
        )

    def f_stream(self, *args, **kwargs):
        """Streaming with the face overrides (src/funlbm/flow/d3/base.py:161-323)."""
        self._ensure_explicit()
        if self._ftmp is None:
            self._ftmp = torch.empty_like(self._f)
        if self.has_neq and self._feq is None:
            raise RuntimeError("NON_EQUILIBRIUM streaming needs the equilibrium of a previous cul_equ")
        nat.call(
            "funlbm_stream", self.lat32, nat.ptr(self._f), nat.ptr(self._ftmp), nat.ptr(self._rho),
            nat.ptr(self._feq) if self.has_neq else None, self._stream(),
        )
        self._f, self._ftmp = self._ftmp, self._f

    def cul_equ2(self, *args, **kwargs):
        """Forcing term with FOL (src/funlbm/flow/d3/base.py:103-114)."""
        self._ensure_explicit()
        nat.call(
            "funlbm_forcing", self.lat32, nat.ptr(self._f), nat.ptr(self._u),
            self._fol_ptr(self._fc), self._stream(),
        )

    # ------------------------------------------------------------------
    # fused representation
    # ------------------------------------------------------------------
    def _alloc_state(self):
        dt = nat.STORAGE_DTYPES[self.lat.storage]
        for i in range(2):
            if self._state[i] is None:
                plane = nat.state_plane(self.lat)
                self._state[i] = torch.empty((self.q, plane), device=self.device, dtype=dt)
        if self.has_moving_wall:
            for i in range(2):
                if self._rho_face[i] is None:
                    self._rho_face[i] = torch.zeros((self.n,), device=self.device, dtype=_F32)

    def begin_step(self):
        """Opens a fused step.  From the explicit representation this is the
        reference's first call of the step, cul_equ (src/funlbm/lbm/base.py:161),
        followed by the conversion of f* into the state storage."""
        if self._mode == "fused":
            return
        if self.has_neq:
            raise NotImplementedError("NON_EQUILIBRIUM faces run on the explicit path (config fused: false)")
        self._alloc_state()
        s = self._stream()
        if self._pristine and self.lat.storage != nat.STORE_F32:
            # the initial state computed directly in the storage's arithmetic
            # type (for float32 storage it equals the explicit sequence below)
            self.cul_equ()
            nat.call("funlbm_init_state", self.lat, nat.ptr(self._state[self._cur]), s)
        else:
            self.cul_equ()
            nat.call("funlbm_pack", self.lat, nat.ptr(self._f), nat.ptr(self._state[self._cur]), s)
        if self.has_moving_wall:
            self._rho_face[self._cur].copy_(self._rho)
        self._mode = "fused"
        self._mat_valid = True

    def state_source(self):
        """(state, face density) that the coming fused step streams from."""
        return self._state[self._cur], self._rho_face[self._cur]

    def fol_old(self):
        """FOL of the previous step (None when known to be zero)."""
        return self._fol[self._fc] if self._fol_live[self._fc] else None

    def fol_next_buffer(self):
        """The FOL buffer the coming step's spreading writes (FOL_new)."""
        j = 1 - self._fc
        if self._fol[j] is None:
            self._fol[j] = torch.zeros((3, self.n), device=self.device, dtype=_F32)
        return self._fol[j]

    def publish_fol(self, written: bool, tiled: bool = False):
        """Makes FOL_new the current FOL.  Without a spread it is zero, as the
        reference's lagrange_to_flow zeroes FOL every step
        (src/funlbm/lbm/lbm3d.py:104).  `tiled`: the spreading recorded its
        support in the tile table given to finish_step."""
        j = 1 - self._fc
        if written:
            self._fol_live[j] = True
        elif self._fol_live[j]:
            self._fol[j].zero_()
            self._fol_live[j] = False
        self._fol_tiled[j] = bool(written and tiled)
        self._fc = j

    def finish_step(self, fol_new_written: bool, tiles=None, ctl=None):
        """The fused kernel: stream, forcing with FOL_new, moments, collide
        (src/funlbm/lbm/base.py:162-175 and the next step's :161).  tiles/ctl:
        the spreading's tile table, which limits the FOL reads to its support
        (used when both FOL buffers come from consecutive spreadings)."""
        if self._mode != "fused":
            raise RuntimeError("finish_step without begin_step")
        a, b = self._cur, 1 - self._cur
        fol_old = self._fol_ptr(self._fc)
        fol_new = nat.ptr(self._fol[1 - self._fc]) if fol_new_written else None
        use_tiles = (
            tiles is not None and fol_new_written
            and (fol_old is None or self._fol_tiled[self._fc])
        )
        self._tiles = (tiles, ctl) if use_tiles else (None, None)
        self.lat.seed = nat.step_seed(self.seed, self.steps_done)
        if self.stats_next:
            nb = -(-self.n // 256)
            nb += min(nb, 256)  # and the first stage of the final reduction
            if getattr(self, "_fstat_part", None) is None or self._fstat_part.numel() < 9 * nb:
                self._fstat_part = torch.empty(9 * nb, device=self.device, dtype=torch.float64)
            self._ensure_stat_out()
            nat.call(
                "funlbm_fused_step_stats", self.lat, nat.ptr(self._state[a]), nat.ptr(self._state[b]),
                fol_old, fol_new, nat.ptr(self._tiles[0]), nat.ptr(self._tiles[1]),
                nat.ptr(self._rho_face[a]), nat.ptr(self._rho_face[b]), nat.ptr(self._fstat_part), nb,
                nat.ptr(self._stat_out), self._stream(),
            )
        else:
            nat.call(
                "funlbm_fused_step", self.lat, nat.ptr(self._state[a]), nat.ptr(self._state[b]),
                fol_old, fol_new, nat.ptr(self._tiles[0]), nat.ptr(self._tiles[1]),
                nat.ptr(self._rho_face[a]), nat.ptr(self._rho_face[b]), self._stream(),
            )
        self._stats_valid = bool(self.stats_next)
        self._cur = b
        self.steps_done += 1
        self.publish_fol(fol_new_written, tiled=use_tiles)
        self._mat_valid = False

    def _materialize(self):
        """The reference's f, u, rou, p after the last fused step."""
        if self._mat_valid:
            return
        a = 1 - self._cur  # state streamed by the step just taken
        nat.call(
            "funlbm_materialize", self.lat, nat.ptr(self._state[a]),
            self._fol_ptr(1 - self._fc), self._fol_ptr(self._fc), nat.ptr(self._tiles[0]),
            nat.ptr(self._tiles[1]), nat.ptr(self._rho_face[a]), nat.ptr(self._f), nat.ptr(self._u),
            nat.ptr(self._rho), nat.ptr(self._p), self._stream(),
        )
        self._mat_valid = True

    # ------------------------------------------------------------------
    # the reference's tensors (views [X, Y, Z, *] of the SoA buffers)
    # ------------------------------------------------------------------
    def _view(self, t, k):
        X, Y, Z = self.nodes
        return t.view(k, X, Y, Z).permute(1, 2, 3, 0)

    @property
    def f(self):
        if self._mode == "fused":
            self._materialize()
        return self._view(self._f, self.q)

    @f.setter
    def f(self, value):
        self._ensure_explicit()
        v = torch.as_tensor(value, device=self.device, dtype=_F32).reshape(*self.nodes, self.q)
        self._f.copy_(v.permute(3, 0, 1, 2).reshape(self.q, self.n))
        self._pristine = False
        self._stats_valid = False

    @property
    def u(self):
        if self._mode == "fused":
            self._materialize()
        return self._view(self._u, 3)

    @property
    def rou(self):
        if self._mode == "fused":
            self._materialize()
        return self._view(self._rho, 1)

    @property
    def p(self):
        if self._mode == "fused":
            self._materialize()
        return self._view(self._p, 1)

    @property
    def FOL(self):
        # the caller may write through the view
        self._fol_live[self._fc] = True
        self._fol_tiled[self._fc] = False
        self._pristine = False
        self._stats_valid = False
        return self._view(self._fol[self._fc], 3)

    def fol_array(self):
        """A copy of the current FOL, [X, Y, Z, 3], without marking it written."""
        if not self._fol_live[self._fc] or self._fol[self._fc] is None:
            return torch.zeros((*self.nodes, 3), device=self.device, dtype=_F32)
        return self._view(self._fol[self._fc], 3).clone()

    @FOL.setter
    def FOL(self, value):
        v = torch.as_tensor(value, device=self.device, dtype=_F32).reshape(*self.nodes, 3)
        self._fol[self._fc].copy_(v.permute(3, 0, 1, 2).reshape(3, self.n))
        self._fol_live[self._fc] = True
        self._fol_tiled[self._fc] = False
        self._pristine = False
        self._stats_valid = False

    @property
    def feq(self):
        """Equilibrium of the last explicit collision.  After fused steps the
        equilibrium of the current moments (the one the next collision uses)."""
        if self._mode == "fused":
            self._materialize()
            out = torch.empty_like(self._f)
            nat.call("funlbm_equilibrium_state", self.lat32, nat.ptr(self._rho), nat.ptr(self._u),
                     nat.ptr(out), self._stream())
            return self._view(out, self.q)
        if self._feq is None:
            return None
        return self._view(self._feq, self.q)

    @property
    def tau(self):
        X, Y, Z = self.nodes
        return self._tau.expand(X, Y, Z, 1)

    @property
    def x(self):
        if self._x is None:
            X, Y, Z = self.nodes
            gx, gy, gz = torch.meshgrid(
                torch.arange(X, device=self.device),
                torch.arange(Y, device=self.device),
                torch.arange(Z, device=self.device),
                indexing="ij",
            )
            self._x = torch.stack([gx, gy, gz], dim=-1)
        return self._x

    def stats(self):
        """min, mean, max of the reference's f, u and rho as a double[9]
        device tensor.  After fused steps it is one read of the state
        (funlbm_flow_stats) instead of materialising the fields."""
        if self._mode == "fused" and not self._mat_valid:
            if self._stats_valid:  # produced by the step itself
                return self._stat_out
            if getattr(self, "_stat_part", None) is None:
                self._stat_part = torch.empty(9 * 2048, device=self.device, dtype=torch.float64)
            self._ensure_stat_out()
            a = 1 - self._cur
            nat.call(
                "funlbm_flow_stats", self.lat, nat.ptr(self._state[a]), self._fol_ptr(1 - self._fc),
                self._fol_ptr(self._fc), nat.ptr(self._tiles[0]), nat.ptr(self._tiles[1]),
                nat.ptr(self._rho_face[a]), nat.ptr(self._stat_part), 2048, nat.ptr(self._stat_out),
                self._stream(),
            )
            return self._stat_out
        vals = []
        for t in (self.f, self.u, self.rou):
            vals += [t.min().double(), t.double().mean(), t.max().double()]
        return torch.stack(vals)

    def conservation(self):
        """Global mass and momentum as a double[4] device tensor: sum of rho,
        then sum of rho u_d (the reference's rho and u, reduced in double).
        After fused steps it is one read of the state
        (funlbm_flow_conservation)."""
        if self._mode == "fused" and not self._mat_valid:
            if getattr(self, "_cons_part", None) is None:
                self._cons_part = torch.empty(4 * 2048, device=self.device, dtype=torch.float64)
                self._cons_out = torch.empty(4, device=self.device, dtype=torch.float64)
            a = 1 - self._cur
            nat.call(
                "funlbm_flow_conservation", self.lat, nat.ptr(self._state[a]), self._fol_ptr(1 - self._fc),
                self._fol_ptr(self._fc), nat.ptr(self._tiles[0]), nat.ptr(self._tiles[1]),
                nat.ptr(self._rho_face[a]), nat.ptr(self._cons_part), 2048, nat.ptr(self._cons_out),
                self._stream(),
            )
            return self._cons_out
        rho = self.rou.double().reshape(-1)
        u = self.u.double().reshape(-1, 3)
        return torch.stack([rho.sum()] + [(rho * u[:, d]).sum() for d in range(3)])

    def _ensure_stat_out(self):
        if getattr(self, "_stat_out", None) is None:
            self._stat_out = torch.empty(9, device=self.device, dtype=torch.float64)

    def to_json(self):
        """min/mean/max of f, u, rho (src/funlbm/flow/base.py:183-206)."""
        v = self.stats().tolist()
        return {"f": v[0:3], "u": v[3:6], "rho": v[6:9]}


def _lattice_class(name):
    class _C(FlowD3):
        lattice_name = name

    _C.__name__ = f"Flow{name}"
    _C.__qualname__ = f"Flow{name}"
    return _C

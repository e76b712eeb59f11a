"""Solver configuration and the run loop (src/funlbm/lbm/base.py:17-227)."""

import logging
from typing import List

from ..base.work import Worker
from ..config.base import BaseConfig, FileConfig, deep_get
from ..flow import FlowConfig, FlowD3, create_flow

logger = logging.getLogger("funlbm_amd")


class Config(BaseConfig):
    """src/funlbm/lbm/base.py:17-40.

    Extension keys of this package (absent from the reference's configs):
      fused      run_step uses the fused HIP step (default true); false runs
                 the reference's method sequence, one kernel per method
      log_every  cadence of the per-step statistics of run() (default 1, the
                 reference's; 0 disables them)
      db_path    SQLite file for the per-step records of run() (the
                 reference's flow / particle tables; default none)
      checkpoint_every  write a checkpoint every this many steps into
                 file.cache_dir/checkpoint (default 0: none; the reference
                 writes every file.per_steps = 100 steps, 1.3 GB each for a
                 256^3 D3Q19 lattice)
      output_every  write the reference's save_bak file set (Tecplot and VTK
                 files of u, rho, p and the markers, under file.cache_dir)
                 every this many steps (default 0: never, as the reference,
                 whose save_bak is not called)
      graph      replay the fused particle step from HIP graphs once it is
                 periodic (default false: no faster on ROCm 7.2; see LBMD3)
      async_log  run() copies its per-step records back without waiting for
                 the GPU and emits them up to 8 steps late, in order
                 (default true; false waits for each record, as the
                 reference); the weight-sum failure stops the run when its
                 record is emitted
    The reference's `checkpoint` key (a file to resume from) is honoured.
    """

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.dt: float = 1.0
        self.dx: float = 1.0
        self.checkpoint = None
        self.max_step = 10000
        self.device: str = "auto"
        self.file_config = FileConfig()
        self.flow_config = FlowConfig()
        self.particles: List = []
        self.fused: bool = True
        self.log_every: int = 1
        self.checkpoint_every: int = 0
        self.db_path = None
        self.async_log: bool = True
        self.output_every: int = 0
        self.graph: bool = False

    def _from_json(self, config_json: dict, *args, **kwargs) -> "Config":
        from ..particle import ParticleConfig

        self.dt = deep_get(config_json, "dt") or self.dt
        self.dx = deep_get(config_json, "dx") or self.dx
        self.device = deep_get(config_json, "device") or self.device
        self.max_step = deep_get(config_json, "max_step") or self.max_step
        self.checkpoint = deep_get(config_json, "checkpoint") or self.checkpoint
        self.file_config = FileConfig().from_json(deep_get(config_json, "file") or {})
        self.flow_config = FlowConfig().from_json(deep_get(config_json, "flow") or {})
        for config in deep_get(config_json, "particles") or []:
            self.particles.append(ParticleConfig().from_json(config_json=config))
        fused = deep_get(config_json, "fused")
        self.fused = True if fused is None else bool(fused)
        le = deep_get(config_json, "log_every")
        self.log_every = 1 if le is None else int(le)
        self.checkpoint_every = int(deep_get(config_json, "checkpoint_every") or 0)
        self.db_path = deep_get(config_json, "db_path") or self.db_path
        al = deep_get(config_json, "async_log")
        self.async_log = True if al is None else bool(al)
        self.output_every = int(deep_get(config_json, "output_every") or 0)
        self.graph = bool(deep_get(config_json, "graph") or False)
        return self


def create_lbm_config(path="./config.json") -> Config:
    return Config().from_file(path)


class LBMBase(Worker):
    """Solver object and step orchestration (src/funlbm/lbm/base.py:43-227)."""

    def __init__(self, config: Config = None, flow: FlowD3 = None, particle_swarm=None,
                 device=None, *args, **kwargs):
        from ..particle import create_particle_swarm

        self.config: Config = config or create_lbm_config()
        dev = device
        if dev is None and self.config.device not in (None, "auto", "cpu"):
            dev = self.config.device
        super().__init__(device=dev)
        self.step = 1
        self.flow = flow or create_flow(flow_config=self.config.flow_config, device=self.device)
        fc = self.config.flow_config
        self.particle_swarm = particle_swarm or create_particle_swarm(
            self.config.particles, device=self.device, gravity=fc.gravity, gl=fc.gl
        )
        self.run_status = True
        self.history = []
        self.store = None
        self._steplog = None
        if self.config.db_path:
            from .store import StepStore

            self.store = StepStore(self.config.db_path)

    def run(self, max_steps: int = 1000000, *args, **kwargs) -> None:
        """src/funlbm/lbm/base.py:86-103.  A logged step asks the flow for
        its statistics within the step (no second pass over the state)."""
        self.init()
        total_steps = min(max_steps, self.config.max_step)
        every = self.config.log_every
        if self.config.async_log and every and getattr(self.flow, "device", None) is not None \
                and str(self.flow.device).startswith("cuda"):
            from .steplog import StepLog

            self._steplog = StepLog(self)
        try:
            for _ in range(total_steps):
                if hasattr(self.flow, "stats_next"):
                    self.flow.stats_next = bool(every) and self.step % every == 0
                self.run_step(step=self.step)
                self._log_step_info(self.step)
                if self.run_status is False:
                    break
                if self.step >= total_steps:
                    break
                self.step += 1
        finally:
            if hasattr(self.flow, "stats_next"):
                self.flow.stats_next = False
            if self._steplog is not None:
                self._steplog.drain(wait=True)
                self._steplog = None

    def _log_step_info(self, *args, **kwargs) -> None:
        """Per-step statistics (src/funlbm/lbm/base.py:105-144) at the
        configured cadence."""
        every = self.config.log_every
        if not every or self.step % every:
            return
        if self._steplog is not None:
            self._steplog.record(self.step)
            return
        self._check_run_status()
        flow = self.flow.to_json()
        particles = [p.to_json() for p in self.particle_swarm.particles]
        self._emit_record(self.step, flow, particles)

    def _emit_record(self, step: int, flow: dict, particles: list) -> None:
        """History entry, SQLite rows and the reference's log line of one step."""
        from ..particle.base.base import format_particle

        self.history.append({"step": step, "flow": flow, "particles": particles})
        if self.store is not None:
            self.store.put(step, flow, particles)
        if not logger.isEnabledFor(logging.INFO):
            return
        info = [f"step={step:6d}"] + [
            f"{k}=" + ",".join(f"{v:.6f}" for v in flow[k]) for k in ("f", "u", "rho")
        ]
        info += [format_particle(d) for d in particles]
        logger.info("\t".join(info))

    def _check_run_status(self):
        pass

    def run_step(self, *args, **kwargs) -> None:
        """src/funlbm/lbm/base.py:146-157."""
        # NON_EQUILIBRIUM faces are implemented on the explicit path only
        if self.config.fused and not getattr(self.flow, "has_neq", False):
            self._run_step_fused()
        else:
            self._compute_flow()
            self._handle_immersed_boundary()
            self._update_particles()
        self.save(self.step)
        every = self.config.output_every
        if every and self.step % every == 0:
            self.save_bak(self.step, force=True)

    def _compute_flow(self) -> None:
        """src/funlbm/lbm/base.py:159-163."""
        self.flow.cul_equ(step=self.step)
        self.flow.f_stream()
        self.flow.update_u_rou(step=self.step)

    def _handle_immersed_boundary(self) -> None:
        """src/funlbm/lbm/base.py:165-175."""
        self.flow_to_lagrange()
        self.particle_to_wall()
        for particle in self.particle_swarm.particles:
            particle.update_from_lar(dt=self.config.dt, gl=self.config.flow_config.gl)
        self.lagrange_to_flow()
        self.flow.cul_equ2()
        self.flow.update_u_rou()

    def _update_particles(self) -> None:
        """src/funlbm/lbm/base.py:177-179."""
        self.particle_swarm.update(dt=self.config.dt)

    def _run_step_fused(self) -> None:
        raise NotImplementedError()

    def init(self, *args, **kwargs) -> None:
        """src/funlbm/lbm/base.py:181-184: set up, then resume from
        config.checkpoint if given."""
        self._init()
        if self.config.checkpoint is not None:
            self.load_checkpoint(self.config.checkpoint)

    def _init(self, *args, **kwargs):
        raise NotImplementedError()

    def flow_to_lagrange(self, *args, **kwargs):
        raise NotImplementedError()

    def lagrange_to_flow(self, *args, **kwargs):
        raise NotImplementedError()

    def particle_to_wall(self, *args, **kwargs):
        raise NotImplementedError()

    def save(self, step=10, *args, **kwargs):
        """Checkpoint every config.checkpoint_every steps
        (src/funlbm/lbm/base.py:198-203); the file is
        checkpoint_{step:010d}.npz in file.cache_dir/checkpoint."""
        every = self.config.checkpoint_every
        if not every or step % every:
            return None
        path = f"{self.config.file_config.checkpoint_path}/checkpoint_{str(step).zfill(10)}.npz"
        return self.dump_checkpoint(path)

    def save_bak(self, step=10, force=False, *args, **kwargs):
        """Tecplot and VTK files of the flow and the markers at a step
        divisible by file.per_steps (src/funlbm/lbm/lbm3d.py:169-244), or
        at any step with force; returns the paths written (funlbm_amd.lbm.output)."""
        if not force and step % self.config.file_config.per_steps > 0:
            return []
        from .output import save_fields

        return save_fields(self, step)

    def dump_checkpoint(self, path: str) -> str:
        """src/funlbm/lbm/base.py:205-210."""
        from .checkpoint import dump

        return dump(self, path)

    def load_checkpoint(self, checkpoint_path: str = None, *args, **kwargs) -> None:
        """src/funlbm/lbm/base.py:212-227."""
        import os

        from .checkpoint import load

        if checkpoint_path is None or not os.path.exists(checkpoint_path):
            logger.error("load failed: checkpoint %r not found", checkpoint_path)
            return
        load(self, checkpoint_path)
        if hasattr(self, "reset_graphs"):
            self.reset_graphs()

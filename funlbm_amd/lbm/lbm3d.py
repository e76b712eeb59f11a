"""3-D solver with immersed-boundary particles (src/funlbm/lbm/lbm3d.py:11-264).

run_step (fused, the default) issues per step:
  [particles]  ibm_interp_boxed (pull + moments with FOL_old once per node of
               each particle's box, then the marker stencils read them),
               wall_spring, particle_forces (double), ibm_spread into FOL_new
  fused_step   stream, forcing with FOL_new, moments, collide -- one pass
  [particles]  particle_pose (double rotation, float markers)
The explicit path (config "fused": false) runs the reference's method
sequence with one kernel per method; both give the same results.

With particles, the fused step can be replayed from a HIP graph once it is
periodic (config "graph", default false): after two steps both FOL buffers
and both stencil-origin buffers are live, and a step's launches then depend
only on the parity of the two ping-pong pairs and on whether the weight-sum
check is armed (step > 100).  The first step of each such kind is captured
(its launches go into the graph, its host bookkeeping runs as usual) and
replayed at once; later steps of that kind run the host bookkeeping with
launches switched off and replay the graph.  Steps with in-pass statistics,
and 16-bit states with stochastic rounding (whose dither seed changes every
step), run eagerly.  On the MI355X with ROCm 7.2 a graph launch costs about
10 us of host time per kernel node, as much as launching the kernel
directly, so a replayed step is no faster than an eager one (the free-fall
case, 64 x 64 x 128 nodes and 452 markers: 0.069 ms per step eager, 0.0675
ms per step for the replay alone of graphs of 1 to 32 steps); hence the
default.
"""

import torch

from .. import _native as nat
from ..particle import boxes
from .base import LBMBase


# markers from which the boxed interpolation beats the per-node one: below,
# the per-node kernel takes a few microseconds and the box pass and the
# fallback pass are two more launches
BOXED_MIN_MARKERS = 4096


class LBMD3(LBMBase):
    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        # interpolation from the fused state through the particles' moment
        # boxes, and spreading claims per particle box (same results as the
        # per-node pull and per-marker claims, which False selects); None:
        # boxes from BOXED_MIN_MARKERS markers up, where they pay for their
        # extra launches
        self.boxed_interp = None
        self._graphs = {}

    def _init(self):
        """flow.init, one extra collision, particles (src/funlbm/lbm/lbm3d.py:23-30)."""
        self._graphs = {}
        self.flow.init()
        self.flow.cul_equ()
        self.particle_swarm.init()
        sw = self.particle_swarm
        self._has_particles = sw.n_markers > 0
        if self._has_particles:
            dev = self.device
            self._owner = torch.zeros(self.flow.n, device=dev, dtype=torch.int64)
            self._origin = [
                torch.zeros((sw.n_markers, 3), device=dev, dtype=torch.int32) for _ in range(2)
            ]
            self._origin_ok = [False, False]
            self._tiles = torch.zeros((self.flow.n + 255) // 256, device=dev, dtype=torch.int32)
            # moment boxes of the interpolation from the fused state
            radii = [boxes.box_radius(p._lagrange) for p in sw.particles]
            rows, chunks, nodes = boxes.layout(radii)
            self._box_rows = torch.from_numpy(rows.reshape(-1)).to(dev)
            self._box_chunks = torch.from_numpy(chunks).to(dev) if len(chunks) else None
            self._box_cache = torch.empty(4 * max(nodes, 1), device=dev, dtype=torch.float32)
            self._box_nodes = nodes
            self._box_miss = torch.zeros(sw.n_markers + 1, device=dev, dtype=torch.int32)
            self._claims_per_marker = 0 if all(boxes.claims_in_box(R) for R in radii) else 1

    # ------------------------------------------------------------------
    # IBM phases
    # ------------------------------------------------------------------
    def _boxed(self) -> bool:
        if self.boxed_interp is None:
            return self.particle_swarm.n_markers >= BOXED_MIN_MARKERS
        return bool(self.boxed_interp)

    def _interp(self, from_state: bool):
        sw, fl = self.particle_swarm, self.flow
        if from_state and self._boxed():
            state, rface = fl.state_source()
            nch = 0 if self._box_chunks is None else int(self._box_chunks.numel())
            nat.call(
                "funlbm_ibm_interp_boxed", fl.lat, nat.ptr(state), nat.ptr(fl.fol_old()), nat.ptr(rface),
                sw.n_markers, nat.ptr(sw.lx), nat.ptr(sw.pid), len(sw.particles), nat.ptr(sw.pstate),
                nat.ptr(self._box_rows), nch, nat.ptr(self._box_chunks) if nch else None,
                nat.ptr(self._box_cache), self._box_nodes, nat.ptr(self._box_miss), float(self.config.dt),
                1 if self.step > 100 else 0, nat.ptr(sw.lu), nat.ptr(sw.lF), nat.ptr(sw.ctl),
                fl._stream(),
            )
            return
        if from_state:
            state, rface = fl.state_source()
            args = (nat.ptr(state), nat.ptr(fl.fol_old()), nat.ptr(rface), None, None)
        else:
            args = (None, None, None, nat.ptr(fl._u), nat.ptr(fl._rho))
        nat.call(
            "funlbm_ibm_interp", fl.lat if from_state else fl.lat32, *args, sw.n_markers,
            nat.ptr(sw.lx), nat.ptr(sw.pid), nat.ptr(sw.pstate), float(self.config.dt),
            1 if self.step > 100 else 0, nat.ptr(sw.lu), nat.ptr(sw.lF), nat.ptr(sw.ctl),
            fl._stream(),
        )

    def flow_to_lagrange(self, *args, **kwargs):
        """Interpolation and direct forcing (src/funlbm/lbm/lbm3d.py:66-100)."""
        if self._has_particles:
            self._interp(from_state=False)

    def particle_to_wall(self, *args, **kwargs):
        """Wall spring (src/funlbm/lbm/lbm3d.py:121-167)."""
        sw, fl = self.particle_swarm, self.flow
        if not self._has_particles:
            return
        nat.call(
            "funlbm_wall_spring", fl.lat32, len(sw.particles), nat.ptr(sw.offsets), nat.ptr(sw.lx),
            nat.ptr(sw.lF), float(self.config.dx), fl._stream(),
        )

    def _spread(self):
        sw, fl = self.particle_swarm, self.flow
        j = 1 - fl._fc
        buf = fl.fol_next_buffer()
        if fl._fol_live[j] and not self._origin_ok[j]:
            buf.zero_()
        clear = nat.ptr(self._origin[j]) if (fl._fol_live[j] and self._origin_ok[j]) else None
        args = (fl.lat32, sw.n_markers, nat.ptr(sw.lx), nat.ptr(sw.lF), nat.ptr(sw.lm), nat.ptr(buf), clear,
                nat.ptr(self._origin[j]), nat.ptr(self._owner), nat.ptr(self._tiles), nat.ptr(sw.ctl))
        if self._boxed():
            # claims per particle box in shared memory (same ownership)
            nat.call("funlbm_ibm_spread_boxed", *args, len(sw.particles), nat.ptr(sw.offsets),
                     nat.ptr(sw.pid), nat.ptr(sw.pstate), nat.ptr(self._box_rows),
                     self._claims_per_marker, fl._stream())
        else:
            nat.call("funlbm_ibm_spread", *args, fl._stream())
        self._origin_ok[j] = True

    def lagrange_to_flow(self, *args, **kwargs):
        """Spreading into a zeroed FOL (src/funlbm/lbm/lbm3d.py:102-119)."""
        if self._has_particles:
            self._spread()
        self.flow.publish_fol(self._has_particles, tiled=self._has_particles)

    def _graph_key(self):
        """The kind of a fused particle step whose launches a graph can
        replay, or None while the step must run eagerly."""
        fl = self.flow
        if not (self.config.graph and self._has_particles and fl._mode == "fused"):
            return None
        if fl.stats_next or fl.lat.rounding == nat.ROUND_STOCHASTIC:
            return None
        if not (all(fl._fol_live) and all(fl._fol_tiled) and all(self._origin_ok)):
            return None
        return (fl._cur, fl._fc, self.step > 100, self._boxed())

    def reset_graphs(self):
        """Drops the captured steps (after a change of the lattice or the
        particles' buffers)."""
        self._graphs = {}

    def _run_step_fused(self):
        key = self._graph_key()
        if key is None:
            self._step_launches()
            return
        g = self._graphs.get(key)
        if g is None:
            dev = self.device
            g = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream(device=dev)
            side.wait_stream(torch.cuda.current_stream(dev))
            with torch.cuda.stream(side):
                with torch.cuda.graph(g, stream=side):
                    self._step_launches()
            torch.cuda.current_stream(dev).wait_stream(side)
            self._graphs[key] = g
        else:
            with nat.launches_off():
                self._step_launches()
        g.replay()

    def _step_launches(self):
        fl, sw = self.flow, self.particle_swarm
        fl.begin_step()
        if self._has_particles:
            self._interp(from_state=True)
            self.particle_to_wall()
            sw.forces(self.config.dt, 1.0)
            self._spread()
        if self._has_particles:
            fl.finish_step(True, tiles=self._tiles, ctl=sw.ctl)
            sw.pose()
        else:
            fl.finish_step(False)

    def _handle_immersed_boundary(self) -> None:
        """Explicit path: the reference's sequence (src/funlbm/lbm/base.py:165-175)."""
        self.flow_to_lagrange()
        self.particle_to_wall()
        if self._has_particles:
            self.particle_swarm.forces(self.config.dt, 1.0)
        self.lagrange_to_flow()
        self.flow.cul_equ2()
        self.flow.update_u_rou()

    def _check_run_status(self):
        """The weight-sum failure check of flow_to_lagrange
        (src/funlbm/lbm/lbm3d.py:84-86), read back at the logging cadence."""
        if self._has_particles and int(self.particle_swarm.ctl[1].item()) != 0:
            self.run_status = False


class LBMD3Q27(LBMD3):
    pass


class LBMD3Q19(LBMD3):
    pass


class LBMD3Q15(LBMD3):
    pass


class LBMD3Q13(LBMD3):
    pass

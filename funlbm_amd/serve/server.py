"""Continuous-batching server for settling-case requests.

Requests arrive over time (each a CaseSpec: lattice, particles, step count).
The server keeps their state resident in the BatchEngine's pools and admits a
waiting request as soon as its HBM footprint fits under the budget and the
pools have room; admitted cases step together, each to its own step count, in
ticks of `tick_steps` steps.  By default a tick is `tick_steps` eager step
launches whose grids the engine sizes from the live counts (one block per
active tile of each lattice: the population changes between ticks, and a
persistent grid of fixed size ran the fused kernels 15 % slower).  With
`graph=True` a tick is one replay of a captured HIP graph whose kernels read
the live work lists from device memory, so admissions between ticks need no
re-capture, but the grids stay those of the capture.

The loop is pipelined: up to `depth` ticks are in flight on the stream.
Behind each tick the server enqueues asynchronous copies of the per-case
counters, the particle states and the trajectory ring into pinned host
buffers of its own, and an event.  While the GPU runs the next tick, the host
waits for the oldest tick's event only, streams out the trajectory rows of
the steps taken in that tick, retires the requests that finished in it and
admits waiting ones (admission enqueues its copies asynchronously too).  The
host knows how many steps each admitted case takes per tick, so it launches
a tick only while some case still has steps left.  Admission is first come,
first served, under the HBM budget and, optionally, a cap on the lattice
nodes stepping at once.

No reference counterpart: the reference runs one case per process
(src/funlbm/server/script.py:10-24).
"""

import dataclasses
import time
from collections import deque
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from .. import _native as nat
from ..batch.engine import BatchEngine, CaseSpec


@dataclasses.dataclass
class Request:
    rid: int
    spec: CaseSpec
    arrival: float  # seconds since the start of the run
    admitted: Optional[float] = None
    finished: Optional[float] = None
    cid: Optional[int] = None
    # per tick, float32 [particles, steps taken in the tick, 4] rows of
    # (cx, cy, cz, step), in step order
    trajectory: List[np.ndarray] = dataclasses.field(default_factory=list)
    result: Optional[np.ndarray] = None  # final particle rows (double)
    first_tick: int = 0  # the first tick that steps the request
    left: int = 0  # steps not yet launched
    streamed: int = 0  # last step whose trajectory row was streamed out

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class _Snapshot:
    """Pinned host copies of what one tick leaves on the device."""

    def __init__(self, eng: BatchEngine, with_traj: bool):
        pin = dict(pin_memory=True)
        self.left = torch.empty(eng.case_left.shape, dtype=eng.case_left.dtype, **pin)
        self.pstate = torch.empty(eng.pstate.shape, dtype=eng.pstate.dtype, **pin)
        self.traj = torch.empty(eng.traj.shape, dtype=eng.traj.dtype, **pin) if with_traj else None
        self.event = torch.cuda.Event()
        self.tick = -1

    def record(self, eng: BatchEngine, tick: int):
        self.left.copy_(eng.case_left, non_blocking=True)
        self.pstate.copy_(eng.pstate, non_blocking=True)
        if self.traj is not None:
            self.traj.copy_(eng.traj, non_blocking=True)
        self.event.record(torch.cuda.current_stream(eng.device))
        self.tick = tick


class Server:
    def __init__(self, engine: BatchEngine, hbm_budget: float, tick_steps: int = 50,
                 stream_trajectories: bool = True, depth: int = 2, max_active_nodes: Optional[int] = None,
                 graph: bool = False):
        if engine.traj_cap < tick_steps:
            raise ValueError("the trajectory ring must hold one tick of steps")
        if depth < 1:
            raise ValueError("at least one tick must be in flight")
        self.engine = engine
        self.graph = bool(graph)
        engine.exact_grids = not self.graph
        self.budget = float(hbm_budget)
        self.tick_steps = int(tick_steps)
        self.stream = stream_trajectories
        self.depth = int(depth)
        # lattice nodes admitted at once (None: the HBM budget alone limits
        # admission): enough to keep the GPU busy, few enough that requests
        # finish in arrival order instead of all sharing the GPU
        self.max_active_nodes = max_active_nodes
        self.active_nodes = 0
        self.in_use = 0.0
        self.queue: deque = deque()
        self.running: Dict[int, Request] = {}
        self.done: List[Request] = []
        self._graph_ready = False
        self._storage_bytes = torch.tensor([], dtype=engine.pool[0].dtype).element_size()
        self._free = [_Snapshot(engine, stream_trajectories) for _ in range(self.depth)]
        self._inflight: deque = deque()
        self._ticks = 0  # ticks launched

    def footprint(self, spec: CaseSpec) -> float:
        return BatchEngine.case_bytes_needed(spec, self._storage_bytes)

    def submit(self, req: Request):
        self.queue.append(req)

    def _admit(self, now: float):
        while self.queue:
            req = self.queue[0]
            need = self.footprint(req.spec)
            if self.in_use + need > self.budget or not self.engine.fits(req.spec):
                break
            if (self.max_active_nodes is not None and self.running
                    and self.active_nodes + req.spec.n > self.max_active_nodes):
                break
            self.queue.popleft()
            req.cid = self.engine.add(req.spec)
            req.admitted = now
            req.first_tick = self._ticks
            req.left = int(req.spec.steps)
            self.in_use += need
            self.active_nodes += req.spec.n
            self.running[req.cid] = req

    def _launch(self):
        """Enqueues one tick and the snapshot copies behind it."""
        eng = self.engine
        if not self.graph:
            eng.step(self.tick_steps)
        else:
            if not self._graph_ready:
                eng.capture(self.tick_steps)
                self._graph_ready = True
            eng.run_graph()
        snap = self._free.pop()
        snap.record(eng, self._ticks)
        self._inflight.append(snap)
        for req in self.running.values():
            req.left -= min(req.left, self.tick_steps)
        self._ticks += 1

    def _complete(self, clock: Callable[[], float], t0: float):
        """Waits for the oldest tick in flight and handles its snapshot."""
        snap = self._inflight.popleft()
        snap.event.synchronize()
        now = clock() - t0
        eng = self.engine
        left = snap.left.numpy()
        finished = []
        for cid, req in self.running.items():
            if snap.tick < req.first_tick:
                continue  # admitted after this tick was launched
            s = eng.slots[cid]
            if snap.traj is not None and s.p_count:
                rows = snap.traj.view(-1, eng.traj_cap, 4)[s.p_off : s.p_off + s.p_count].numpy()
                steps = rows[0, :, 3]
                new = np.nonzero(steps > req.streamed)[0]
                if len(new):
                    order = new[np.argsort(steps[new])]
                    req.trajectory.append(rows[:, order].copy())
                    req.streamed = int(steps[order[-1]])
            if left[cid] <= 0:
                finished.append(cid)
        ps = snap.pstate.numpy().reshape(-1, nat.PSTRIDE)
        for cid in finished:
            req = self.running.pop(cid)
            s = eng.slots[cid]
            req.result = ps[s.p_off : s.p_off + s.p_count].copy()
            req.finished = now
            self.in_use -= self.footprint(req.spec)
            self.active_nodes -= req.spec.n
            eng.remove(cid)
            self.done.append(req)
        self._free.append(snap)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            max_seconds: float = 1e9) -> Dict:
        """Feeds `requests` at their arrival times and serves until all are
        done (or max_seconds pass).  Returns throughput and latency figures:
        solves/s over the wall time, p50/p99 latency of the completed
        requests, and the requests still queued or running at the end."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = clock()
        busy = 0.0
        while pending or self.queue or self.running or self._inflight:
            now = clock() - t0
            if now > max_seconds:
                break
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            self._admit(now)
            steps_left = any(r.left > 0 for r in self.running.values())
            if steps_left and len(self._inflight) < self.depth:
                self._launch()
                continue
            if self._inflight:
                ts = clock()
                self._complete(clock, t0)
                busy += clock() - ts
                continue
            if pending:
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                continue
            if self.queue:
                raise MemoryError("a request does not fit the pools even when the server is idle")
            break
        # drain what is still in flight (a run cut at max_seconds)
        while self._inflight:
            self._complete(clock, t0)
        wall = clock() - t0
        lat = np.array([r.latency for r in self.done if r.latency is not None])
        lups = sum(r.spec.n * r.spec.steps for r in self.done)
        return {
            "completed": len(self.done),
            "submitted": len(requests),
            "backlog": len(self.queue) + len(self.running) + len(pending),
            "wall_s": wall,
            "wait_s": busy,
            "ticks": self._ticks,
            "solves_per_s": len(self.done) / wall if wall > 0 else 0.0,
            "mlups": lups / wall / 1e6 if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        }

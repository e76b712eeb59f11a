"""Continuous-batching server for settling-case requests.

Requests arrive over time (each a CaseSpec: lattice, particles, step count).
The server keeps their state resident in the pools of one or more
BatchEngines and admits a waiting request as soon as its HBM footprint fits
under the budget and an engine's pools have room; admitted cases step
together, each to its own step count, in ticks of `tick_steps` steps.  By
default a tick is `tick_steps` eager step launches whose grids the engine
sizes from the live counts (one block per active tile of each lattice: the
population changes between ticks, and a persistent grid of fixed size ran the
fused kernels 15 % slower).  With `graph=True` a tick is one replay of a
captured HIP graph whose kernels read the live work lists from device memory,
so admissions between ticks need no re-capture, but the grids stay those of
the capture.

With several engines each runs on a stream of its own and a request goes to
the engine with the fewest active nodes.  An engine's step is a chain of
small, latency-bound IBM kernels before its bandwidth-bound fused kernels;
on several streams one engine's IBM kernels run beside another's fused
kernels.  Measured on the MI355X with the serving workload at 200 req/s
(2^26 active nodes): one engine 42.1 GLUPS stepped, two 44.2, three 45.4,
four 42.7.  Every operation on an engine (admission, removal, tick, snapshot
copies) is issued on its stream, so its work lists are never rewritten under
a tick in flight.

The loop is pipelined: up to `depth` ticks are in flight.  Behind each
engine's tick the server enqueues asynchronous copies of the per-case
counters, the particle states and the trajectory ring into pinned host
buffers of its own, and an event.  While the GPU runs the next tick, the host
waits for the oldest tick's events only, streams out the trajectory rows of
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
from contextlib import nullcontext
from typing import Callable, Dict, List, Optional, Sequence, Tuple, Union

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
    engine: Optional[int] = None  # index of the engine that holds the case
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


class _Lane:
    """One engine, its stream and its snapshots."""

    def __init__(self, eng: BatchEngine, stream, depth: int, with_traj: bool):
        self.eng = eng
        self.stream = stream
        self.free = [_Snapshot(eng, with_traj) for _ in range(depth)]
        self.graph_ready = False
        self.active_nodes = 0

    def on_stream(self):
        return torch.cuda.stream(self.stream) if self.stream is not None else nullcontext()


class Server:
    def __init__(self, engine: Union[BatchEngine, Sequence[BatchEngine]], hbm_budget: float, tick_steps: int = 50,
                 stream_trajectories: bool = True, depth: int = 2, max_active_nodes: Optional[int] = None,
                 graph: bool = False):
        engines = [engine] if isinstance(engine, BatchEngine) else list(engine)
        if not engines:
            raise ValueError("the server needs at least one engine")
        if any(e.traj_cap < tick_steps for e in engines):
            raise ValueError("the trajectory ring must hold one tick of steps")
        if depth < 1:
            raise ValueError("at least one tick must be in flight")
        if len({e.pool[0].dtype for e in engines}) != 1:
            raise ValueError("the engines must share one storage format")
        self.engines = engines
        self.engine = engines[0]
        self.graph = bool(graph)
        for e in engines:
            e.exact_grids = not self.graph
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
        self.running: Dict[Tuple[int, int], Request] = {}  # (engine, cid) -> request
        self.done: List[Request] = []
        self._storage_bytes = torch.tensor([], dtype=engines[0].pool[0].dtype).element_size()
        # one engine steps on the caller's current stream; several on streams of their own
        streams = [None] if len(engines) == 1 else [torch.cuda.Stream(device=e.device) for e in engines]
        for st, e in zip(streams, engines):
            if st is not None:
                st.wait_stream(torch.cuda.current_stream(e.device))  # the engines' set-up
        self._lanes = [_Lane(e, st, self.depth, stream_trajectories) for e, st in zip(engines, streams)]
        self._inflight: deque = deque()  # per tick: [(lane index, snapshot)]
        self._ticks = 0  # ticks launched

    def footprint(self, spec: CaseSpec) -> float:
        return BatchEngine.case_bytes_needed(spec, self._storage_bytes)

    def submit(self, req: Request):
        self.queue.append(req)

    def _admit(self, now: float):
        while self.queue:
            req = self.queue[0]
            need = self.footprint(req.spec)
            if self.in_use + need > self.budget:
                break
            if (self.max_active_nodes is not None and self.running
                    and self.active_nodes + req.spec.n > self.max_active_nodes):
                break
            fits = [i for i, ln in enumerate(self._lanes) if ln.eng.fits(req.spec)]
            if not fits:
                break
            li = min(fits, key=lambda i: self._lanes[i].active_nodes)
            lane = self._lanes[li]
            self.queue.popleft()
            with lane.on_stream():
                req.cid = lane.eng.add(req.spec)
            req.engine = li
            req.admitted = now
            req.first_tick = self._ticks
            req.left = int(req.spec.steps)
            self.in_use += need
            self.active_nodes += req.spec.n
            lane.active_nodes += req.spec.n
            self.running[(li, req.cid)] = req

    def _launch(self):
        """Enqueues one tick of every engine with steps left, and the
        snapshot copies behind each."""
        busy = {li for (li, _), r in self.running.items() if r.left > 0}
        tick = []
        for li in sorted(busy):
            lane = self._lanes[li]
            eng = lane.eng
            with lane.on_stream():
                if not self.graph:
                    eng.step(self.tick_steps)
                else:
                    if not lane.graph_ready:
                        eng.capture(self.tick_steps)
                        lane.graph_ready = True
                    eng.run_graph()
                snap = lane.free.pop()
                snap.record(eng, self._ticks)
            tick.append((li, snap))
        self._inflight.append(tick)
        for (li, _), req in self.running.items():
            if li in busy:
                req.left -= min(req.left, self.tick_steps)
        self._ticks += 1

    def _complete(self, clock: Callable[[], float], t0: float):
        """Waits for the oldest tick in flight and handles its snapshots."""
        tick = self._inflight.popleft()
        for li, snap in tick:
            snap.event.synchronize()
        now = clock() - t0
        for li, snap in tick:
            lane = self._lanes[li]
            eng = lane.eng
            left = snap.left.numpy()
            finished = []
            for (lj, cid), req in self.running.items():
                if lj != li or snap.tick < req.first_tick:
                    continue  # another engine's, or admitted after this tick was launched
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
                req = self.running.pop((li, cid))
                s = eng.slots[cid]
                req.result = ps[s.p_off : s.p_off + s.p_count].copy()
                req.finished = now
                self.in_use -= self.footprint(req.spec)
                self.active_nodes -= req.spec.n
                lane.active_nodes -= req.spec.n
                with lane.on_stream():
                    eng.remove(cid)
                self.done.append(req)
            lane.free.append(snap)

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

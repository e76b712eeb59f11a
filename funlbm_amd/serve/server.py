"""Continuous-batching server for settling-case requests.

Requests arrive over time (each a CaseSpec: lattice, particles, step count).
The server keeps their state resident in the BatchEngine's pools and admits a
waiting request as soon as its HBM footprint fits under the budget and the
pools have room; admitted cases step together, each to its own step count, in
ticks of `tick_steps` steps (one replay of a captured HIP graph whose kernels
read the live work lists from device memory, so admissions between ticks need
no re-capture).  After every tick the trajectories of the running cases are
copied back, finished cases are retired and their memory is reused.

No reference counterpart: the reference runs one case per process
(src/funlbm/server/script.py:10-24).
"""

import dataclasses
import time
from collections import deque
from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from ..batch.engine import BatchEngine, CaseSpec


@dataclasses.dataclass
class Request:
    rid: int
    spec: CaseSpec
    arrival: float  # seconds since the start of the run
    admitted: Optional[float] = None
    finished: Optional[float] = None
    cid: Optional[int] = None
    trajectory: List[np.ndarray] = dataclasses.field(default_factory=list)
    result: Optional[np.ndarray] = None  # final particle rows (double)

    @property
    def latency(self):
        return None if self.finished is None else self.finished - self.arrival


class Server:
    def __init__(self, engine: BatchEngine, hbm_budget: float, tick_steps: int = 50,
                 stream_trajectories: bool = True):
        if engine.traj_cap < tick_steps:
            raise ValueError("the trajectory ring must hold one tick of steps")
        self.engine = engine
        self.budget = float(hbm_budget)
        self.tick_steps = int(tick_steps)
        self.stream = stream_trajectories
        self.in_use = 0.0
        self.queue: deque = deque()
        self.running: Dict[int, Request] = {}
        self.done: List[Request] = []
        self._graph_ready = False
        self._storage_bytes = torch.tensor([], dtype=engine.pool[0].dtype).element_size()

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
            self.queue.popleft()
            req.cid = self.engine.add(req.spec)
            req.admitted = now
            self.in_use += need
            self.running[req.cid] = req

    def _collect(self, now: float):
        eng = self.engine
        if self.stream and self.running:
            # one copy of the whole ring per tick; each running case keeps the
            # entries of the steps it took in this tick
            traj = eng.traj.view(-1, eng.traj_cap, 4).cpu().numpy()
            for cid, req in self.running.items():
                s = eng.slots[cid]
                rows = traj[s.p_off : s.p_off + s.p_count]
                req.trajectory.append(rows.copy())
        left = eng.left()
        for cid in [c for c in self.running if left[c] <= 0]:
            req = self.running.pop(cid)
            req.result = eng.particle_rows(cid).cpu().numpy().copy()
            req.finished = now
            self.in_use -= self.footprint(req.spec)
            eng.remove(cid)
            self.done.append(req)

    def tick(self):
        eng = self.engine
        if not self._graph_ready:
            eng.capture(self.tick_steps)
            self._graph_ready = True
        eng.run_graph()
        torch.cuda.synchronize(eng.device)

    def run(self, requests: List[Request], clock: Callable[[], float] = time.perf_counter,
            max_seconds: float = 1e9) -> Dict:
        """Feeds `requests` at their arrival times and serves until all are
        done (or max_seconds pass).  Returns throughput and latency figures."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        t0 = clock()
        busy = 0.0
        while pending or self.queue or self.running:
            now = clock() - t0
            if now > max_seconds:
                break
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            self._admit(now)
            if not self.running:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.01)))
                    continue
                if self.queue:
                    raise MemoryError("a request does not fit the pools even when the server is idle")
                break
            ts = clock()
            self.tick()
            busy += clock() - ts
            self._collect(clock() - t0)
        wall = clock() - t0
        lat = np.array([r.latency for r in self.done if r.latency is not None])
        lups = sum(r.spec.n * r.spec.steps for r in self.done)
        return {
            "completed": len(self.done),
            "submitted": len(requests),
            "wall_s": wall,
            "busy_s": busy,
            "solves_per_s": len(self.done) / wall if wall > 0 else 0.0,
            "mlups": lups / wall / 1e6 if wall > 0 else 0.0,
            "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        }

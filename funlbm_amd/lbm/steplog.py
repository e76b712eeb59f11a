"""Per-step records of the run loop (src/funlbm/lbm/base.py:105-144), copied
back from the GPU asynchronously.

A record is the min/mean/max of the reference's f, u and rho (computed on
the device, by the fused step itself when it was asked for them), each
particle's state row (centre, velocities, force, rotation vector) and the
min/mean/max of its marker forces, and the weight-sum failure flag.  They are
gathered into one device vector, copied into a pinned host buffer behind an
event, and turned into the reference's log line, SQLite rows and history
entry once the copy has landed, so the step loop never waits for the GPU.
Records are emitted in step order, at most `depth` steps behind the loop
(older ones are waited for); run() flushes the rest before it returns.
"""

from collections import deque

import numpy as np
import torch

from .. import _native as nat


class StepLog:
    def __init__(self, lbm, depth: int = 8):
        self.lbm = lbm
        self.depth = max(1, int(depth))
        self.pending = deque()
        self.free = []
        self.events = []

    # ---------------------------------------------------------------- device side
    def _gather(self):
        """Device double vector: flow stats [9], then per particle lF min,
        mean, max [3P], state rows [40P], and the failure flag [1]."""
        lbm = self.lbm
        parts = [lbm.flow.stats().to(torch.float64)]
        sw = lbm.particle_swarm
        P = len(sw.particles)
        if not P:
            # copied in stream order, before a later step can overwrite it
            return parts[0]
        dev = sw.lF.device
        lf = sw.lF.reshape(-1)
        idx = sw.pid.long().repeat_interleave(3)
        inf = torch.full((P,), float("inf"), device=dev, dtype=torch.float32)
        lo = inf.scatter_reduce(0, idx, lf, "amin", include_self=True)
        hi = (-inf).scatter_reduce(0, idx, lf, "amax", include_self=True)
        # per-particle sums from a double prefix sum: fixed order
        cs = torch.cat([torch.zeros(1, device=dev, dtype=torch.float64), lf.double().cumsum(0)])
        off = sw.offsets.long() * 3
        cnt = (off[1:] - off[:-1]).clamp_min(1).double()
        mean = (cs[off[1:]] - cs[off[:-1]]) / cnt
        parts += [torch.stack([lo.double(), mean, hi.double()], 1).reshape(-1),
                  sw.pstate[:P].reshape(-1).to(torch.float64),
                  sw.ctl[1:2].to(torch.float64)]
        return torch.cat(parts)

    def record(self, step: int) -> None:
        """Queues the record of `step` (no host wait)."""
        vec = self._gather()
        host = None
        for i, h in enumerate(self.free):
            if h.numel() == vec.numel():
                host = self.free.pop(i)
                break
        if host is None:
            host = torch.empty(vec.numel(), dtype=torch.float64, pin_memory=True)
        host.copy_(vec, non_blocking=True)
        ev = self.events.pop() if self.events else torch.cuda.Event()
        ev.record()
        self.pending.append((step, host, ev, vec))
        self.drain()

    def drain(self, wait: bool = False) -> None:
        """Emits the records whose copies have landed (all of them if wait,
        and always enough to stay within `depth` records)."""
        while self.pending:
            step, host, ev, _vec = self.pending[0]
            if not (wait or len(self.pending) > self.depth or ev.query()):
                break
            ev.synchronize()
            self.pending.popleft()
            self._emit(step, host.numpy().copy())
            self.free.append(host)
            self.events.append(ev)

    # ------------------------------------------------------------------ host side
    def _emit(self, step: int, v: np.ndarray) -> None:
        lbm = self.lbm
        flow = {"f": v[0:3].tolist(), "u": v[3:6].tolist(), "rho": v[6:9].tolist()}
        particles = []
        sw = lbm.particle_swarm
        P = len(sw.particles)
        if P:
            lfs = v[9 : 9 + 3 * P].reshape(P, 3)
            rows = v[9 + 3 * P : 9 + 43 * P].reshape(P, nat.PSTRIDE)
            if v[9 + 43 * P] != 0:
                lbm.run_status = False
            for i, p in enumerate(sw.particles):
                r = rows[i]
                particles.append({
                    "m": float(p.mass),
                    "cu": r[nat.P_CU : nat.P_CU + 3].tolist(),
                    "cx": r[nat.P_CX : nat.P_CX + 3].tolist(),
                    "cf": r[nat.P_CF : nat.P_CF + 3].tolist(),
                    "lF": lfs[i].tolist(),
                    "cw": r[nat.P_CW : nat.P_CW + 3].tolist(),
                    "coord": {"center": r[nat.P_CX : nat.P_CX + 3].tolist(),
                              "angle": r[nat.P_ANGLE : nat.P_ANGLE + 3].tolist()},
                })
        lbm._emit_record(step, flow, particles)

"""Particle swarm (src/funlbm/particle/swarm.py:12-47) with concatenated
device arrays: all markers of all particles, particle-major, so every IBM
kernel covers the whole swarm in one launch."""

from typing import List

import numpy as np
import torch

from .. import _native as nat
from .base import Particle, ParticleConfig, create_particle


class ParticleSwarm:
    def __init__(self, configs: List[ParticleConfig] = (), device=None, gravity=None, gl=0.0):
        self.device = device
        self.particles: List[Particle] = [create_particle(c, device=device) for c in configs]
        self.gravity = gravity
        self.gl = gl
        self.n_markers = 0
        self.ready = False

    def init(self, *args, **kwargs):
        """Per-particle constants, then the device arrays."""
        for p in self.particles:
            p.init(*args, **kwargs)
        self._build()

    def _build(self):
        dev = self.device
        P = len(self.particles)
        counts = [p.n_markers for p in self.particles]
        offsets = np.zeros(P + 1, dtype=np.int32)
        offsets[1:] = np.cumsum(counts)
        M = int(offsets[-1])
        self.n_markers = M
        self.offsets_host = offsets
        self.offsets = torch.tensor(offsets, device=dev, dtype=torch.int32)
        pid = np.repeat(np.arange(P, dtype=np.int32), counts)
        self.pid = torch.tensor(pid, device=dev, dtype=torch.int32)
        lag = np.concatenate([p._lagrange for p in self.particles]) if M else np.zeros((0, 3), np.float32)
        self.lag = torch.tensor(lag, device=dev, dtype=torch.float32).reshape(M, 3)
        # the reference starts lx, lF, lu at zero (src/funlbm/particle/base/base.py:112-114);
        # lrou is never written by the reference
        self.lx = torch.zeros((M, 3), device=dev, dtype=torch.float32)
        self.lF = torch.zeros((M, 3), device=dev, dtype=torch.float32)
        self.lu = torch.zeros((M, 3), device=dev, dtype=torch.float32)
        self.lrou = torch.zeros((M, 1), device=dev, dtype=torch.float32)
        lm = np.concatenate([np.full(c, p.lm_value, dtype=np.float32) for c, p in zip(counts, self.particles)]) \
            if M else np.zeros(0, np.float32)
        self.lm = torch.tensor(lm, device=dev, dtype=torch.float32)
        ps = np.zeros((P, nat.PSTRIDE), dtype=np.float64)
        for i, p in enumerate(self.particles):
            ps[i, nat.P_CX : nat.P_CX + 3] = p.coord.initial_center
            ps[i, nat.P_ANGLE : nat.P_ANGLE + 3] = p.coord.initial_angle
            ps[i, nat.P_MASS] = p.mass
            ps[i, nat.P_INERTIA : nat.P_INERTIA + 3] = p.I
            ps[i, nat.P_RHO] = p.rou
            g = self.gravity if self.gravity is not None else [self.gl, 0.0, 0.0]
            ps[i, nat.P_GRAV : nat.P_GRAV + 3] = np.array(g, dtype=np.float32).astype(np.float64)
        self.pstate = torch.tensor(ps, device=dev, dtype=torch.float64)
        self.ctl = torch.zeros(4, device=dev, dtype=torch.int32)
        for i, p in enumerate(self.particles):
            p.bind(self, i, self.pstate[i])
        self.ready = P > 0
        if P:
            # rotation matrices of the initial rotation vectors (lx stays 0)
            lx_keep = torch.zeros_like(self.lx)
            nat.call("funlbm_particle_pose", P, nat.ptr(self.offsets), nat.ptr(self.lag), nat.ptr(lx_keep),
                     nat.ptr(self.pstate), 0, nat.stream_handle(dev))

    def set_gravity_from_gl(self, index, gl):
        if self.gravity is None:
            g = np.array([gl, 0.0, 0.0], dtype=np.float32).astype(np.float64)
            self.pstate[index, nat.P_GRAV : nat.P_GRAV + 3] = torch.tensor(g, device=self.device)

    def _range(self, only):
        if only is None:
            return len(self.particles), self.offsets, self.pstate
        i = int(only)
        return 1, self.offsets[i:], self.pstate[i:]

    def forces(self, dt, rouf=1.0, only=None):
        """Rigid-body update in double (particle_forces kernel)."""
        n, off, ps = self._range(only)
        if n == 0:
            return
        nat.call("funlbm_particle_forces", n, nat.ptr(off), nat.ptr(self.lx), nat.ptr(self.lF),
                 nat.ptr(self.lm), nat.ptr(ps), float(dt), float(rouf), nat.stream_handle(self.device))

    def pose(self, only=None):
        """angle += cw, rotation, lx = R lag + cx (particle_pose kernel)."""
        n, off, ps = self._range(only)
        if n == 0:
            return
        nat.call("funlbm_particle_pose", n, nat.ptr(off), nat.ptr(self.lag), nat.ptr(self.lx),
                 nat.ptr(ps), 1, nat.stream_handle(self.device))

    def update(self, *args, **kwargs):
        """src/funlbm/particle/swarm.py:23-25."""
        self.pose()

    def save_particle(self):
        """The reference writes the body-frame markers to an h5 file
        (src/funlbm/particle/swarm.py:27-31); not written here."""
        return None


def create_particle_swarm(configs: List[ParticleConfig] = (), device=None, *args, **kwargs):
    return ParticleSwarm(configs=configs, device=device, *args, **kwargs)

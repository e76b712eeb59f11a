"""Checkpoint archive layout (CPU): the reference's dataset paths
(src/funlbm/lbm/base.py:205-210, src/funlbm/particle/swarm.py:34-41,
src/funlbm/particle/base/base.py:190-202) as keys, and a dump -> load ->
dump round trip that reproduces every array."""

import types

import numpy as np
import torch

from funlbm_amd.lbm import checkpoint as ck


class _Flow:
    """f only (no FOL), as a flow without the fol_array hook."""
    def __init__(self, f):
        self._f = f
        self.updated = 0

    @property
    def f(self):
        return self._f

    @f.setter
    def f(self, v):
        self._f = v.clone()

    def update_u_rou(self):
        self.updated += 1


def _solver(seed):
    g = torch.Generator().manual_seed(seed)
    parts = []
    for m in (5, 3):
        p = types.SimpleNamespace()
        for key in ck.PARTICLE_KEYS:
            shape = (m, 1) if key == "lrou" else ((m, 3) if key.startswith("l") else (3,))
            setattr(p, key, torch.rand(shape, generator=g, dtype=torch.float64 if key[0] == "c" else torch.float32))
        parts.append(p)
    sw = types.SimpleNamespace(particles=parts, pstate=torch.rand((2, 40), generator=g, dtype=torch.float64))
    return types.SimpleNamespace(step=17, flow=_Flow(torch.rand((4, 3, 5, 19), generator=g)), particle_swarm=sw)


def test_layout_and_round_trip(tmp_path):
    a = _solver(0)
    path = ck.dump(a, str(tmp_path / "checkpoint_0000000017.npz"))
    with np.load(path, allow_pickle=False) as z:
        keys = set(z.files)
        assert {"step", "flow/f"} <= keys
        for i in range(2):
            for key in ck.PARTICLE_KEYS + ("state",):
                assert f"particle/particle_{i:06d}/{key}" in keys
        assert z["flow/f"].shape == (4, 3, 5, 19) and z["step"][0] == 17
    b = _solver(1)
    ck.load(b, path)
    assert b.step == 17 and b.flow.updated == 1
    assert torch.equal(b.flow.f, a.flow.f)
    assert torch.equal(b.particle_swarm.pstate, a.particle_swarm.pstate)
    for pa, pb in zip(a.particle_swarm.particles, b.particle_swarm.particles):
        for key in ck.PARTICLE_KEYS:
            assert torch.equal(getattr(pa, key), getattr(pb, key))

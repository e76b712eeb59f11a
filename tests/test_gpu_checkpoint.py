"""Checkpoint resume (needs a GPU): a resumed run equals the uninterrupted
one bit for bit, with and without particles (the archive also keeps FOL,
which the reference's load leaves at zero)."""

import numpy as np
import pytest
import torch

from oracle import lbm_oracle as orc
from tests.helpers import flow_json

pytestmark = pytest.mark.gpu

W, P = orc.WALL, orc.PERIODICAL


def solver(particles=(), **extra):
    from funlbm_amd.lbm import LBMD3, Config

    flow = flow_json(19, (9, 10, 11), 0.05, [W, W, P, P, P, P])
    flow["body_force"] = [0.0, 2e-5, 0.0]
    cfg = {"flow": flow, "particles": list(particles), "log_every": 0}
    cfg.update(extra)
    lbm = LBMD3(config=Config().from_json(cfg))
    lbm.init()
    return lbm


def advance(lbm, first, last):
    for s in range(first, last + 1):
        lbm.step = s
        lbm.run_step()


def test_resume_without_particles_is_exact(tmp_path):
    a = solver()
    advance(a, 1, 12)
    b = solver()
    advance(b, 1, 7)
    path = b.dump_checkpoint(str(tmp_path / "c.npz"))
    c = solver(checkpoint=path)
    assert c.step == 7
    advance(c, 8, 12)
    assert torch.equal(a.flow.f, c.flow.f)


def test_resume_with_a_sphere(tmp_path):
    sphere = [{"type": "sphere", "r": 3, "rou": 1.5, "coord": {"center": [5, 5, 6]}}]
    b = solver(sphere)
    advance(b, 1, 6)
    path = b.dump_checkpoint(str(tmp_path / "c.npz"))
    c = solver(sphere, checkpoint=path)
    assert torch.equal(c.particle_swarm.pstate, b.particle_swarm.pstate)
    assert torch.equal(c.particle_swarm.lx, b.particle_swarm.lx)
    assert torch.equal(c.flow.f, b.flow.f)
    advance(b, 7, 10)
    advance(c, 7, 10)
    assert torch.equal(c.particle_swarm.pstate, b.particle_swarm.pstate)
    assert torch.equal(c.flow.f, b.flow.f)


def test_periodic_checkpoints_are_written(tmp_path):
    lbm = solver(checkpoint_every=3, file={"cache_dir": str(tmp_path)})
    advance(lbm, 1, 7)
    files = sorted(p.name for p in (tmp_path / "checkpoint").iterdir())
    assert files == ["checkpoint_0000000003.npz", "checkpoint_0000000006.npz"]


def test_run_writes_step_records(tmp_path):
    from funlbm_amd.lbm.store import StepStore

    sphere = [{"type": "sphere", "r": 3, "rou": 1.5, "coord": {"center": [5, 5, 6]}}]
    db = str(tmp_path / "run.db")
    lbm = solver(sphere, db_path=db, log_every=2, max_step=6)
    lbm.run()
    lbm.store.close()
    s = StepStore(db)
    for step in (2, 4, 6):
        rec = s.flow(step)
        assert set(rec) == {"f", "u", "rho"} and len(rec["rho"]) == 3
        assert "cx" in s.particle(step, 1)
    assert s.flow(3) is None

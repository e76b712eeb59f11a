"""Continuous-batching server (needs a GPU): requests arriving while others
run are admitted as memory frees up, and every result equals the same request
solved alone."""

import numpy as np
import pytest
import torch

from funlbm_amd.batch.engine import BatchEngine, CaseSpec, ParticleSpec
from funlbm_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def small_requests():
    specs = [
        CaseSpec(nodes=(20, 20, 24), q=19, tau=0.6, steps=120,
                 particles=[ParticleSpec(kind="sphere", r=3.5, rou=1.5, center=(10, 10, 15))]),
        CaseSpec(nodes=(24, 18, 20), q=27, tau=0.7, steps=75,
                 particles=[ParticleSpec(kind="ellipsoid", abc=(2.5, 2.5, 4), rou=2.0, center=(12, 9, 11),
                                         angle=(0.2, 0.4, 0.1))]),
        CaseSpec(nodes=(28, 16, 30), q=19, tau=0.9, steps=200,
                 particles=[ParticleSpec(kind="sphere", r=3.0, rou=2.5, center=(8, 8, 20)),
                            ParticleSpec(kind="sphere", r=3.0, rou=2.0, center=(20, 8, 20))]),
        CaseSpec(nodes=(22, 22, 22), q=27, tau=0.55, steps=50,
                 particles=[ParticleSpec(kind="sphere", r=4.0, rou=3.0, center=(11, 11, 13))]),
        CaseSpec(nodes=(18, 26, 18), q=19, tau=0.8, steps=150,
                 particles=[ParticleSpec(kind="ellipsoid", abc=(3, 2, 2), rou=1.7, center=(9, 13, 10))]),
    ]
    arrivals = [0.0, 0.0, 0.05, 0.1, 0.3]
    return [Request(rid=i, spec=s, arrival=a) for i, (s, a) in enumerate(zip(specs, arrivals))]


def engine():
    return BatchEngine(storage="f16", max_nodes=1 << 16, max_markers=8192, max_particles=16, max_cases=8,
                       traj_cap=32)


def test_server_mix_equals_alone_and_respects_budget():
    reqs = small_requests()
    eng = engine()
    # room for about two of these cases at a time
    budget = 2.2 * max(BatchEngine.case_bytes_needed(r.spec) for r in reqs)
    srv = Server(eng, hbm_budget=budget, tick_steps=25)
    peak = []
    orig_admit = srv._admit

    def admit(now):
        orig_admit(now)
        peak.append(srv.in_use)

    srv._admit = admit
    stats = srv.run(reqs)
    assert stats["completed"] == len(reqs)
    assert max(peak) <= budget
    # requests overlapped: someone was admitted while another was running
    by_adm = sorted(reqs, key=lambda r: r.admitted)
    assert any(b.admitted < a.finished for a, b in zip(by_adm, by_adm[1:]))
    for r in reqs:
        alone = engine()
        cid = alone.add(r.spec)
        alone.step(r.spec.steps)
        want = alone.particle_rows(cid).cpu().numpy()
        assert np.array_equal(r.result[:, :15], want[:, :15]), r.rid
        # the streamed trajectory ends at the final centre
        last = r.trajectory[-1]
        slot = (r.spec.steps - 1) % eng.traj_cap
        assert last[0, slot, 3] == r.spec.steps
        assert np.allclose(last[:, slot, :3], want[:, 0:3].astype(np.float32))
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0

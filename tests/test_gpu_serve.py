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


@pytest.mark.parametrize("graph,engines", [(False, 1), (True, 1), (False, 2), (True, 2)],
                         ids=["eager", "graph", "eager-2-engines", "graph-2-engines"])
def test_server_mix_equals_alone_and_respects_budget(graph, engines):
    """Eager ticks (grids from the live counts) and graph-replayed ticks
    (grids of the capture), with one engine or two on streams of their own,
    give the same bits as each request alone."""
    reqs = small_requests()
    engs = [engine() for _ in range(engines)]
    # room for about two of these cases at a time
    budget = 2.2 * max(BatchEngine.case_bytes_needed(r.spec) for r in reqs)
    srv = Server(engs if engines > 1 else engs[0], hbm_budget=budget, tick_steps=25, graph=graph)
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
    assert {r.engine for r in reqs} == set(range(engines))
    for r in reqs:
        alone = engine()
        cid = alone.add(r.spec)
        alone.step(r.spec.steps)
        want = alone.particle_rows(cid).cpu().numpy()
        assert np.array_equal(r.result[:, :15], want[:, :15]), r.rid
        # the streamed trajectory holds every step once, in order, and ends
        # at the final centre
        traj = np.concatenate(r.trajectory, axis=1)
        assert np.array_equal(traj[0, :, 3], np.arange(1, r.spec.steps + 1, dtype=np.float32))
        assert np.allclose(traj[:, -1, :3], want[:, 0:3].astype(np.float32))
    assert stats["p99_latency_s"] >= stats["p50_latency_s"] > 0


def _engine_for(specs, storage="f16", traj_cap=32):
    from funlbm_amd.batch.engine import markers_for

    nodes = sum(s.n for s in specs)
    markers = sum(len(markers_for(p)[0]) for s in specs for p in s.particles)
    return BatchEngine(storage=storage, max_nodes=nodes + 4096 * len(specs), max_markers=markers + 64,
                       max_particles=4 * len(specs) + 4, max_cases=len(specs) + 1, max_q=27, traj_cap=traj_cap)


def test_serving_stream_mix_equals_alone_and_f16_error():
    """A short Poisson stream from the serving workload's generator (sides
    32-48, 300-600 steps; D3Q19 and D3Q27 requests with spheres and
    ellipsoids), served with continuous batching: every result equals the
    request solved alone, and on a D3Q27 request the f16 velocity field stays
    within 1e-3 (relative L-inf) of the float64 solve, the centroid within
    1e-5 lattice units."""
    from funlbm_amd.batch.workloads import serving_requests

    stream = serving_requests(40.0, 0.25, seed=0, side=(32, 48), steps=(300, 600))
    specs = [s for _, s in stream]
    assert {s.q for s in specs} == {19, 27}
    assert {p.kind for s in specs for p in s.particles} == {"sphere", "ellipsoid"}
    reqs = [Request(rid=i, spec=s, arrival=t) for i, (t, s) in enumerate(stream)]
    srv = Server([_engine_for(specs), _engine_for(specs)], hbm_budget=200e9, tick_steps=25)
    stats = srv.run(reqs)
    assert stats["completed"] == len(reqs) and stats["backlog"] == 0
    for r in reqs:
        alone = _engine_for([r.spec])
        cid = alone.add(r.spec)
        alone.step(r.spec.steps)
        want = alone.particle_rows(cid).cpu().numpy()
        assert np.array_equal(r.result[:, :15], want[:, :15]), r.rid
        del alone
    torch.cuda.empty_cache()
    spec = next(s for s in specs if s.q == 27)
    out = {}
    for st in ("f64", "f16"):
        e = _engine_for([spec], storage=st)
        cid = e.add(spec)
        e.step(spec.steps)
        out[st] = (e.fields(cid)[0].double().cpu().numpy(), e.particle_rows(cid)[:, 0:3].cpu().numpy())
        del e
    u64, c64 = out["f64"]
    u16, c16 = out["f16"]
    err = float(np.abs(u16 - u64).max() / np.abs(u64).max())
    cerr = float(np.abs(c16 - c64).max())
    print(f"D3Q27 {spec.nodes} {len(spec.particles)} particles {spec.steps} steps: f16 u rel L-inf {err:.3e} "
          f"(max|u| {np.abs(u64).max():.3e}), centroid {cerr:.3e}")
    assert err <= 1e-3, err
    assert cerr <= 1e-5, cerr

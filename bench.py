"""Benchmark of the fused D3Q19 BGK timestep on one MI355X.

Workload (the headline metric's configuration): D3Q19, float32 populations,
256 x 256 x 256 nodes, all six faces PERIODICAL, tau = 0.6, no particles,
initial state f = feq(rho, u) with rho = 1 + 1e-3 N(0,1), u = 1e-3 N(0,1)^3
(seed 0).  A step is one LBMD3.run_step: one fused pull-stream + moments +
collision pass over the lattice.  Inputs are resident in HBM before timing.

Prints one JSON line with MLUPS, the HBM roofline of the fused kernel (HIP
events on the launch stream; algorithmic bytes 2*Q*sizeof(storage) per node)
and a CPU baseline (the reference's torch op sequence on the host's threads,
a bounded sample at 256^3).  Further legs: f16/bf16 deviation storage, D3Q27,
the 64-ellipsoid suspension (timing and, at 2000 steps, its precision against
float64), 256 batched settling cases against one-at-a-time solves, and a
serving sweep over arrival rates.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & %HBM roofline (D3Q19 256³); batched solves/s + p99 at fixed error"


def build_solver(n, q, storage, tau=0.6, seed=0, rounding=None):
    import torch

    from funlbm_amd import _native as nat
    from funlbm_amd.lbm import LBMD3, Config

    mu = (tau - 0.5) / 3.0
    faces = {f: {"code": "PERIODICAL"} for f in ("input", "output", "back", "forward", "bottom", "top")}
    cfg = Config().from_json(
        {
            "flow": {
                "param_type": f"D3Q{q}",
                "size": [n - 1, n - 1, n - 1],
                "param": {"mu": mu},
                "boundary": faces,
                "storage": storage,
                **({"rounding": rounding} if rounding else {}),
            },
            "log_every": 0,
        }
    )
    lbm = LBMD3(config=cfg)
    lbm.init()
    fl = lbm.flow
    g = torch.Generator(device=fl.device)
    g.manual_seed(seed)
    rho = 1.0 + 1e-3 * torch.randn(fl.n, device=fl.device, generator=g)
    u = 1e-3 * torch.randn(3, fl.n, device=fl.device, generator=g)
    nat.call("funlbm_equilibrium_state", fl.lat32, nat.ptr(rho), nat.ptr(u), nat.ptr(fl._f), fl._stream())
    fl._pristine = False
    fl.update_u_rou()
    return lbm


def time_steps(lbm, steps, warmup, repeats=5):
    """Per-step time of exactly `steps` timed steps (after `warmup` untimed
    ones), split into `repeats` consecutive chunks each bracketed by HIP events
    on the launch stream; the median chunk's per-step time."""
    import torch

    for _ in range(warmup):
        lbm.run_step()
    torch.cuda.synchronize()
    repeats = max(1, min(repeats, steps))
    chunks = [steps // repeats] * repeats
    chunks[-1] += steps - sum(chunks)
    events = [torch.cuda.Event(enable_timing=True) for _ in range(repeats + 1)]
    t0 = time.perf_counter()
    events[0].record()
    for i, c in enumerate(chunks):
        for _ in range(c):
            lbm.run_step()
        events[i + 1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    runs = [events[i].elapsed_time(events[i + 1]) / c for i, c in enumerate(chunks)]
    return float(np.median(runs)), wall


def time_fused_kernel(lbm, reps=50):
    """Average duration (ms) of one funlbm_fused_step launch: `reps` launches
    back to back on the stream the step runs on, ping-ponging the flow's two
    state buffers, bracketed by events recorded on that same stream."""
    import torch

    from funlbm_amd import _native as nat

    fl = lbm.flow
    fl.begin_step()
    a, b = fl._state[fl._cur], fl._state[1 - fl._cur]
    stream = torch.cuda.current_stream()
    s = nat.stream_handle()

    def launch(src, dst):
        nat.call("funlbm_fused_step", fl.lat, nat.ptr(src), nat.ptr(dst), None, None, None, None, None, None, s)

    for i in range(4):
        launch(a, b) if i % 2 == 0 else launch(b, a)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for i in range(reps):
        launch(a, b) if i % 2 == 0 else launch(b, a)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "k_fused_d3q19_256_traffic.md")


def profiled_traffic(storage, size, q):
    """HBM bytes per k_fused launch from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (copy-calibrated) of this workload, or None."""
    if size != 256 or q != 19 or not os.path.exists(TRAFFIC_PROFILE):
        return None
    for line in open(TRAFFIC_PROFILE):
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == storage:
            return int(parts[2])
    return None


def storage_leg(size, q, storage, steps, warmup, ref_mlups=None, rounding=None):
    """Step and kernel timing of one lattice/storage at size^3 (periodic, no particles)."""
    import torch

    lbm = build_solver(size, q, storage, rounding=rounding)
    ms, _ = time_steps(lbm, steps, max(warmup, 50))
    kms = time_fused_kernel(lbm)
    n = lbm.flow.n
    del lbm
    torch.cuda.empty_cache()
    elem = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}[storage]
    mlups = n / (ms * 1e-3) / 1e6
    ach = n * 2 * q * elem / (kms * 1e-3) / 1e9
    eff = rounding or ("stochastic" if storage in ("f16", "bf16") else None)
    rd = f", {eff} rounding" if eff else ""
    out = {"workload": f"D3Q{q} {size}^3 periodic, tau 0.6, {storage} state{rd}", "mlups": round(mlups, 1),
           "ms_per_step": round(ms, 4), "kernel_ms": round(kms, 4), "bytes_per_lu": 2 * q * elem,
           "achieved_gbs": round(ach, 1), "frac_of_8tbs": round(ach / HBM_PEAK_GBS, 4)}
    if ref_mlups:
        out["speedup_vs_f32"] = round(mlups / ref_mlups, 3)
    return out


def bench_suspension(size=256, n_particles=64, steps=30, warmup=5, storage="f32"):
    """The D3Q27 suspension: 64 prolate ellipsoids (4, 4, 8) at seeded
    non-overlapping positions and random orientations in a periodic 256^3
    box, whole LBMD3 steps (IBM included)."""
    import torch

    from funlbm_amd import precision
    from funlbm_amd.lbm import LBMD3, Config

    cfg = precision.suspension(n=size, n_particles=n_particles, mu=0.05 / 3)
    cfg["flow"]["storage"] = storage
    t0 = time.perf_counter()
    lbm = LBMD3(config=Config().from_json(cfg))
    lbm.init()
    init_s = time.perf_counter() - t0
    for k in range(1, warmup + 1):
        lbm.step = k
        lbm.run_step()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for k in range(warmup + 1, warmup + steps + 1):
        lbm.step = k
        lbm.run_step()
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / steps
    n = lbm.flow.n
    m = lbm.particle_swarm.n_markers
    p = len(lbm.particle_swarm.particles)
    del lbm
    torch.cuda.empty_cache()
    return {"workload": f"D3Q27 {size}^3 periodic, {p} ellipsoids (4,4,8), rho_p 1.2, body force 1e-6 x, "
                        f"{storage} state, whole steps with IBM", "particles": p, "markers": m,
            "ms_per_step": round(ms, 4), "mlups": round(n / (ms * 1e-3) / 1e6, 1), "init_s": round(init_s, 2)}


def bench_sphere(steps=200, warmup=20):
    """The free-fall configuration: D3Q19 64 x 64 x 128 nodes, tau 0.6,
    walls, a sphere of radius 6 at (32, 32, 100) on Fibonacci markers with
    Peskin's 4-point kernel, density ratio 1.5, gravity 2e-5 along -z, f32;
    whole LBMD3 steps, launched eagerly and replayed from HIP graphs."""
    import torch

    from funlbm_amd.lbm import LBMD3, Config

    out = {"workload": "D3Q19 64x64x128 walls, tau 0.6, sphere r 6 at (32,32,100), Fibonacci markers, "
                       "Peskin 4-point kernel, rho_p 1.5, gravity 2e-5 -z, f32 state, whole steps"}
    for graph in (False, True):
        faces = {f: {"code": "WALL"} for f in ("input", "output", "back", "forward", "bottom", "top")}
        cfg = Config().from_json({
            "flow": {"param_type": "D3Q19", "size": [63, 63, 127], "param": {"mu": 0.1 / 3, "gravity": [0, 0, 2e-5]},
                     "boundary": faces, "ibm_delta": "peskin4"},
            "particles": [{"type": "sphere", "r": 6, "rou": 1.5, "markers": "fibonacci",
                           "coord": {"center": [32, 32, 100]}}],
            "log_every": 0, "graph": graph,
        })
        lbm = LBMD3(config=cfg)
        lbm.init()
        for k in range(1, warmup + 1):
            lbm.step = k
            lbm.run_step()
        torch.cuda.synchronize()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for k in range(warmup + 1, warmup + steps + 1):
            lbm.step = k
            lbm.run_step()
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / steps
        out["graph_ms_per_step" if graph else "eager_ms_per_step"] = round(ms, 4)
        out["markers"] = lbm.particle_swarm.n_markers
        out["mlups_graph" if graph else "mlups_eager"] = round(lbm.flow.n / (ms * 1e-3) / 1e6, 1)
        del lbm
        torch.cuda.empty_cache()
    return out


def copy_bandwidth(nbytes=2 << 30, reps=10):
    """Measured copy peak: read + write bytes of a plain copy kernel / time."""
    import torch

    from funlbm_amd import _native as nat

    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    s = nat.stream_handle()
    for _ in range(2):
        nat.call("funlbm_copy_calibrate", nat.ptr(src), nat.ptr(dst), nbytes, s)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        nat.call("funlbm_copy_calibrate", nat.ptr(src), nat.ptr(dst), nbytes, s)
    b.record()
    torch.cuda.synchronize()
    ms = a.elapsed_time(b) / reps
    del src, dst
    return 2.0 * nbytes / (ms * 1e-3) / 1e9


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def numpy_oracle_rate(n=64, q=19, budget_s=8.0):
    """The numpy oracle's fluid step on an n^3 periodic lattice, one core (bounded)."""
    from oracle import lbm_oracle as orc

    nodes = (n, n, n)
    st = orc.init_flow(q, nodes, tau=np.float32(0.6), bc=[orc.PERIODICAL] * 6)
    rng = np.random.default_rng(0)
    rho = (1.0 + 1e-3 * rng.standard_normal(nodes + (1,))).astype(np.float32)
    u = (1e-3 * rng.standard_normal(nodes + (3,))).astype(np.float32)
    st.f = orc.equilibrium(st.lat, rho, u)
    orc.update_u_rou(st)
    orc.fluid_step(st)  # warm
    steps = 0
    t0 = time.perf_counter()
    while True:
        orc.fluid_step(st)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 200:
            break
    return {"value": round(n ** 3 * steps / el / 1e6, 3), "unit": "MLUPS", "cores": 1,
            "sample": f"numpy oracle fluid step, D3Q{q} float32, {n}^3 periodic, {steps} steps in {el:.1f} s"}


def cpu_baseline(n=256, q=19, budget_s=12.0):
    """The reference's op sequence on CPU torch (oracle/torch_ref.py: matmul,
    norm, roll, the per-direction forcing loop) at the headline lattice, on
    all the threads torch uses on this host, timed for about budget_s after
    one warm-up step; the numpy oracle on one core beside it."""
    import torch

    from oracle import lbm_oracle as orc
    from oracle.torch_ref import TorchFlow

    nodes = (n, n, n)
    tf = TorchFlow(q, nodes, tau=np.float32(0.6), bc=[orc.PERIODICAL] * 6)
    g = torch.Generator().manual_seed(0)
    tf.rou = 1.0 + 1e-3 * torch.randn(nodes + (1,), generator=g)
    tf.u = 1e-3 * torch.randn(nodes + (3,), generator=g)
    tf.cul_equ()
    tf.f = tf.feq.clone()  # f = feq(rho, u)
    tf.update_u_rou()
    t0 = time.perf_counter()
    tf.step()  # warm
    warm = time.perf_counter() - t0
    steps = 0
    t0 = time.perf_counter()
    while True:
        tf.step()
        steps += 1
        el = time.perf_counter() - t0
        if el + el / steps > budget_s:
            break
    del tf
    mlups = n ** 3 * steps / el / 1e6
    return {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": torch.get_num_threads(),
        "kind": "port",
        "sample": f"the reference's torch op sequence (oracle/torch_ref.py), D3Q{q} float32, {n}^3 periodic, "
                  f"tau 0.6, {steps} steps in {el:.1f} s after a {warm:.1f} s warm-up step",
        "torch_threads": torch.get_num_threads(),
        "os_cpu_count": os.cpu_count(),
        "cpu_model": cpu_model(),
        "numpy_oracle": numpy_oracle_rate(),
    }


def batch_on_streams(specs, n_engines=2, storage="f16", tick=50):
    """The same cases split over `n_engines` engines (alternately by size),
    each replaying its own captured graph on a stream of its own, so one
    engine's latency-bound IBM kernels run beside another's fused kernels.
    Seconds until every case has taken its steps."""
    import torch

    from funlbm_amd.batch.engine import BatchEngine, markers_for

    order = sorted(range(len(specs)), key=lambda i: specs[i].n)
    engs = []
    for j in range(n_engines):
        group = [specs[i] for i in order[j::n_engines]]
        nodes = sum(s.n for s in group)
        markers = sum(len(markers_for(p)[0]) for s in group for p in s.particles)
        e = BatchEngine(storage=storage, max_nodes=nodes + 256 * len(group) + 4096, max_markers=markers + 64,
                        max_particles=len(group) + 1, max_cases=len(group), max_q=19, traj_cap=tick)
        for s in group:
            e.add(s)
        e.exact_grids = True  # one block per tile: the grid the autotuner picks for this workload
        engs.append(e)
    streams = [torch.cuda.Stream() for _ in engs]
    main = torch.cuda.current_stream()
    for e, st in zip(engs, streams):
        st.wait_stream(main)
        with torch.cuda.stream(st):
            e.capture(tick)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for st in streams:
        st.wait_stream(main)
    rounds = -(-max(s.steps for s in specs) // tick)
    for _ in range(rounds):
        for e, st in zip(engs, streams):
            with torch.cuda.stream(st):
                e.run_graph()
    for st in streams:
        main.wait_stream(st)
    b.record()
    torch.cuda.synchronize()
    assert all(int(e.case_left.max().item()) == 0 for e in engs)
    del engs
    torch.cuda.empty_cache()
    return a.elapsed_time(b) * 1e-3


def bench_batch(n_cases=256, steps=3000, seed=0, storage="f16", tick=50, baseline_per_lattice=8,
                baseline_steps=200):
    """Batched independent settling cases (all cases advance in one persistent
    launch sequence per step, replayed from a HIP graph) against plain
    one-case-at-a-time solves of the same cases with LBMD3."""
    import torch

    from funlbm_amd.batch.engine import BatchEngine, markers_for, prepare_markers
    from funlbm_amd.batch.workloads import settling_batch
    from funlbm_amd.lbm import LBMD3, Config

    specs = settling_batch(n_cases=n_cases, seed=seed, steps=steps)
    nodes = sum(s.n for s in specs)
    # every case's particle shapes, generated in one launch (timed cold)
    all_particles = [p for s in specs for p in s.particles]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prepare_markers(all_particles)
    marker_ms = (time.perf_counter() - t0) * 1e3
    markers = sum(len(markers_for(p)[0]) for s in specs for p in s.particles)
    eng = BatchEngine(storage=storage, max_nodes=nodes + 256 * n_cases, max_markers=markers + 64,
                      max_particles=n_cases, max_cases=n_cases, max_q=19, traj_cap=tick)
    for s in specs:
        eng.add(s)
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    a.record()
    tune = eng.autotune(steps=5)
    eng.capture(tick)
    while int(eng.case_left.max().item()) > 0:
        eng.run_graph()
    b.record()
    torch.cuda.synchronize()
    t = a.elapsed_time(b) * 1e-3
    wall = time.perf_counter() - t0
    assert all(eng.left()[c] == 0 for c in eng.slots)
    grid = dict(eng.grid)
    lups = sum(s.n * s.steps for s in specs)
    # one case at a time: the single-lattice solver, eager, same storage, on
    # a sample of each lattice's cases at evenly spaced size quantiles
    # (smallest and largest included); the mean solve time weights each
    # lattice's sample mean by its share of the cases
    sample = []
    for q in sorted({s.q for s in specs}):
        ids = sorted((i for i, s in enumerate(specs) if s.q == q), key=lambda i: specs[i].n)
        k = min(baseline_per_lattice, len(ids))
        picks = sorted({ids[round(j * (len(ids) - 1) / max(k - 1, 1))] for j in range(k)})
        sample.append((q, len(ids), picks))
    per_case = {}
    if baseline_per_lattice <= 0:
        sample = []
    for s in [specs[i] for _, _, picks in sample for i in picks]:
        p = s.particles[0]
        faces = {f: {"code": "WALL"} for f in ("input", "output", "back", "forward", "bottom", "top")}
        cfg = Config().from_json({
            "flow": {"param_type": f"D3Q{s.q}", "size": [n - 1 for n in s.nodes],
                     "param": {"mu": (s.tau - 0.5) / 3, "gravity": list(s.gravity)},
                     "boundary": faces, "storage": storage},
            "particles": [{"type": "sphere", "r": p.r, "rou": p.rou, "coord": {"center": list(p.center)}}],
            "log_every": 0,
        })
        lbm = LBMD3(config=cfg)
        lbm.init()
        for k in range(1, 11):
            lbm.step = k
            lbm.run_step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(11, 11 + baseline_steps):
            lbm.step = k
            lbm.run_step()
        torch.cuda.synchronize()
        per_case[id(s)] = (time.perf_counter() - t0) / baseline_steps * s.steps
        del lbm
    mean_s = sum(cnt / len(specs) * float(np.mean([per_case[id(specs[i])] for i in picks]))
                 for _, cnt, picks in sample)
    base_solves = 1.0 / mean_s if sample else None
    n_base = sum(len(p) for _, _, p in sample)
    sizes = [specs[i].n for _, _, p in sample for i in p] or [0]
    solves = n_cases / t
    del eng
    torch.cuda.empty_cache()
    t2 = batch_on_streams(specs, 2, storage=storage, tick=tick)
    return {
        "workload": f"{n_cases} settling cases, sides log-uniform 32-96, D3Q15/19, one sphere each, "
                    f"{steps} steps, {storage}-deviation storage",
        "cases": n_cases,
        "nodes": nodes,
        "markers": markers,
        "marker_generation_ms": round(marker_ms, 2),
        "marker_shapes": len({(p.kind, p.r, tuple(p.abc or ())) for p in all_particles}),
        "seconds": round(t, 3),
        "wall_seconds": round(wall, 3),
        "solves_per_s": round(solves, 2),
        "mlups": round(lups / t / 1e6, 1),
        "one_at_a_time_solves_per_s": round(base_solves, 3) if base_solves else None,
        "one_at_a_time_sample": f"{n_base} cases ({', '.join(f'{len(p)} D3Q{q}' for q, _, p in sample)}; "
                                f"{min(sizes)} to {max(sizes)} nodes, evenly spaced size quantiles) x "
                                f"{baseline_steps} steps with LBMD3, scaled to {steps} steps, weighted by lattice",
        "speedup_vs_one_at_a_time": round(solves / base_solves, 2) if base_solves else None,
        "grid": grid,
        "autotune": tune,
        "two_streams": {
            "workload": "the same cases split over two engines, each replaying its own graph on a stream of its "
                        "own (one engine's IBM kernels beside the other's fused kernels)",
            "seconds": round(t2, 3),
            "solves_per_s": round(n_cases / t2, 2),
            "speedup_vs_one_at_a_time": round(n_cases / t2 / base_solves, 2) if base_solves else None,
        },
    }


def serve_rate(rate, duration, seed=0, steps=(1000, 20000), budget=200e9, storage="f16", tick=50,
               drain_s=6.0, pool_nodes_cap=1 << 28, max_active_nodes=None, engines=3, depth=2):
    """Poisson stream of settling requests at `rate` req/s for `duration` s,
    served with continuous batching by `engines` engines on streams of their
    own (each with pools for the whole stream); the run is cut `drain_s`
    after the last arrival (requests still queued or running then are the
    backlog)."""
    import torch

    from funlbm_amd.batch.engine import BatchEngine, markers_for, prepare_markers
    from funlbm_amd.batch.workloads import serving_requests
    from funlbm_amd.serve import Request, Server

    reqs = serving_requests(rate, duration, seed=seed, steps=steps)
    prepare_markers([p for _, s in reqs for p in s.particles])
    need_nodes = sum(s.n for _, s in reqs)
    pool_nodes = min(need_nodes + 4096 * len(reqs), pool_nodes_cap)
    markers = sum(len(markers_for(p)[0]) for _, s in reqs for p in s.particles)
    engs = [BatchEngine(storage=storage, max_nodes=pool_nodes, max_markers=markers + 64,
                        max_particles=4 * len(reqs) + 4, max_cases=len(reqs) + 1, max_q=27, traj_cap=tick)
            for _ in range(engines)]
    srv = Server(engs, hbm_budget=budget, tick_steps=tick, max_active_nodes=max_active_nodes, depth=depth)
    stats = srv.run([Request(rid=i, spec=s, arrival=t) for i, (t, s) in enumerate(reqs)],
                    max_seconds=duration + drain_s)
    # lattice updates done by the GPU in the run, finished or not
    stats["glups_stepped"] = sum(r.spec.n * (r.spec.steps - r.left) for r in srv.done + list(srv.running.values())) \
        / stats["wall_s"] / 1e9
    # steady-state throughput: the lattice updates stepped per second over the
    # mean lattice updates of a request of this stream (completions in a short
    # window from a cold start over-count short requests and miss the ramp)
    mean_work = float(np.mean([s.n * s.steps for _, s in reqs]))
    stats["solves_per_s_steady"] = stats["glups_stepped"] * 1e9 / mean_work
    del srv, engs
    torch.cuda.empty_cache()
    out = {"rate": rate}
    for k in ("submitted", "completed", "backlog", "solves_per_s", "solves_per_s_steady", "mlups", "glups_stepped",
              "p50_latency_s", "p99_latency_s", "wall_s", "wait_s", "ticks"):
        v = stats.get(k)
        out[k] = round(v, 3) if isinstance(v, float) else v
    return out


def serve_error_check(seed=0, steps=(1000, 20000), storage="f16", index=0):
    """The served storage's error against float64 on request `index` of the
    stream, solved alone (float32 storage beside it)."""
    import torch

    from funlbm_amd.batch.engine import BatchEngine, markers_for
    from funlbm_amd.batch.workloads import serving_requests

    spec = serving_requests(20.0, 1.0, seed=seed, steps=steps)[index][1]
    fields = {}
    for st in ("f64", "f32", storage):
        e = BatchEngine(storage=st, max_nodes=spec.n + 16384,
                        max_markers=sum(len(markers_for(p)[0]) for p in spec.particles) + 8,
                        max_particles=len(spec.particles) + 1, max_cases=1, max_q=27, traj_cap=8)
        cid = e.add(spec)
        e.step(spec.steps)
        u, _ = e.fields(cid)
        fields[st] = (u.double().cpu().numpy(), e.particle_rows(cid)[:, 0:3].cpu().numpy())
        del e
        torch.cuda.empty_cache()
    u64, c64 = fields["f64"]
    us, cs = fields[storage]
    u32, c32 = fields["f32"]
    scale = float(np.abs(u64).max())
    return {
        "sample": f"request {index} alone: {spec.nodes[0]}^3 D3Q{spec.q}, {len(spec.particles)} particles "
                  f"({', '.join(p.kind for p in spec.particles)}), {spec.steps} steps",
        "max_abs_u_f64": scale,
        "u_rel_linf_vs_f64": float(np.abs(us - u64).max() / scale) if scale > 0 else None,
        "centroid_abs_vs_f64": float(np.abs(cs - c64).max()) if len(c64) else None,
        # float32 storage on the same request: at settling velocities (~1e-5)
        # float32 rounding of the populations sets its error
        "f32_u_rel_linf_vs_f64": float(np.abs(u32 - u64).max() / scale) if scale > 0 else None,
        "f32_centroid_abs_vs_f64": float(np.abs(c32 - c64).max()) if len(c64) else None,
    }


def serve_one_at_a_time(seed=0, steps=(1000, 20000), storage="f16", per_lattice=8, timed_steps=200,
                        rate=50.0, duration=4.0):
    """Plain one-case-at-a-time solves of the serving mix: requests of the
    stream at `rate` for `duration` s, `per_lattice` per lattice at evenly
    spaced size quantiles (smallest and largest included), each solved alone
    with LBMD3 (eager steps, same storage); a request's time is its measured
    step time times its step count, and the mix's mean weights each lattice's
    sample by its share of the stream."""
    import torch

    from funlbm_amd.batch.workloads import lbm_config_for, serving_requests
    from funlbm_amd.lbm import LBMD3, Config

    specs = [s for _, s in serving_requests(rate, duration, seed=seed, steps=steps)]
    mean_s = 0.0
    picked = []
    for q in sorted({s.q for s in specs}):
        ids = sorted((i for i, s in enumerate(specs) if s.q == q), key=lambda i: specs[i].n)
        k = min(per_lattice, len(ids))
        picks = sorted({ids[round(j * (len(ids) - 1) / max(k - 1, 1))] for j in range(k)})
        times = []
        for i in picks:
            spec = specs[i]
            lbm = LBMD3(config=Config().from_json(lbm_config_for(spec, storage)))
            lbm.init()
            for n in range(1, 11):
                lbm.step = n
                lbm.run_step()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for n in range(11, 11 + timed_steps):
                lbm.step = n
                lbm.run_step()
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) / timed_steps * spec.steps)
            picked.append(spec)
            del lbm
            torch.cuda.empty_cache()
        mean_s += len(ids) / len(specs) * float(np.mean(times))
    sizes = [s.nodes[0] for s in picked]
    return {
        "solves_per_s": round(1.0 / mean_s, 3),
        "mean_solve_s": round(mean_s, 3),
        "sample": f"{len(picked)} requests of the {rate:g} req/s stream ({per_lattice} per lattice, sides "
                  f"{min(sizes)}-{max(sizes)}, evenly spaced size quantiles), each alone with LBMD3 "
                  f"for {timed_steps} timed steps, scaled to its step count, weighted by lattice",
    }


def bench_serve(rates, duration, steps=(1000, 20000), budget=200e9, storage="f16", tick=50, drain_s=6.0,
                max_active_nodes=1 << 26, baseline_per_lattice=8, error_check=True, engines=3):
    sweep = [serve_rate(r, duration, steps=steps, budget=budget, storage=storage, tick=tick, drain_s=drain_s,
                        max_active_nodes=max_active_nodes, engines=engines) for r in rates]
    base = None
    if baseline_per_lattice > 0:
        base = serve_one_at_a_time(steps=steps, storage=storage, per_lattice=baseline_per_lattice)
        best = max(r["solves_per_s"] for r in sweep)
        base["speedup_of_best_rate"] = round(best / base["solves_per_s"], 2)
        # the like-for-like figure: both sides as lattice updates per second
        # over the stream's mean request
        best_steady = max(r["solves_per_s_steady"] for r in sweep)
        base["speedup_steady_of_best_rate"] = round(best_steady / base["solves_per_s"], 2)
    return {
        "workload": f"Poisson arrivals for {duration} s per rate (run cut {drain_s} s after the last arrival), "
                    f"sides log-uniform 32-128, D3Q19/27, 1-4 spheres or ellipsoids, steps {steps[0]}-{steps[1]}, "
                    f"{storage}-deviation storage, {budget / 1e9:.0f} GB budget, at most {max_active_nodes} "
                    f"nodes stepping at once (first come, first served), ticks of {tick} steps, "
                    f"two ticks in flight, {engines} engines on streams of their own",
        "sweep": sweep,
        "one_at_a_time": base,
        "error_check": serve_error_check(steps=steps, storage=storage) if error_check else None,
        # a D3Q27 request of the same stream with four particles
        "error_check_d3q27": serve_error_check(steps=steps, storage=storage, index=5) if error_check else None,
    }


def bench_suspension_precision(size=256, n_particles=64, steps=2000):
    """The dense-suspension precision study at its configuration: D3Q27
    periodic box, 64 prolate ellipsoids (4, 4, 8), density ratio 1.2, body
    force 1e-6 along x, tau 0.55; float32, f16 and bf16 storage (both
    roundings) against float64 after `steps` steps.  Also with u not zeroed
    on the (periodic) faces, which the reference does after every moment
    update."""
    from funlbm_amd import precision

    out = {}
    for name, zero in (("reference_faces", None), ("no_face_zeroing", [])):
        cfg = precision.suspension(n=size, n_particles=n_particles, mu=0.05 / 3)
        if zero is not None:
            cfg["flow"]["uzero_faces"] = zero
        t0 = time.perf_counter()
        res = precision.compare(cfg, steps)
        res["seconds"] = round(time.perf_counter() - t0, 1)
        out[name] = res
    out["workload"] = (f"D3Q27 {size}^3 periodic, {n_particles} ellipsoids (4,4,8), rho_p 1.2, body force 1e-6 x, "
                       f"tau 0.55, {steps} steps; u rel L-inf and centroid error vs float64")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--q", type=int, default=19)
    ap.add_argument("--storage", default="f32")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-copy", action="store_true")
    ap.add_argument("--batch-cases", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=3000)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-baseline-per-lattice", type=int, default=8,
                    help="one-at-a-time sample cases per lattice (0: no baseline)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="seconds of arrivals per rate")
    ap.add_argument("--serve-steps", default="1000,20000", help="min,max steps of a request")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-baseline-per-lattice", type=int, default=8,
                    help="one-at-a-time sample requests per lattice of the serving leg (0: no baseline)")
    ap.add_argument("--no-serve-error-check", action="store_true",
                    help="skip the served storage's float64 comparison")
    ap.add_argument("--serve-engines", type=int, default=3,
                    help="batch engines of the server, each on a stream of its own")
    ap.add_argument("--no-f16", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the bf16, D3Q27 and suspension legs")
    ap.add_argument("--serve-rates", default="10,20,50,100,200",
                    help="comma-separated arrival rates (req/s) of the serving sweep")
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the D3Q27 suspension precision study")
    ap.add_argument("--precision-steps", type=int, default=2000)
    args = ap.parse_args()
    t_start = time.perf_counter()

    def done(leg):
        # one line per finished leg on stderr, so a long run shows progress
        print(f"[bench] {leg} done at {time.perf_counter() - t_start:.1f} s", file=sys.stderr, flush=True)

    if args.gpus != 1:
        print(
            "bench.py measures one GPU: a lattice runs on one GPU and there is no multi-GPU "
            "path; run with --gpus 1",
            file=sys.stderr,
        )
        sys.exit(2)

    import torch

    lbm = build_solver(args.size, args.q, args.storage)
    ms, wall = time_steps(lbm, args.steps, args.warmup)
    n = lbm.flow.n
    mlups = n / (ms * 1e-3) / 1e6
    elem = {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}[args.storage]
    bytes_per_lu = 2 * args.q * elem
    # roofline of the dominant kernel: algorithmic bytes of one launch (read
    # and write every population of every node once) / its average duration
    kernel_ms = time_fused_kernel(lbm)
    achieved = n * bytes_per_lu / (kernel_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "f32", "f16": "f16-deviation storage, f32 arithmetic",
                  "bf16": "bf16-deviation storage, f32 arithmetic", "f64": "f64"}[args.storage],
        "data": "synthetic (seeded random rho, u equilibrium)",
        "config": {
            "workload": f"D3Q{args.q} BGK fused step, {args.size}^3 nodes, all faces periodic, tau 0.6, no particles",
            "nodes": n,
            "storage": args.storage,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": profiled_traffic(args.storage, args.size, args.q),
            "traffic_source": "profiles/k_fused_d3q19_256_traffic.md (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1)",
            "bytes_per_lu": bytes_per_lu,
            "bytes_per_launch": n * bytes_per_lu,
            "kernel_ms": round(kernel_ms, 4),
            "kernel": "k_fused_step (funlbm_fused_step)",
        },
    }
    del lbm
    torch.cuda.empty_cache()
    done("headline")
    if not args.no_f16 and args.storage == "f32":
        lbm16 = build_solver(args.size, args.q, "f16")
        # a fresh 16-bit state settles over its first few dozen launches
        # (first touch of the new buffers); warm up past that
        ms16, _ = time_steps(lbm16, args.steps, max(args.warmup, 50))
        k16 = time_fused_kernel(lbm16)
        del lbm16
        torch.cuda.empty_cache()
        m16 = n / (ms16 * 1e-3) / 1e6
        out["f16_deviation"] = {
            "mlups": round(m16, 1),
            "ms_per_step": round(ms16, 4),
            "speedup_vs_f32": round(m16 / mlups, 3),
            "kernel_ms": round(k16, 4),
            "achieved_gbs": round(n * 2 * args.q * 2 / (k16 * 1e-3) / 1e9, 1),
        }
        out["f16_deviation"]["rounding"] = "stochastic"
        out["f16_nearest"] = storage_leg(args.size, args.q, "f16", args.steps, args.warmup, mlups,
                                         rounding="nearest")
        done("f16")
    if not args.no_extra and args.storage == "f32":
        try:
            out["bf16_deviation"] = storage_leg(args.size, args.q, "bf16", args.steps, args.warmup, mlups)
            q27 = storage_leg(args.size, 27, "f32", args.steps, args.warmup)
            out["bf16_nearest"] = storage_leg(args.size, args.q, "bf16", args.steps, args.warmup, mlups,
                                              rounding="nearest")
            out["d3q27"] = {"f32": q27,
                            "f16": storage_leg(args.size, 27, "f16", args.steps, args.warmup, q27["mlups"]),
                            "f16_nearest": storage_leg(args.size, 27, "f16", args.steps, args.warmup,
                                                       q27["mlups"], rounding="nearest")}
            out["suspension"] = bench_suspension(size=args.size)
            out["sphere_free_fall"] = bench_sphere()
        except Exception as e:  # reported, never silent
            out["extra_error"] = repr(e)
        torch.cuda.empty_cache()
        done("bf16, D3Q27, suspension and free-fall")
    if not args.no_batch:
        try:
            out["batched"] = bench_batch(n_cases=args.batch_cases, steps=args.batch_steps,
                                         baseline_per_lattice=args.batch_baseline_per_lattice)
        except Exception as e:  # reported, never silent
            out["batched"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        done("batched")
    if not args.no_serve:
        lo, hi = (int(v) for v in args.serve_steps.split(","))
        rates = [float(v) for v in args.serve_rates.split(",") if v]
        try:
            out["serving"] = bench_serve(rates, args.serve_seconds, steps=(lo, hi),
                                         baseline_per_lattice=args.serve_baseline_per_lattice,
                                         error_check=not args.no_serve_error_check,
                                         engines=args.serve_engines)
        except Exception as e:  # reported, never silent
            out["serving"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        done("serving")
    if not args.no_precision:
        try:
            out["suspension_precision"] = bench_suspension_precision(size=args.size,
                                                                     steps=args.precision_steps)
        except Exception as e:  # reported, never silent
            out["suspension_precision"] = {"error": repr(e)}
        torch.cuda.empty_cache()
        done("suspension precision")
    if not args.no_copy:
        cb = copy_bandwidth()
        out["roofline"]["copy_peak_gbs"] = round(cb, 1)
        out["roofline"]["frac_of_copy"] = round(achieved / cb, 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

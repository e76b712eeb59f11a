"""Benchmark of the fused D3Q19 BGK timestep on one MI355X.

Workload (the headline metric's configuration): D3Q19, float32 populations,
256 x 256 x 256 nodes, all six faces PERIODICAL, tau = 0.6, no particles,
initial state f = feq(rho, u) with rho = 1 + 1e-3 N(0,1), u = 1e-3 N(0,1)^3
(seed 0).  A step is one LBMD3.run_step: one fused pull-stream + moments +
collision pass over the lattice.  Inputs are resident in HBM before timing.

Prints one JSON line with MLUPS, the HBM roofline of the fused kernel (HIP
events on the launch stream; algorithmic bytes 2*Q*sizeof(storage) per node)
and a CPU baseline (the numpy oracle on a bounded sample).
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


def build_solver(n, q, storage, tau=0.6, seed=0):
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


def time_steps(lbm, steps, warmup):
    import torch

    for _ in range(warmup):
        lbm.run_step()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    t0 = time.perf_counter()
    for _ in range(steps):
        lbm.run_step()
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    return ms / steps, wall


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


def cpu_baseline(n=64, q=19, budget_s=15.0):
    """The numpy oracle's fluid step on an n^3 periodic lattice (bounded)."""
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
    mlups = n ** 3 * steps / el / 1e6
    return {
        "value": round(mlups, 3),
        "unit": "MLUPS",
        "cores": 1,
        "kind": "port",
        "sample": f"numpy oracle fluid step, D3Q{q} float32, {n}^3 periodic, {steps} steps in {el:.1f} s",
    }


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
    args = ap.parse_args()
    if args.gpus != 1:
        print(
            "bench.py measures one GPU: this build runs one lattice per GPU and no multi-GPU path "
            "(see DESIGN.md for how the lattice would shard); run with --gpus 1",
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
    achieved = mlups * 1e6 * bytes_per_lu / 1e9
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
            "traffic": None,
            "bytes_per_lu": bytes_per_lu,
            "kernel": "k_fused (funlbm_fused_step)",
        },
    }
    del lbm
    torch.cuda.empty_cache()
    if not args.no_copy:
        cb = copy_bandwidth()
        out["roofline"]["copy_peak_gbs"] = round(cb, 1)
        out["roofline"]["frac_of_copy"] = round(achieved / cb, 4)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

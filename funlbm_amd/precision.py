"""Accuracy of the population storage formats against a float64 run.

Each study runs one configuration through LBMD3 once per storage format
(float64 populations and arithmetic as the reference solution, then float32,
f16 and bf16 deviation storage with float32 arithmetic, the 16-bit ones with
round-to-nearest stores and, as "f16-sr" / "bf16-sr", stochastic ones) and
reports, per format, the relative L-inf and L2 errors of the velocity field
and the largest particle-centroid distance from the float64 run.
"""

import math
from typing import Dict, List

import numpy as np
import torch

from .lbm import LBMD3, Config

FACES = ("input", "output", "back", "forward", "bottom", "top")


def _run(cfg_json: dict, steps: int):
    lbm = LBMD3(config=Config().from_json(cfg_json))
    lbm.init()
    for s in range(1, steps + 1):
        lbm.step = s
        lbm.run_step()
    u = lbm.flow.u.double().cpu().numpy()
    cx = [p.cx.cpu().numpy().copy() for p in lbm.particle_swarm.particles]
    del lbm
    torch.cuda.empty_cache()
    return u, cx


def compare(cfg_json: dict, steps: int, storages=("f32", "f16", "bf16", "f16-sr", "bf16-sr")) -> Dict[str, dict]:
    """Runs cfg_json with float64 storage and with each of `storages` (a
    "-sr" suffix selects stochastic rounding of the 16-bit state)."""
    import copy

    ref = copy.deepcopy(cfg_json)
    ref["flow"]["storage"] = "f64"
    u64, cx64 = _run(ref, steps)
    scale = float(np.max(np.abs(u64)))
    norm = float(np.linalg.norm(u64))
    out = {"steps": steps, "max_abs_u_f64": scale}
    for st in storages:
        c = copy.deepcopy(cfg_json)
        name, _, mode = st.partition("-")
        c["flow"]["storage"] = name
        c["flow"]["rounding"] = "stochastic" if mode == "sr" else "nearest"
        u, cx = _run(c, steps)
        err = float(np.max(np.abs(u - u64)) / scale) if scale > 0 else float(np.max(np.abs(u - u64)))
        cxe = max((float(np.max(np.abs(a - b))) for a, b in zip(cx, cx64)), default=0.0)
        l2 = float(np.linalg.norm(u - u64))
        out[st] = {"u_rel_linf": err, "u_rel_l2": l2 / norm if norm > 0 else l2, "centroid_abs": cxe}
    return out


def poiseuille(size=(31, 31, 63), mu=0.1, g=1e-5) -> dict:
    """Poiseuille channel: walls at x, periodic y and z, body force g_z, u
    zeroed on the wall faces only (the reference zeroes it on all six faces,
    which on the periodic faces acts as a momentum sink)."""
    b = {f: {"code": "PERIODICAL"} for f in FACES}
    b["input"] = b["output"] = {"code": "WALL"}
    return {"flow": {"param_type": "D3Q19", "size": list(size), "param": {"mu": mu}, "boundary": b,
                     "body_force": [0.0, 0.0, g], "uzero_faces": ["input", "output"]}, "log_every": 0}


def cavity(size=(47, 47, 47), mu=0.05, lid=0.05) -> dict:
    """Lid-driven cavity: walls on every face, the top one moving along x."""
    b = {f: {"code": "WALL"} for f in FACES}
    b["top"] = {"code": "WALL_WITH_SPEED", "uw": [lid, 0.0, 0.0]}
    return {"flow": {"param_type": "D3Q19", "size": list(size), "param": {"mu": mu}, "boundary": b},
            "log_every": 0}


def settling_sphere(size=(63, 63, 127), mu=1 / 30, r=6, center=(32, 32, 100), rou=1.5, g=2e-5) -> dict:
    """A sphere settling in a walled box (gravity g along the box's z; the
    reference's buoyancy term (1 - rho_p/rho_f) V g moves a heavy sphere
    against g, so g = +z settles it towards z = 0)."""
    b = {f: {"code": "WALL"} for f in FACES}
    return {"flow": {"param_type": "D3Q19", "size": list(size), "param": {"mu": mu, "gravity": [0, 0, g]},
                     "boundary": b},
            "particles": [{"type": "sphere", "r": r, "rou": rou, "coord": {"center": list(center)}}],
            "log_every": 0}


def suspension(n=64, n_particles=4, seed=0, mu=0.05 / 3, abc=(4, 4, 8), rou=1.2, force=1e-6) -> dict:
    """Periodic D3Q27 box with prolate ellipsoids at seeded non-overlapping
    positions and uniformly random orientations, body force along x."""
    from .batch.workloads import _uniform_quat_rotvec

    rng = np.random.default_rng(seed)
    b = {f: {"code": "PERIODICAL"} for f in FACES}
    placed: List[np.ndarray] = []
    parts = []
    lo, hi = abc[2] + 3, n - abc[2] - 3
    tries = 0
    while len(parts) < n_particles and tries < 10000:
        tries += 1
        c = rng.uniform(lo, hi, size=3)
        if all(np.linalg.norm(c - q) > 2 * abc[2] + 2 for q in placed):
            placed.append(c)
            a = _uniform_quat_rotvec(rng)
            parts.append({"type": "ellipsoid", "a": abc[0], "b": abc[1], "c": abc[2], "rou": rou,
                          "coord": {"center": [float(v) for v in c], "alpha": a[0] or 1e-9,
                                    "beta": a[1] or 1e-9, "gamma": a[2] or 1e-9}})
    return {"flow": {"param_type": "D3Q27", "size": [n - 1] * 3, "param": {"mu": mu}, "boundary": b,
                     "body_force": [force, 0.0, 0.0]},
            "particles": parts, "log_every": 0}

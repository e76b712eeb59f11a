"""Seeded generators of the batched and served settling workloads."""

import math

import numpy as np

from .engine import CaseSpec, ParticleSpec


def _uniform_quat_rotvec(rng):
    """Rotation vector of a uniformly random rotation (Shoemake's quaternion)."""
    u1, u2, u3 = rng.random(3)
    q = np.array([
        math.sqrt(1 - u1) * math.sin(2 * math.pi * u2),
        math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
        math.sqrt(u1) * math.sin(2 * math.pi * u3),
        math.sqrt(u1) * math.cos(2 * math.pi * u3),
    ])  # (x, y, z, w)
    if q[3] < 0:
        q = -q
    ang = 2 * math.acos(min(1.0, q[3]))
    s = math.sqrt(max(1e-300, 1 - q[3] * q[3]))
    axis = q[:3] / s if s > 1e-12 else np.array([1.0, 0.0, 0.0])
    return tuple(float(v) for v in axis * ang)


def settling_batch(n_cases=256, seed=0, steps=3000, side=(32, 96), qs=(15, 19), tau=(0.55, 1.0),
                   radius=(3.0, 8.0), rho=(1.1, 3.0), gravity=2e-5):
    """Independent settling cases: cubic-ish boxes with each side
    log-uniform in `side`, lattice by coin flip from `qs`, tau uniform, one
    sphere of uniform radius and density ratio started in the upper middle of
    the box, walls on all faces, settling towards z = 0 (g = +z in the
    reference's buoyancy term (1 - rho_p/rho_f) V g)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_cases):
        nodes = tuple(int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1]))))) for _ in range(3))
        q = int(qs[rng.integers(len(qs))])
        t = float(rng.uniform(*tau))
        r = float(rng.uniform(*radius))
        r = min(r, (min(nodes) - 8) / 2.0)
        dens = float(rng.uniform(*rho))
        center = (nodes[0] / 2.0, nodes[1] / 2.0, max(nodes[2] * 0.7, r + 4.0))
        center = (center[0], center[1], min(center[2], nodes[2] - r - 4.0))
        out.append(CaseSpec(nodes=nodes, q=q, tau=t, steps=steps,
                            particles=[ParticleSpec(kind="sphere", r=r, rou=dens, center=center)],
                            gravity=(0.0, 0.0, gravity)))
    return out


def serving_requests(rate, duration, seed=0, side=(32, 128), qs=(19, 27), n_part=(1, 4),
                     steps=(1000, 20000), rho=(1.1, 3.0), radius=(3.0, 8.0), tau=(0.55, 1.0),
                     gravity=2e-5):
    """A Poisson stream of settling requests at `rate` per second over
    `duration` seconds: side log-uniform, lattice by coin flip, 1-4 spheres
    or ellipsoids with seeded sizes, orientations and density ratios, steps
    log-uniform.  Returns [(arrival_time, CaseSpec)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        L = int(round(math.exp(rng.uniform(math.log(side[0]), math.log(side[1])))))
        nodes = (L, L, L)
        q = int(qs[rng.integers(len(qs))])
        k = int(rng.integers(n_part[0], n_part[1] + 1))
        parts = []
        # particles on a coarse grid of slots in the upper half, non-overlapping
        rmax = min(radius[1], (L - 10) / 4.0)
        rmin = min(radius[0], rmax)
        slots = [(0.3, 0.3), (0.7, 0.3), (0.3, 0.7), (0.7, 0.7)]
        for j in range(k):
            r = float(rng.uniform(rmin, rmax))
            cx, cy = slots[j]
            center = (cx * L, cy * L, L * 0.6)
            if rng.random() < 0.5:
                parts.append(ParticleSpec(kind="sphere", r=r, rou=float(rng.uniform(*rho)), center=center))
            else:
                abc = (r * 0.6, r * 0.6, r)
                parts.append(ParticleSpec(kind="ellipsoid", abc=abc, rou=float(rng.uniform(*rho)), center=center,
                                          angle=_uniform_quat_rotvec(rng)))
        n_steps = int(round(math.exp(rng.uniform(math.log(steps[0]), math.log(steps[1])))))
        out.append((t, CaseSpec(nodes=nodes, q=q, tau=float(rng.uniform(*tau)), steps=n_steps,
                                particles=parts, gravity=(0.0, 0.0, gravity))))
    return out


FACES = ("input", "output", "back", "forward", "bottom", "top")


def lbm_config_for(spec, storage="f16", rounding="stochastic") -> dict:
    """The LBMD3 configuration (the reference's JSON schema) of one case:
    the same lattice, walls, relaxation time, gravity and particles, for
    solving it alone.  A rotation-vector component alpha of 0 is passed as
    1e-9, since the schema reads a zero alpha as its default pi/2."""
    from .. import _native as nat

    codes = {nat.BC_WALL: "WALL", nat.BC_PERIODICAL: "PERIODICAL"}
    parts = []
    for p in spec.particles:
        coord = {"center": [float(v) for v in p.center], "alpha": float(p.angle[0]) or 1e-9,
                 "beta": float(p.angle[1]), "gamma": float(p.angle[2])}
        d = {"rou": float(p.rou), "coord": coord, "markers": p.markers}
        if p.kind == "sphere":
            d.update(type="sphere", r=float(p.r))
        else:
            d.update(type="ellipsoid", a=float(p.abc[0]), b=float(p.abc[1]), c=float(p.abc[2]))
        parts.append(d)
    return {
        "flow": {"param_type": f"D3Q{spec.q}", "size": [int(n) - 1 for n in spec.nodes],
                 "param": {"mu": (float(spec.tau) - 0.5) / 3, "gravity": [float(g) for g in spec.gravity]},
                 "boundary": {f: {"code": codes[int(spec.bcs[i])]} for i, f in enumerate(FACES)},
                 "storage": storage, "rounding": rounding, "seed": int(spec.seed), "ibm_delta": spec.delta,
                 "uzero_faces": [f for i, f in enumerate(FACES) if (spec.uzero_mask >> i) & 1]},
        "particles": parts,
        "log_every": 0,
    }

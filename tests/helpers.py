"""Shared helpers of the test suite (no product logic)."""

import numpy as np

BC_NAMES = {11000: "PERIODICAL", 1200: "WALL", 1201: "WALL_WITH_SPEED", 1400: "NON_EQUILIBRIUM"}
FACES = ("input", "output", "back", "forward", "bottom", "top")


def flow_json(q, size, mu, bcs, uw=None, **ext):
    """A flow JSON block in the reference schema."""
    boundary = {}
    for i, face in enumerate(FACES):
        b = {"code": BC_NAMES[bcs[i]]}
        if uw is not None and uw[i] is not None:
            b["uw"] = list(uw[i])
        boundary[face] = b
    d = {"param_type": f"D3Q{q}", "size": list(size), "param": {"mu": mu}, "boundary": boundary}
    d.update(ext)
    return d


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def random_state(q, nodes, seed, amp=1e-2):
    """Random but physical populations: equilibrium of a perturbed rho/u plus noise."""
    from oracle import lbm_oracle as orc

    rng = np.random.default_rng(seed)
    lat = orc.Lattice(q, np.float32)
    rho = (1.0 + amp * rng.standard_normal(nodes + (1,))).astype(np.float32)
    u = (amp * rng.standard_normal(nodes + (3,))).astype(np.float32)
    f = orc.equilibrium(lat, rho, u)
    f = f + (amp * 0.1 * rng.standard_normal(f.shape) * lat.w).astype(np.float32)
    return f.astype(np.float32)


def deviation_round(f, w, storage, scale=1024.0, rho0=1.0):
    """The state storage's round trip of float32 populations f[..., q]: f16
    and bf16 keep (f - w rho0) * scale rounded to nearest-even, and read back
    as v / scale + w rho0, all in float32 (the kernels' StoreF16/StoreBF16)."""
    f = np.asarray(f, dtype=np.float32)
    wr = (np.asarray(w, dtype=np.float32) * np.float32(rho0)).astype(np.float32)
    s = np.float32(scale)
    inv = np.float32(1.0) / s
    d = ((f - wr) * s).astype(np.float32)
    if storage == "f16":
        r = d.astype(np.float16).astype(np.float32)
    elif storage == "bf16":
        u = d.view(np.uint32).astype(np.uint64)
        u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
        r = u.astype(np.uint32).view(np.float32)
    else:
        raise ValueError(storage)
    return (r * inv + wr).astype(np.float32)

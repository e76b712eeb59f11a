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


# ---------------------------------------------------------------- 16-bit state
# Restatements, for the tests, of the fused state's tile layout and of the
# stochastic rounding's dither and conversion (lbm_common.h).
_M32 = 0xFFFFFFFF


def mix32(x):
    """The 32-bit hash of the dither (x ^= x >> 16; x *= 0x7feb352d;
    x ^= x >> 15; x *= 0x846ca68b; x ^= x >> 16) on uint64 arrays."""
    x = np.asarray(x, dtype=np.uint64) & np.uint64(_M32)
    x = x ^ (x >> np.uint64(16))
    x = (x * np.uint64(0x7FEB352D)) & np.uint64(_M32)
    x = x ^ (x >> np.uint64(15))
    x = (x * np.uint64(0x846CA68B)) & np.uint64(_M32)
    x = x ^ (x >> np.uint64(16))
    return x


DITHER_MULT = (0x8990C1, 0x932181, 0xF1E9DD, 0xFE5ED1, 0xC5D53F, 0xE3D3BB, 0x8BA1E7, 0xD35FBB, 0x9FD5AB,
               0x8BAA7D, 0xA8F9FB, 0xDC746D, 0xBFCFF1, 0xF63447, 0xBD4A6F, 0xA56AE5, 0xABDD59, 0xBB2F91,
               0x8001ED, 0xDD7671, 0xBE2E8F, 0xDBCAC3, 0x9006FB, 0xCE98A9, 0xFF9FE3, 0xBC09CB, 0xDC5B09)


def dither_mult(k):
    return DITHER_MULT[k]


def step_seed(base, steps_done):
    return (int(base) + int(steps_done) * 0x9E3779B9) & _M32


def dither_words(n, q, seed):
    """uint64 [q, n]: the 32-bit dither word of direction k at node j in a step
    of seed `seed`: ((h mod 2^24) * C_k) mod 2^32, h = mix32(j ^ mix32(seed))."""
    h = mix32(np.arange(n, dtype=np.uint64) ^ mix32(np.uint64(seed)))
    h24 = h & np.uint64(0xFFFFFF)
    return np.stack([(h24 * np.uint64(dither_mult(k))) & np.uint64(_M32) for k in range(q)])


def sr_round16(d, r, storage):
    """Stochastic rounding of float32 d to f16 / bf16 bit patterns (uint16):
    truncate toward zero at the format's precision, then add one ulp iff the
    discarded fraction, truncated to 13 (f16) or 16 (bf16) bits, plus the top
    13 / 16 bits of the 32-bit word r reach one ulp."""
    d = np.asarray(d, dtype=np.float32)
    x = np.abs(d.astype(np.float64))
    bits = 13 if storage == "f16" else 16
    mant = 10 if storage == "f16" else 7
    _, e2 = np.frexp(x)
    e = e2.astype(np.int64) - 1  # x in [2^e, 2^(e+1))
    if storage == "f16":
        e = np.maximum(e, -14)  # subnormal f16: ulp 2^-24
    ulp_e = e - mant
    qv = np.ldexp(x, -ulp_e)
    fl = np.floor(qv)
    rem = np.floor((qv - fl) * float(1 << bits))
    rh = (np.asarray(r, dtype=np.uint64) >> np.uint64(32 - bits)).astype(np.float64)
    up = (rem + rh) >= float(1 << bits)
    mag = np.ldexp(fl + up, ulp_e)
    val = np.copysign(mag, d.astype(np.float64))
    with np.errstate(over="ignore"):  # beyond the f16 range: infinity, as the hardware
        if storage == "f16":
            return val.astype(np.float16).view(np.uint16)
        return (val.astype(np.float32).view(np.uint32) >> np.uint32(16)).astype(np.uint16)


def rne_round16(d, storage):
    """Round-to-nearest-even of float32 d to f16 / bf16 bit patterns."""
    d = np.asarray(d, dtype=np.float32)
    if storage == "f16":
        return d.astype(np.float16).view(np.uint16)
    u = d.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def state_tile_shift(n):
    return 12 if n >= (1 << 14) else 8


def state_elems(q, n):
    T = 1 << state_tile_shift(n)
    return (-(-n // T) + 2) * T * q


def state_index(q, n):
    """int64 [q, n]: element of (direction k, node j) in a fused-state buffer."""
    ts = state_tile_shift(n)
    T = 1 << ts
    j = np.arange(n, dtype=np.int64)
    k = np.arange(q, dtype=np.int64)[:, None]
    return (q + (j >> ts) * q + k) * T + (j & (T - 1))


def deviation_round_sr(f, w, storage, seed, scale=1024.0):
    """The state storage's round trip with stochastic rounding of the step of
    seed `seed`: f [X, Y, Z, Q] float32, nodes in the kernels' order."""
    f = np.asarray(f, dtype=np.float32)
    q = f.shape[-1]
    n = int(np.prod(f.shape[:-1]))
    wr = np.asarray(w, dtype=np.float32)
    s = np.float32(scale)
    d = ((f - wr) * s).astype(np.float32).reshape(n, q).T
    bits = sr_round16(d, dither_words(n, q, seed), storage)
    if storage == "f16":
        r = bits.view(np.float16).astype(np.float32)
    else:
        r = (bits.astype(np.uint32) << np.uint32(16)).view(np.float32)
    return (r.T.reshape(f.shape) * (np.float32(1) / s) + wr).astype(np.float32)

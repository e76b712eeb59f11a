"""The 16-bit states' rounding at the fused step's store (needs a GPU).

One fused step through the C ABI from a chosen 16-bit state, against the
oracle's step in float64 from the same populations followed by the test
suite's model of the conversion (round to nearest even, or gfx950's
stochastic conversion fed by the dither words of the step's seed).  The
kernel's float32 deviation arithmetic differs from float64 by a few float32
ulps of the largest terms, so where the exact value lies within that much of
the rounding decision (the midpoint, the stochastic threshold or a 16-bit
value) either neighbour is right; every other population must be stored
exactly as the model rounds it.  With stochastic rounding about a quarter of
the populations round differently from nearest, so a wrong dither word,
direction or node index, or a store that rounds to nearest, fails."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_oracle as orc
from tests.helpers import (dither_words, random_state, rne_round16, sr_round16, state_elems, state_index,
                           step_seed)

pytestmark = pytest.mark.gpu

W, P = orc.WALL, orc.PERIODICAL
BCS = {"periodic": [P] * 6, "channel": [W, W, P, P, W, W]}
TORCH16 = {"f16": torch.float16, "bf16": torch.bfloat16}


def lattice(q, nodes, storage, bcs, tau, rounding, seed):
    from funlbm_amd import _native as nat

    lat = nat.Lattice()
    lat.q = q
    lat.nx, lat.ny, lat.nz = nodes
    lat.storage = nat.STORAGE_CODES[storage]
    for i in range(6):
        lat.bc[i] = bcs[i]
    lat.uzero_mask = 0x3F
    lat.tau = float(np.float32(tau))
    lat.rho0 = 1.0
    lat.dev_scale = 1024.0
    lat.rounding = nat.ROUNDING_CODES[rounding]
    lat.seed = seed
    return lat


def gpu_step(lat, bits_qn):
    """One funlbm_fused_step from the 16-bit patterns bits_qn [q, n]; returns
    the stored patterns [q, n]."""
    from funlbm_amd import _native as nat

    q, n = bits_qn.shape
    idx = state_index(q, n)
    src = np.zeros(state_elems(q, n), np.uint16)
    src[idx] = bits_qn
    dev = torch.device("cuda", 0)
    t_src = torch.from_numpy(src.view(np.int16)).to(dev)
    t_dst = torch.zeros_like(t_src)
    nbytes = ctypes.c_int64()
    nat.check(nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)), "state_bytes")
    assert nbytes.value == 2 * src.size
    nat.call("funlbm_fused_step", lat, nat.ptr(t_src), nat.ptr(t_dst), None, None, None, None, None, None,
             nat.stream_handle(dev))
    torch.cuda.synchronize()
    return t_dst.cpu().numpy().view(np.uint16)[idx]


def oracle_post_collision(q, nodes, bcs, tau, g_qn):
    """The fused step in float64 from f = g / 1024 + w (w the float32 weights
    the kernels use): stream, moments, collide; returns the float32 scaled
    deviations of the post-collision populations, [q, n]."""
    st = orc.init_flow(q, nodes, tau=float(np.float32(tau)), bc=bcs, dtype=np.float64)
    w32 = orc.Lattice(q, np.float32).w.astype(np.float64)
    st.lat.w = w32
    n = int(np.prod(nodes))
    st.f = (g_qn.T.reshape(nodes + (q,)) * 2.0 ** -10 + w32).astype(np.float64)
    orc.f_stream(st)
    orc.update_u_rou(st)
    orc.cul_equ(st)
    return ((st.f - w32) * 1024.0).astype(np.float32).reshape(n, q).T


def decision_distance(g, r, storage):
    """|distance| of float32 g from the nearest point where the model's
    result changes: a 16-bit value, or the midpoint (r None: nearest) or the
    stochastic threshold |a| + (1 - (r >> (32 - bits)) / 2^bits)(|b| - |a|)."""
    bits = 13 if storage == "f16" else 16
    lo = sr_round16(g, np.zeros(g.shape, np.uint64), storage)  # toward zero
    a = np.abs(decode16(lo, storage))
    b = np.abs(decode16(lo + np.uint16(1), storage))
    x = np.abs(g.astype(np.float64))
    if r is None:
        t = 0.5 * (a + b)
    else:
        rh = (np.asarray(r, np.uint64) >> np.uint64(32 - bits)).astype(np.float64)
        t = a + (1.0 - rh / (1 << bits)) * (b - a)
    return np.minimum(np.minimum(np.abs(x - t), x - a), b - x)


def decode16(bits, storage):
    bits = np.asarray(bits, np.uint16)
    if storage == "f16":
        return bits.view(np.float16).astype(np.float64)
    return (bits.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("bcname", ["periodic", "channel"])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("q", [19, 27])
def test_store_rounding_matches_model(q, storage, bcname):
    size = (11, 9, 13)
    nodes = tuple(s + 1 for s in size)
    n = int(np.prod(nodes))
    tau = 0.62
    bcs = BCS[bcname]
    f0 = random_state(q, nodes, seed=21)
    w32 = orc.Lattice(q, np.float32).w
    g0 = ((f0 - w32) * np.float32(1024.0)).astype(np.float32).reshape(n, q).T
    bits0 = rne_round16(g0, storage)
    g0 = torch.from_numpy(bits0.view(np.int16)).view(TORCH16[storage]).float().numpy()
    want_g = oracle_post_collision(q, nodes, bcs, tau, g0)
    seed = step_seed(12345, 7)
    words = dither_words(n, q, seed)
    rne = rne_round16(want_g, storage)
    sr = sr_round16(want_g, words, storage)
    # the two modes differ on a good part of the populations
    assert np.mean(sr != rne) > 0.15
    # bound on the float32 arithmetic error of the kernel relative to float64
    # (a few float32 ulps of the largest deviation; 2e-6 of it was exceeded
    # by 1 or 2 populations in 20000)
    delta = 8e-6 * float(np.max(np.abs(want_g)))
    for rounding, want, r in (("nearest", rne, None), ("stochastic", sr, words)):
        got = gpu_step(lattice(q, nodes, storage, bcs, tau, rounding, seed), bits0)
        far = decision_distance(want_g, r, storage) > delta
        bad = far & (got != want)
        near_diff = int(((got != want) & ~far).sum())
        print(storage, q, bcname, rounding, "far from a decision", int(far.sum()), "of", far.size,
              "mismatches there", int(bad.sum()), "near-decision differences", near_diff)
        # enough populations away from a decision that a wrong dither (a
        # quarter of them rounding the other way) cannot pass
        assert far.sum() > 4000
        assert not bad.any(), (rounding, int(bad.sum()))
        # near a decision the stored value is one of the two neighbours
        err = np.abs(decode16(got, storage) - want_g)
        step = np.abs(decode16(rne_round16(want_g, storage) + np.uint16(1), storage) - decode16(rne, storage))
        assert np.all(err <= step + delta)


def test_stochastic_dither_changes_with_the_seed():
    """Two steps from the same state with different seeds store different
    roundings (and equal seeds equal ones)."""
    q, storage = 19, "f16"
    nodes = (8, 8, 8)
    n = int(np.prod(nodes))
    f0 = random_state(q, nodes, seed=3)
    w32 = orc.Lattice(q, np.float32).w
    bits0 = rne_round16(((f0 - w32) * np.float32(1024.0)).astype(np.float32).reshape(n, q).T, storage)
    a = gpu_step(lattice(q, nodes, storage, [P] * 6, 0.7, "stochastic", 1), bits0)
    b = gpu_step(lattice(q, nodes, storage, [P] * 6, 0.7, "stochastic", 2), bits0)
    c = gpu_step(lattice(q, nodes, storage, [P] * 6, 0.7, "stochastic", 1), bits0)
    assert np.array_equal(a, c)
    assert np.mean(a != b) > 0.2

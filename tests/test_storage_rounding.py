"""The test suite's model of the f16/bf16 state storage, checked against
torch's own float32 -> half / bfloat16 conversions (round to nearest even)."""

import numpy as np
import torch

from oracle import lbm_oracle as orc
from tests.helpers import deviation_round


def test_deviation_round_matches_torch_conversions():
    rng = np.random.default_rng(0)
    lat = orc.Lattice(19, np.float32)
    f = (lat.w * (1 + 1e-2 * rng.standard_normal((64, 19)))).astype(np.float32)
    # ties and near-ties of the 16-bit grids as well
    f[0] = lat.w + np.float32(2.0 ** -21)
    f[1] = lat.w - np.float32(3 * 2.0 ** -22)
    wr = lat.w.astype(np.float32)
    s = np.float32(1024.0)
    d = torch.from_numpy(((f - wr) * s).astype(np.float32))
    for storage, dt in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        want = (d.to(dt).float().numpy() * (np.float32(1) / s) + wr).astype(np.float32)
        got = deviation_round(f, lat.w, storage, 1024.0)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), storage
    # the round trip moves values, so a checker that skipped it would differ
    assert not np.array_equal(deviation_round(f, lat.w, "bf16"), f)


# ------------------------------------------------------------ stochastic rounding
def test_stochastic_rounding_model_matches_recorded_hardware_vectors():
    """The model of gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 against the
    conversions recorded on an MI355X (tests/golden/sr_vectors.npz: crafted
    ties, subnormals, overflow and zeros, and random pairs)."""
    import os

    from tests.helpers import sr_round16

    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "sr_vectors.npz"))
    assert len(z["x"]) > 4000
    assert np.array_equal(sr_round16(z["x"], z["r"], "f16"), z["f16"])
    assert np.array_equal(sr_round16(z["x"], z["r"], "bf16"), z["bf16"])


def _neighbours(d, storage):
    """The two adjacent 16-bit values around float32 d (toward and away from zero)."""
    from tests.helpers import sr_round16

    lo = sr_round16(d, np.zeros(d.shape, np.uint64), storage)  # truncation
    return lo, lo + np.uint16(1)


def test_stochastic_rounding_is_unbiased_and_lands_on_a_neighbour():
    """Over all 2^13 (f16) or 2^16 (bf16) top-bit values of the random word,
    the mean of the rounded value is the input up to the truncation of the
    discarded fraction; every result is one of the two neighbours; exact
    values never move; the sign is symmetric."""
    from tests.helpers import rne_round16, sr_round16

    rng = np.random.default_rng(1)
    d = np.concatenate([
        (rng.standard_normal(40) * 10.0 ** rng.uniform(-6, 3, 40)).astype(np.float32),
        np.float32([1.0, 1.0 + 2.0 ** -11, 3e-6, 6e-8, 65519.0]),
    ])
    for storage, bits, dt in (("f16", 13, np.float16), ("bf16", 16, None)):
        words = (np.arange(1 << bits, dtype=np.uint64) << np.uint64(32 - bits))
        lo, hi = _neighbours(d, storage)
        for i, v in enumerate(d):
            out = sr_round16(np.full(len(words), v, np.float32), words, storage)
            assert set(np.unique(out)) <= {lo[i], hi[i]}
            vals = _as_float(out, storage)
            a, b = _as_float(np.array([lo[i], hi[i]], np.uint16), storage)
            if b == a or not np.isfinite(b):
                continue
            frac = (abs(float(v)) - abs(float(a))) / (abs(float(b)) - abs(float(a)))
            p_up = np.mean(out == hi[i])
            # floor(frac * 2^bits) / 2^bits of the draws round up
            assert abs(p_up - np.floor(frac * (1 << bits)) / (1 << bits)) < 1e-12, (storage, v)
            assert abs(np.mean(vals) - float(v)) <= (abs(float(b)) - abs(float(a))) * 2.0 ** -bits * 1.0001
        # negatives mirror positives
        r = rng.integers(0, 1 << 32, size=len(d), dtype=np.uint64)
        pos, neg = sr_round16(d, r, storage), sr_round16(-d, r, storage)
        assert np.array_equal(pos & np.uint16(0x7FFF), neg & np.uint16(0x7FFF))
        assert np.all((pos ^ neg) == np.uint16(0x8000))
        # values on the 16-bit grid stay put for every word
        grid = _as_float(rne_round16(d, storage), storage).astype(np.float32)
        for w in (0, 1 << 31, (1 << 32) - 1):
            assert np.array_equal(sr_round16(grid, np.full(len(d), w, np.uint64), storage),
                                  rne_round16(grid, storage))


def _as_float(bits16, storage):
    bits16 = np.asarray(bits16, dtype=np.uint16)
    if storage == "f16":
        return bits16.view(np.float16).astype(np.float64)
    return (bits16.astype(np.uint32) << np.uint32(16)).view(np.float32).astype(np.float64)


def test_dither_words_are_uniform_and_uncorrelated_across_directions():
    """The top 13 / 16 bits that the conversions read from the dither words
    (mix32 of node and step seed, times C_k mod 2^32) are uniform over nodes,
    nearly uncorrelated between directions and between consecutive steps, and
    change with the seed; the multipliers are odd and in [2^23, 2^24)."""
    from tests.helpers import DITHER_MULT

    assert len(set(DITHER_MULT)) == 27
    assert all(c & 1 and (1 << 23) <= c < (1 << 24) for c in DITHER_MULT)
    from tests.helpers import dither_words, step_seed

    n, q = 1 << 16, 27
    w0 = dither_words(n, q, step_seed(0, 0))
    w1 = dither_words(n, q, step_seed(0, 1))
    for bits in (13, 16):
        top = (w0 >> np.uint64(32 - bits)).astype(np.float64) / (1 << bits)
        assert np.all(np.abs(top.mean(axis=1) - 0.5) < 0.01)
        assert np.all(np.abs(top.var(axis=1) - 1 / 12) < 0.005)
        c = np.corrcoef(top)
        assert np.max(np.abs(c - np.eye(q))) < 0.03
        top1 = (w1 >> np.uint64(32 - bits)).astype(np.float64) / (1 << bits)
        assert np.max(np.abs([np.corrcoef(top[k], top1[k])[0, 1] for k in range(q)])) < 0.03
    assert np.mean(w0 == w1) < 1e-3

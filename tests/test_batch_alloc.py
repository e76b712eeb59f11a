"""Host logic of the batch pools (CPU)."""

import random

import pytest

from funlbm_amd.batch.alloc import RangeAllocator


def test_first_fit_and_coalescing():
    a = RangeAllocator(100)
    x = a.alloc(30)
    y = a.alloc(30)
    z = a.alloc(30)
    assert (x, y, z) == (0, 30, 60)
    assert a.alloc(20) is None and a.fits(10)
    a.free(y, 30)
    assert a.alloc(10) == 30  # first fit in the hole
    a.free(30, 10)
    a.free(x, 30)
    a.free(z, 30)
    assert a.used == 0 and a.alloc(100) == 0


def test_alignment_and_high_water():
    a = RangeAllocator(1024, align=256)
    assert a.alloc(1) == 0 and a.alloc(300) == 256
    assert a.high_water == 768
    a.free(256, 300)
    assert a.high_water == 256
    with pytest.raises(ValueError):
        a.free(256, 300)


def test_random_workload_never_overlaps():
    rng = random.Random(0)
    a = RangeAllocator(10_000, align=8)
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off = rng.choice(list(live))
            a.free(off, live.pop(off))
        else:
            size = rng.randint(1, 400)
            off = a.alloc(size)
            if off is not None:
                live[off] = size
        spans = sorted((o, o + ((s + 7) // 8) * 8) for o, s in live.items())
        assert all(spans[i][1] <= spans[i + 1][0] for i in range(len(spans) - 1))
        assert all(e <= 10_000 for _, e in spans)
    assert a.used == sum(((s + 7) // 8) * 8 for s in live.values())

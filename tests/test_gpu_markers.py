"""Parity of the GPU marker generator (funlbm_surface_markers) with the
oracle's closed-form generator (bit-exact float32 markers, same order) and
with its fsolve generator (the reference's method), for spheres and
ellipsoids, unit and fractional spacing, one shape per call and many."""

import numpy as np
import pytest
import torch

from oracle import ibm_oracle as ibo

pytestmark = pytest.mark.gpu

SHAPES = [
    ((3.0, 3.0, 3.0), 1.0),
    ((6.0, 6.0, 6.0), 1.0),
    ((4.0, 4.0, 8.0), 1.0),
    ((2.5, 3.5, 4.0), 1.0),
    ((8.0, 8.0, 8.0), 1.0),
    ((4.37, 4.37, 4.37), 1.0),
    ((3.07, 3.07, 3.07), 0.7),
    ((1.3, 2.9, 2.1), 0.45),
]


def raw_call(semi, dx, cap, fill=np.nan):
    from funlbm_amd import _native as nat

    dev = torch.device("cuda", 0)
    s = torch.tensor(semi, dtype=torch.float64, device=dev).reshape(-1, 3)
    out = torch.full((s.shape[0], cap, 3), fill, dtype=torch.float32, device=dev)
    counts = torch.full((s.shape[0],), -7, dtype=torch.int32, device=dev)
    nat.call("funlbm_surface_markers", s.shape[0], nat.ptr(s), float(dx), cap, nat.ptr(out), nat.ptr(counts),
             nat.stream_handle(dev))
    torch.cuda.synchronize()
    return out.cpu().numpy(), counts.cpu().numpy()


@pytest.mark.parametrize("abc,dx", SHAPES)
def test_markers_bitwise_equal_closed_form_oracle(abc, dx):
    want = ibo.surface_points_closed_form(*abc, dx=dx).astype(np.float32)
    out, cnt = raw_call([abc], dx, len(want) + 16)
    assert cnt[0] == len(want)
    got = out[0, : cnt[0]]
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # the rows past the count are untouched
    assert np.all(np.isnan(out[0, cnt[0]:]))


def test_many_shapes_in_one_launch_match_one_at_a_time():
    cap = 2048
    out, cnt = raw_call([s for s, dx in SHAPES if dx == 1.0], 1.0, cap)
    for i, (abc, dx) in enumerate([s for s in SHAPES if s[1] == 1.0]):
        want = ibo.surface_points_closed_form(*abc, dx=dx).astype(np.float32)
        assert cnt[i] == len(want)
        assert np.array_equal(out[i, : cnt[i]].view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("abc,dx", [((3.0, 3.0, 6.0), 1.0), ((3.07, 3.07, 3.07), 0.7)])
def test_markers_match_fsolve_generator(abc, dx):
    """Against the reference's method (fsolve from t = 0.5 per vertex pair):
    same count and order, positions within float32 rounding."""
    want = ibo.surface_points(*abc, dx=dx)
    out, cnt = raw_call([abc], dx, len(want) + 16)
    assert cnt[0] == len(want)
    np.testing.assert_allclose(out[0, : cnt[0]], want, rtol=0, atol=2e-6)


def test_capacity_short_writes_the_leading_markers_and_reports_the_count():
    want = ibo.surface_points_closed_form(6.0, 6.0, 6.0).astype(np.float32)
    cap = len(want) // 3
    out, cnt = raw_call([(6.0, 6.0, 6.0)], 1.0, cap)
    assert cnt[0] == len(want)
    assert np.array_equal(out[0].view(np.uint32), want[:cap].view(np.uint32))


def test_invalid_shapes_report_minus_one():
    out, cnt = raw_call([(3.0, 3.0, 3.0), (-1.0, 2.0, 2.0), (2.0, np.inf, 2.0), (600.0, 2.0, 2.0), (0.0, 1.0, 1.0)],
                        1.0, 512)
    want = ibo.surface_points_closed_form(3.0, 3.0, 3.0).astype(np.float32)
    assert cnt[0] == len(want)
    assert np.array_equal(out[0, : cnt[0]].view(np.uint32), want.view(np.uint32))
    assert list(cnt[1:]) == [-1, -1, -1, -1]
    assert np.all(np.isnan(out[1:]))


def test_wrapper_caches_and_grows_its_capacity(monkeypatch):
    from funlbm_amd.particle import markers

    shapes = [(5.5, 5.5, 5.5), (2.0, 3.0, 9.5), (5.5, 5.5, 5.5)]
    got = markers.surface_markers(shapes)
    assert got[0] is got[2]
    for g, abc in zip(got, shapes):
        want = ibo.surface_points_closed_form(*abc).astype(np.float32)
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    # a first capacity estimate that is too short: the second pass sizes the
    # output to the reported counts and the markers still come back whole
    monkeypatch.setattr(markers, "_area_estimate", lambda a, b, c: 1.0)
    new_shapes = [(4.75, 4.75, 4.75), (2.25, 3.25, 6.5)]
    got = markers.surface_markers(new_shapes)
    for g, abc in zip(got, new_shapes):
        want = ibo.surface_points_closed_form(*abc).astype(np.float32)
        assert len(want) > 67
        assert np.array_equal(g.view(np.uint32), want.view(np.uint32))


def test_particles_take_their_markers_from_the_generator():
    from funlbm_amd.particle.base import ParticleConfig, create_particle

    cfg = ParticleConfig().from_json({"type": "ellipsoid", "a": 3, "b": 4, "c": 5, "rou": 1.2})
    p = create_particle(cfg, device="cuda")
    p.init()
    want = ibo.surface_points_closed_form(3, 4, 5).astype(np.float32)
    assert np.array_equal(p._lagrange.view(np.uint32), want.view(np.uint32))
    assert p.lm_value == np.float32(np.float32(p.area) / np.float32(len(want)))

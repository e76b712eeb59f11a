"""Fibonacci-sphere markers (the non-default marker option), pinned by
closed forms: the count rule (one marker per dx^2 of area), every point on
the surface, equal-area z bands, nearest neighbours about dx apart."""

import math

import numpy as np
import pytest

from funlbm_amd.particle.markers import fibonacci_markers


@pytest.mark.parametrize("r,dx", [(6.0, 1.0), (3.0, 1.0), (8.0, 0.7), (4.5, 1.3)])
def test_sphere_count_surface_spacing(r, dx):
    m = fibonacci_markers(r, r, r, dx).astype(np.float64)
    n = round(4 * math.pi * r * r / (dx * dx))
    assert len(m) == n
    assert np.allclose(np.linalg.norm(m, axis=1), r, rtol=2e-7)
    # nearest-neighbour distances about dx (0.95 dx for the golden-angle lattice)
    d = np.linalg.norm(m[:, None, :] - m[None, :, :], axis=-1)
    np.fill_diagonal(d, np.inf)
    nn = d.min(axis=1)
    assert 0.9 * dx < float(np.median(nn)) < 1.1 * dx
    assert float(nn.min()) > 0.5 * dx
    # the centroid is near the centre and the z bands hold equal counts
    assert np.linalg.norm(m.mean(axis=0)) < 0.02 * r
    counts, _ = np.histogram(m[:, 2], bins=4, range=(-r, r))
    assert counts.max() - counts.min() <= 2


def test_first_points_and_golden_angle():
    m = fibonacci_markers(1.0, 1.0, 1.0, 0.25).astype(np.float64)
    n = len(m)
    z0 = 1 - 1 / n
    assert np.allclose(m[0], [math.sqrt(1 - z0 * z0), 0.0, z0], atol=1e-7)
    t1 = math.pi * (3 - math.sqrt(5))
    z1 = 1 - 3 / n
    r1 = math.sqrt(1 - z1 * z1)
    assert np.allclose(m[1], [r1 * math.cos(t1), r1 * math.sin(t1), z1], atol=1e-7)


def test_ellipsoid_points_lie_on_the_surface():
    a, b, c = 4.0, 4.0, 8.0
    m = fibonacci_markers(a, b, c).astype(np.float64)
    v = (m[:, 0] / a) ** 2 + (m[:, 1] / b) ** 2 + (m[:, 2] / c) ** 2
    assert np.allclose(v, 1.0, atol=1e-6)
    p = 1.6075
    area = 4 * math.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)
    assert len(m) == round(area)


def test_unknown_marker_method_is_refused():
    from funlbm_amd.particle.markers import markers

    with pytest.raises(ValueError):
        markers("spiral", 3, 3, 3)

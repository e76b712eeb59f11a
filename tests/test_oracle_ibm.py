"""Pins of the oracle's immersed-boundary and particle functions (CPU),
among them the closed-form marker generator against the fsolve generator."""

import math

import numpy as np
import pytest

from oracle import ibm_oracle as ibo
from oracle import lbm_oracle as orc


# ---------------------------------------------------------------- markers
@pytest.mark.parametrize("abc,dx", [((3, 3, 3), 1.0), ((4.5, 4.5, 4.5), 1.0), ((3, 3, 6), 1.0),
                                    ((2.5, 3.5, 4), 1.0), ((3.07, 3.07, 3.07), 0.7), ((1.3, 2.9, 2.1), 0.45)])
def test_closed_form_marker_generator_matches_fsolve_generator(abc, dx):
    """Closed-form in-segment roots vs scipy fsolve from t = 0.5 (the
    reference's method): same markers, same order."""
    got = ibo.surface_points_closed_form(*abc, dx=dx)
    want = ibo.surface_points(*abc, dx=dx)
    assert got.shape == want.shape
    assert np.max(np.abs(got - want)) < 1e-7


def test_markers_lie_near_the_surface_and_cover_it():
    for a, b, c in [(6, 6, 6), (4, 4, 8)]:
        m = ibo.surface_points_closed_form(a, b, c)
        val = (m[:, 0] / a) ** 2 + (m[:, 1] / b) ** 2 + (m[:, 2] / c) ** 2
        # centroids of the cut polygon sit within one cell of the surface
        assert np.all(np.abs(np.sqrt(val) - 1) < 1.0 / min(a, b, c))
        # about one marker per unit of area (cube-intersection sampling)
        area = 4 * math.pi * ((a * b) ** 1.6 + (a * c) ** 1.6 + (b * c) ** 1.6) ** (1 / 1.6) / 3 ** (1 / 1.6)
        assert 0.8 * area < len(m) < 2.0 * area
        # symmetric point set
        assert np.allclose(np.sort(m[:, 0]), np.sort(-m[:, 0]), atol=1e-9)


# ---------------------------------------------------------------- weights
def test_cosine_kernel_partition_of_unity():
    """(1 + cos(pi r/2))/4 over the 4 nodes around x sums to 1 per axis, so the
    unnormalised 3-D weights of an interior marker already sum to 1."""
    rng = np.random.default_rng(0)
    for _ in range(50):
        x = rng.uniform(3, 10, size=3).astype(np.float32)
        rl, rr = ibo.region_bounds(x[None], [20, 20, 20])
        g = ibo._grid(rl[0], rr[0]).astype(np.float32)
        r = g - x
        phi = (1 + np.cos(np.abs(r * np.float32(np.pi) / 2))) / 4
        w = np.prod(phi, axis=-1)
        assert g.shape[:3] == (4, 4, 4)
        assert abs(w.sum() - 1) < 1e-5
        wn = ibo.weight_function(r)
        assert abs(wn.sum() - 1) < 1e-6


def test_region_bounds_clamp_and_empty_stencil():
    lx = np.array([[0.3, 5.5, 18.9], [25.0, 1.0, 2.0]], dtype=np.float32)
    rl, rr = ibo.region_bounds(lx, [20, 20, 20])
    assert rl[0].tolist() == [0, 4, 17] and rr[0].tolist() == [4, 8, 20]
    # a marker beyond the upper bound gets an empty stencil: weight sum 0
    assert rl[1, 0] == 24 and rr[1, 0] == 20
    g = ibo._grid(rl[1], rr[1])
    assert g.size == 0


# ---------------------------------------------------------------- wall spring
def test_wall_spring_duplicate_index_last_wins():
    p = ibo.Particle(kind="sphere", r=3, pdtype=np.float64, lagrange=np.zeros((3, 3), np.float32))
    p.lx = np.array([[0.5, 0.5, 5.5], [5.0, 5.0, 5.0], [6.0, 6.0, 6.0]], dtype=np.float32)
    p.lF = np.zeros((3, 3), np.float32)
    ibo.particle_to_wall([p], [20, 20, 20], dx=1.0)
    d = 300 * (1 - 0.5 / 2)
    # marker 0 is argmin of x and y: rows j=0 (+x d) and j=1 (+y d), j=1 wins
    assert p.lF[0].tolist() == [0.0, d, 0.0]
    assert np.all(p.lF[1:] == 0)
    # a third duplicate (argmin z too) with d = 0 wipes both pushes
    p.lx[0, 2] = 4.0
    p.lF[:] = 0
    ibo.particle_to_wall([p], [20, 20, 20], dx=1.0)
    assert np.all(p.lF == 0)


def test_wall_spring_far_from_walls_is_noop():
    p = ibo.Particle(kind="sphere", r=3, lagrange=ibo.surface_points(3, 3, 3))
    p.lx = p.lag + np.float32(10)
    p.lF = np.ones_like(p.lx)
    ibo.particle_to_wall([p], [20, 20, 20])
    assert np.all(p.lF == 1)


# ---------------------------------------------------------------- rigid body
def test_update_from_lar_closed_form():
    """No marker force: cF = (1 - rho_p) V g, cu = cF/V dt, cx += cu dt, cT = 0;
    a uniform marker force F on every marker adds -F sum(lm) = -F area."""
    p = ibo.Particle(kind="sphere", r=3, rou=1.5, center=(10, 10, 10), pdtype=np.float64)
    p.lx = p.lag + np.float32(10)
    p.update_from_lar(dt=1.0, gl=2e-5)
    V = float(np.float32(4 / 3 * math.pi * 27))
    assert np.allclose(p.cF, [(1 - 1.5) * V * np.float32(2e-5), 0, 0], rtol=1e-6)
    assert np.allclose(p.cu, [(1 - 1.5) * np.float32(2e-5), 0, 0], rtol=1e-6)
    assert np.allclose(p.cx, [10 + p.cu[0], 10, 10], rtol=1e-12)
    assert np.all(p.cT == 0)
    q = ibo.Particle(kind="sphere", r=3, rou=1.0, center=(10, 10, 10), pdtype=np.float64)
    q.lx = q.lag + np.float32(10)
    q.lF = np.tile(np.array([1e-3, 0, 0], np.float32), (len(q.lx), 1))
    q.update_from_lar(dt=1.0, gl=0.0)
    assert np.allclose(q.cF, [-1e-3 * float(q.lm.sum()), 0, 0], rtol=1e-6)
    assert np.allclose(float(q.lm.sum()), float(np.float32(4 * math.pi * 9)), rtol=1e-5)


def test_torque_of_tangential_forces():
    """Forces f = e_z x r on the markers give torque -sum(r x f) lm along -z
    times sum(|r_perp|^2 lm)."""
    p = ibo.Particle(kind="sphere", r=4, rou=1.0, center=(0, 0, 0), pdtype=np.float64)
    r = p.lag.astype(np.float64)
    p.lx = p.lag.copy()
    p.lF = np.cross(np.array([0, 0, 1.0]), r).astype(np.float32)
    p.update_from_lar(dt=1.0, gl=0.0)
    expect = -np.sum((r[:, 0] ** 2 + r[:, 1] ** 2) * p.lm[:, 0])
    assert np.allclose(p.cT, [0, 0, expect], rtol=1e-5, atol=1e-6 * abs(expect))


def test_pose_update_rotates_markers():
    """angle += cw (no dt); a quarter turn about z maps (x, y) -> (-y, x)."""
    lag = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 3.0]], np.float32)
    p = ibo.Particle(kind="sphere", r=3, angle=(0, 0, 0), center=(5, 5, 5), lagrange=lag)
    p.cw = np.array([0, 0, np.pi / 2])
    p.update()
    assert np.allclose(p.lx, [[5, 6, 5], [3, 5, 5], [5, 5, 8]], atol=1e-6)


# ---------------------------------------------------------------- coupling
def test_first_step_markers_at_origin():
    """lx starts at 0 (src/funlbm/particle/base/base.py:112), so the first
    step's interpolation and spreading use the corner stencil [0, 4)^3 and the
    particle only feels gravity."""
    size = [15, 15, 15]
    st = orc.init_flow(19, (16, 16, 16), mu=0.05, bc=[orc.WALL] * 6)
    orc.init_solver(st)
    p = ibo.Particle(kind="sphere", r=3, rou=1.5, center=(8, 8, 8), lagrange=ibo.surface_points(3, 3, 3))
    assert np.all(p.lx == 0)
    ibo.run_step(st, [p], size, gl=1e-4)
    assert np.all(st.FOL[4:] == 0) and np.all(st.FOL[:, 4:] == 0) and np.all(st.FOL[:, :, 4:] == 0)
    assert np.allclose(p.cu, [-0.5 * np.float32(1e-4), 0, 0], rtol=1e-6)
    assert np.allclose(p.lx.mean(0), p.cx.astype(np.float32), atol=1e-5)


def test_spreading_overwrites_in_marker_order():
    """Two markers sharing stencil nodes: the later one's value survives."""
    st = orc.init_flow(19, (12, 12, 12), mu=0.05, bc=[orc.WALL] * 6)
    p = ibo.Particle(kind="sphere", r=3, lagrange=np.zeros((2, 3), np.float32))
    p.lx = np.array([[5.2, 5.0, 5.0], [5.7, 5.0, 5.0]], np.float32)
    p.lF = np.array([[1.0, 0, 0], [2.0, 0, 0]], np.float32)
    ibo.lagrange_to_flow(st, [p], [11, 11, 11])
    rl, rr = ibo.region_bounds(p.lx, [11, 11, 11])
    w1 = ibo.weight_function(ibo._grid(rl[1], rr[1]).astype(np.float32) - p.lx[1])
    sl = tuple(slice(rl[1, d], rr[1, d]) for d in range(3))
    assert np.array_equal(st.FOL[sl][..., 0:1], w1 * np.float32(2.0) * p.lm[1])

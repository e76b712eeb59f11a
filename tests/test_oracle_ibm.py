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


# ---------------------------------------------------------------- interpolation pins
# flow_to_lagrange (src/funlbm/lbm/lbm3d.py:66-100) against closed forms that
# do not restate its formula: a uniform field interpolates to itself under
# any normalised kernel, so s_u is the component sum a + b + c (F7) and
# s_rho the density; lu is the rigid-body velocity; and on a node the
# cosine kernel's per-axis weights are (1/4, 1/2, 1/4, 0).
UNIFORM_U = (0.01, -0.02, 0.005)
UNIFORM_RHO = 1.07
RIGID = dict(cu=(0.003, -0.001, 0.002), cw=(0.001, 0.002, -0.0015), cx=(6.2, 7.1, 5.9))
MARKERS = np.array([[6.7, 7.3, 8.2], [4.1, 9.6, 5.5], [8.05, 6.45, 3.9], [5.0, 6.0, 7.0]], np.float32)
SIZE = [14, 15, 13]
DT = 0.5


def _uniform_case():
    nodes = tuple(s + 1 for s in SIZE)
    st = orc.init_flow(19, nodes, mu=0.05, bc=[orc.PERIODICAL] * 6)
    st.u = np.broadcast_to(np.array(UNIFORM_U, np.float32), nodes + (3,)).copy()
    st.rou = np.full(nodes + (1,), UNIFORM_RHO, np.float32)
    p = ibo.Particle(kind="sphere", r=3, lagrange=np.zeros((len(MARKERS), 3), np.float32))
    p.lx = MARKERS.copy()
    p.cu = np.array(RIGID["cu"])
    p.cw = np.array(RIGID["cw"])
    p.cx = np.array(RIGID["cx"])
    return st, p


def _closed_form_lu_lF():
    cu, cw, cx = (np.array(RIGID[k]) for k in ("cu", "cw", "cx"))
    lu = cu + np.cross(cw, MARKERS.astype(np.float64) - cx)
    s_u = sum(UNIFORM_U)
    lF = 2.0 * UNIFORM_RHO * (lu - s_u) / DT
    return lu, lF


def _interp_pins_hold(f2l):
    """True if flow_to_lagrange `f2l` meets the closed forms above."""
    st, p = _uniform_case()
    f2l(st, [p], SIZE, DT, 1)
    lu, lF = _closed_form_lu_lF()
    scale = np.max(np.abs(lF))
    return bool(np.allclose(p.lu, lu, rtol=0, atol=1e-7) and np.max(np.abs(p.lF - lF)) < 1e-5 * scale)


def test_interpolation_of_uniform_field_closed_form():
    """s_u = a + b + c (the components summed into one scalar, F7),
    s_rho = rho, lu = cu + cw x (lx - cx), lF = 2 rho (lu - s_u)/dt."""
    assert _interp_pins_hold(ibo.flow_to_lagrange)
    st, p = _uniform_case()
    ibo.flow_to_lagrange(st, [p], SIZE, DT, 1)
    lu, lF = _closed_form_lu_lF()
    # every component of lF carries the same scalar s_u
    s_u = lu - p.lF.astype(np.float64) * DT / (2.0 * UNIFORM_RHO)
    assert np.allclose(s_u, sum(UNIFORM_U), rtol=0, atol=2e-7)


def test_cosine_weights_on_a_node_and_clamped_at_a_face():
    """A marker on a node: per-axis weights (1/4, 1/2, 1/4, 0) on the stencil
    [x-1, x+3), so the 3-D weights are their products and need no
    normalisation.  A marker at x = 0.3 (stencil clamped to [0, 4)): the
    weights phi(0.3), phi(0.7), phi(1.7) and 0 (|r| = 2.7 > 2) along x, divided
    by their sum (the node at x = -1 is missing)."""
    x = np.array([[5.0, 6.0, 7.0]], np.float32)
    rl, rr = ibo.region_bounds(x, SIZE)
    assert rl[0].tolist() == [4, 5, 6] and rr[0].tolist() == [8, 9, 10]
    w = ibo.weight_function(ibo._grid(rl[0], rr[0]).astype(np.float32) - x[0])[..., 0]
    axis = np.array([0.25, 0.5, 0.25, 0.0])
    want = axis[:, None, None] * axis[None, :, None] * axis[None, None, :]
    assert np.allclose(w, want, rtol=0, atol=1e-7)

    x = np.array([[0.3, 6.0, 7.0]], np.float32)
    rl, rr = ibo.region_bounds(x, SIZE)
    assert rl[0].tolist() == [0, 5, 6] and rr[0].tolist() == [4, 9, 10]
    w = ibo.weight_function(ibo._grid(rl[0], rr[0]).astype(np.float32) - x[0])[..., 0]
    phi = [(1 + math.cos(math.pi * r / 2)) / 4 for r in (0.3, 0.7, 1.7)]
    ax = np.array(phi + [0.0]) / sum(phi)
    want = ax[:, None, None] * axis[None, :, None] * axis[None, None, :]
    assert np.allclose(w, want, rtol=0, atol=2e-7)
    assert abs(float(w.sum()) - 1.0) < 1e-6


def _mutant_f2l(per_component=False, factor=2.0, sign=1.0, kernel_scale=0.5):
    """flow_to_lagrange with one planted mistake: components interpolated
    separately, another prefactor, the sign of (lu - s_u) flipped, or the
    kernel's argument pi r scaled differently."""

    def weights(r):
        mask = np.all(np.abs(r) <= 2, axis=-1, keepdims=True)
        tmp = np.where(mask, (np.float32(1) + np.cos(np.abs(r * np.float32(np.pi) * np.float32(kernel_scale))))
                       / np.float32(4), np.float32(0)).astype(np.float32)
        w = (tmp[..., 0] * tmp[..., 1] * tmp[..., 2])[..., None]
        s = np.sum(w)
        return (w / s if s > 0 else w).astype(np.float32)

    def f2l(st, particles, size, dt, step):
        for p in particles:
            rl, rr = ibo.region_bounds(p.lx, size)
            lu = np.zeros_like(p.lu)
            lrou = np.zeros((len(p.lx), 1), np.float32)
            for m, lar in enumerate(p.lx):
                sl = tuple(slice(rl[m, d], rr[m, d]) for d in range(3))
                w = weights(ibo._grid(rl[m], rr[m]).astype(np.float32) - lar)
                lu[m, :] = np.sum(st.u[sl] * w, axis=(0, 1, 2)) if per_component else np.sum(st.u[sl] * w)
                lrou[m, :] = np.sum(st.rou[sl] * w)
            p.lu = (p.cu + np.cross(p.cw, p.lx.astype(np.float64) - p.cx)).astype(np.float32)
            p.lF = (np.float32(factor) * lrou * np.float32(sign) * (p.lu - lu) / np.float32(dt)).astype(np.float32)
        return True

    return f2l


@pytest.mark.parametrize("mutation", [dict(per_component=True), dict(factor=1.0), dict(sign=-1.0),
                                      dict(kernel_scale=1.0), dict(kernel_scale=0.25)])
def test_interpolation_pins_catch_planted_mistakes(mutation):
    """Each planted mistake fails a pin above: the uniform field catches the
    per-component sum, the factor and the sign; the on-node weights catch a
    wrong kernel argument (the mutant itself is normalised, so the uniform
    field alone cannot)."""
    assert _interp_pins_hold(_mutant_f2l())  # the unmutated copy passes
    f2l = _mutant_f2l(**mutation)
    caught = not _interp_pins_hold(f2l)
    if "kernel_scale" in mutation:
        # the on-node pin, through the mutant's weights on a marker at a node
        st, p = _uniform_case()
        st.u = np.zeros_like(st.u)
        st.u[..., 0] = np.arange(st.u.shape[0], dtype=np.float32)[:, None, None]  # u_x = x
        p.lx = np.array([[5.0, 6.0, 7.0]], np.float32)
        p.cu = np.zeros(3)
        p.cw = np.zeros(3)
        p.lu = np.zeros((1, 3), np.float32)
        f2l(st, [p], SIZE, 1.0, 1)
        # the (1/4, 1/2, 1/4, 0) weights reproduce u_x = 5 on the node
        caught = caught or abs(float(p.lF[0, 0]) / (-2.0 * UNIFORM_RHO) - 5.0) > 1e-5
    assert caught, mutation


def test_on_node_interpolation_of_a_linear_field():
    """u_x = x sampled at a node: the weights (1/4, 1/2, 1/4, 0) average
    x - 1, x, x + 1 to x exactly (the oracle's result, to float32)."""
    st, p = _uniform_case()
    st.u = np.zeros_like(st.u)
    st.u[..., 0] = np.arange(st.u.shape[0], dtype=np.float32)[:, None, None]
    p.lx = np.array([[5.0, 6.0, 7.0]], np.float32)
    p.cu = np.zeros(3)
    p.cw = np.zeros(3)
    ibo.flow_to_lagrange(st, [p], SIZE, 1.0, 1)
    assert abs(float(p.lF[0, 0]) - (-2.0 * UNIFORM_RHO * 5.0)) < 1e-5


# ---------------------------------------------------------------- Peskin kernel
def test_peskin_kernel_moment_conditions():
    """Peskin's 4-point kernel: over the four nodes around any point its
    values sum to 1 (even and odd nodes to 1/2 each), its first moment
    vanishes and its squares sum to 3/8 (Acta Numerica 11 (2002) 479-517,
    eqs. 6.8-6.12); phi(0) = 1/2, phi(1) = 1/4, phi(2) = 0, continuous at 1."""
    x = np.linspace(0.0, 1.0, 101, dtype=np.float64)[:, None]
    j = np.array([-1, 0, 1, 2], dtype=np.float64)[None, :]
    r = (j - x).astype(np.float32)
    phi = ibo.peskin4(r).astype(np.float64)
    assert np.allclose(phi.sum(1), 1.0, atol=2e-7)
    assert np.allclose(phi[:, ::2].sum(1), 0.5, atol=2e-7)
    assert np.allclose((phi * (j - x)).sum(1), 0.0, atol=5e-7)
    assert np.allclose((phi * phi).sum(1), 3.0 / 8.0, atol=5e-7)
    v = ibo.peskin4(np.array([0.0, 1.0, 2.0, 1.0 - 1e-6, 1.0 + 1e-6], np.float32))
    assert np.allclose(v[:3], [0.5, 0.25, 0.0], atol=1e-7)
    assert abs(float(v[3]) - float(v[4])) < 1e-5


def test_peskin_weights_interpolate_linear_fields_exactly():
    """With Peskin weights an interior marker interpolates u_x = x exactly
    (first moment condition), unlike the cosine kernel off the nodes."""
    st, p = _uniform_case()
    st.u = np.zeros_like(st.u)
    st.u[..., 0] = np.arange(st.u.shape[0], dtype=np.float32)[:, None, None]
    p.lx = np.array([[5.3, 6.6, 7.1], [8.85, 4.2, 3.5], [3.5, 9.9, 6.25], [10.05, 5.5, 4.75]], np.float32)
    p.cu = np.zeros(3)
    p.cw = np.zeros(3)
    ibo.flow_to_lagrange(st, [p], SIZE, 1.0, 1, delta="peskin4")
    s_u = -p.lF[:, 0].astype(np.float64) / (2.0 * UNIFORM_RHO)
    assert np.allclose(s_u, p.lx[:, 0], atol=2e-5)
    p.lu = np.zeros_like(p.lu)
    ibo.flow_to_lagrange(st, [p], SIZE, 1.0, 1)  # the cosine kernel is not exact
    s_cos = -p.lF[:, 0].astype(np.float64) / (2.0 * UNIFORM_RHO)
    assert np.max(np.abs(s_cos - p.lx[:, 0])) > 1e-3

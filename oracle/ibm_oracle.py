"""CPU oracle of the immersed-boundary / rigid-particle part of the funlbm
timestep.  TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg may import it, as the checker.

Restates, with per-marker Python loops like the reference:

  marker generator     src/funlbm/particle/base/ellipsoid.py:10-95 (scipy
                       fsolve per vertex pair, as the reference; and the
                       closed-form root, vectorised, pinned against it)
  particle constants   src/funlbm/particle/base/sphere.py:15-53,
                       ellipsoid.py:107-143, base.py:108-118
  stencil / weights    src/funlbm/lbm/lbm3d.py:32-64 (and Peskin's 4-point
                       kernel as the build's option)
  flow_to_lagrange     src/funlbm/lbm/lbm3d.py:66-100
  particle_to_wall     src/funlbm/lbm/lbm3d.py:121-167
  update_from_lar      src/funlbm/particle/base/base.py:120-151
  lagrange_to_flow     src/funlbm/lbm/lbm3d.py:102-119
  particle update      src/funlbm/particle/base/base.py:153-156,
                       src/funlbm/particle/coord.py:73-108 (scipy Rotation)
  run_step             src/funlbm/lbm/base.py:146-179

The particle state (cx, cu, cw, cF, cT, rotation vector) is kept in
`pdtype`: float32 reproduces the reference's tensors, float64 the build's
double-precision particle state.  Marker arrays are float32 as the reference's.
"""

import math

import numpy as np
from scipy.optimize import fsolve
from scipy.spatial.transform import Rotation

from . import lbm_oracle as lbo


def find_intersection(p1, p2, value):
    """fsolve from t = 0.5 on the segment (src/funlbm/particle/base/ellipsoid.py:10-28)."""

    def eq(t):
        x = p1[0] + t * (p2[0] - p1[0])
        y = p1[1] + t * (p2[1] - p1[1])
        z = p1[2] + t * (p2[2] - p1[2])
        return value(x, y, z)

    t = fsolve(eq, 0.5)[0]
    return np.array([p1[0] + t * (p2[0] - p1[0]), p1[1] + t * (p2[1] - p1[1]), p1[2] + t * (p2[2] - p1[2])])


def surface_points(a, b, c, dx=1.0):
    """generate_uniform_points_on_ellipsoid (src/funlbm/particle/base/ellipsoid.py:31-95)."""

    def value(x, y, z):
        return x ** 2 / a ** 2 + y ** 2 / b ** 2 + z ** 2 / c ** 2 - 1

    xs = np.arange(-a - 2 * dx, a + 2 * dx, dx)
    ys = np.arange(-b - 2 * dx, b + 2 * dx, dx)
    zs = np.arange(-c - 2 * dx, c + 2 * dx, dx)
    pts = []
    for i in range(len(xs) - 1):
        for j in range(len(ys) - 1):
            for k in range(len(zs) - 1):
                x0, x1 = xs[i], xs[i + 1]
                y0, y1 = ys[j], ys[j + 1]
                z0, z1 = zs[k], zs[k + 1]
                verts = [(x0, y0, z0), (x0, y0, z1), (x0, y1, z0), (x0, y1, z1),
                         (x1, y0, z0), (x1, y0, z1), (x1, y1, z0), (x1, y1, z1)]
                inside = [value(*v) < 0 for v in verts]
                if any(inside) and not all(inside):
                    inter = []
                    for i1 in range(8):
                        for i2 in range(i1 + 1, 8):
                            if inside[i1] != inside[i2]:
                                inter.append(find_intersection(np.array(verts[i1]), np.array(verts[i2]), value))
                    if len(inter) >= 3:
                        pts.append(np.mean(inter, axis=0))
    return np.array(pts)


# vertex order of a cube (src/funlbm/particle/base/ellipsoid.py:55-64)
_CUBE_VERTS = np.array(
    [[0, 0, 0], [0, 0, 1], [0, 1, 0], [0, 1, 1], [1, 0, 0], [1, 0, 1], [1, 1, 0], [1, 1, 1]],
    dtype=np.int64,
)
_PAIRS = [(i, j) for i in range(8) for j in range(i + 1, 8)]


def _segment_root(p1, p2, inv2):
    """t in [0, 1] with value(p1 + t (p2 - p1)) = 0 for value = sum x^2 inv2 - 1;
    p1, p2 [n, 3] on opposite sides."""
    d = p2 - p1
    A = np.sum(d * d * inv2, axis=1)
    B = 2.0 * np.sum(p1 * d * inv2, axis=1)
    C = np.sum(p1 * p1 * inv2, axis=1) - 1.0
    disc = np.sqrt(np.maximum(B * B - 4.0 * A * C, 0.0))
    # numerically stable pair of roots
    qv = -0.5 * (B + np.copysign(disc, B))
    with np.errstate(divide="ignore", invalid="ignore"):
        r1 = qv / A
        r2 = C / qv
    # exactly one root lies in [0, 1]; rounding can push it just outside, so
    # take the root nearer to the interval
    d1 = np.maximum(np.maximum(-r1, r1 - 1.0), 0.0)
    d2 = np.maximum(np.maximum(-r2, r2 - 1.0), 0.0)
    d1 = np.where(np.isfinite(r1), d1, np.inf)
    d2 = np.where(np.isfinite(r2), d2, np.inf)
    return np.where(d1 <= d2, r1, r2)


def surface_points_closed_form(a, b, c, dx=1.0):
    """surface_points with each segment root in closed form, vectorised over
    the cubes: markers [M, 3] (float64) in the same order.  On a segment whose
    end points lie on opposite sides of the convex quadric the quadratic in t
    has exactly one root in [0, 1], the one fsolve from t = 0.5 converges to."""
    semi = np.array([a, b, c], dtype=np.float64)
    inv2 = 1.0 / (semi * semi)
    axes = [np.arange(-s - 2 * dx, s + 2 * dx, dx) for s in (a, b, c)]
    nx, ny, nz = (len(ax) for ax in axes)
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    val = X * X / (a * a) + Y * Y / (b * b) + Z * Z / (c * c) - 1
    inside = val < 0
    # cubes (i, j, k) with i < nx-1 ...
    ci, cj, ck = np.meshgrid(np.arange(nx - 1), np.arange(ny - 1), np.arange(nz - 1), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    ins = np.stack(
        [inside[ci + v[0], cj + v[1], ck + v[2]] for v in _CUBE_VERTS], axis=1
    )  # [ncubes, 8]
    cut = ins.any(axis=1) & ~ins.all(axis=1)
    ci, cj, ck, ins = ci[cut], cj[cut], ck[cut], ins[cut]
    verts = np.stack(
        [
            np.stack([axes[0][ci + v[0]], axes[1][cj + v[1]], axes[2][ck + v[2]]], axis=1)
            for v in _CUBE_VERTS
        ],
        axis=1,
    )  # [ncut, 8, 3]
    total = np.zeros((len(ci), 3))
    count = np.zeros(len(ci), dtype=np.int64)
    for i, j in _PAIRS:
        diff = ins[:, i] != ins[:, j]
        if not diff.any():
            continue
        p1 = verts[diff, i]
        p2 = verts[diff, j]
        t = _segment_root(p1, p2, inv2)
        pt = p1 + t[:, None] * (p2 - p1)
        total[diff] = total[diff] + pt
        count[diff] += 1
    keep = count >= 3
    return total[keep] / count[keep][:, None]


class Particle:
    """Particle constants and state (src/funlbm/particle/base/base.py:22-156)."""

    def __init__(self, kind="sphere", r=None, abc=None, rou=1.0, center=(0, 0, 0),
                 angle=(np.pi / 2, 0, 0), gravity=None, pdtype=np.float64, lagrange=None):
        self.pd = np.dtype(pdtype)
        if kind == "sphere":
            r = 10 if r is None else r
            a = b = c = r
        else:
            a, b, c = abc if abc is not None else (10, 10, 10)
        self.kind = kind
        lag = surface_points(a, b, c) if lagrange is None else lagrange
        self.lag = np.asarray(lag, dtype=np.float32)
        M = len(self.lag)
        V = 4.0 / 3.0 * math.pi * a * b * c
        self.mass = np.float32(V)
        if kind == "sphere":
            self.area = np.float32(4.0 * math.pi * a * a)
            I = np.array([a * a] * 3) * self.mass / 5.0
        else:
            self.area = np.float32(4.0 / 3.0 * math.pi * math.pow(a * b * c, 2.0 / 3.0))
            I = np.array([b * c, a * c, a * b]) * self.mass / 5.0
        self.I = I.astype(np.float32)
        self.rou = float(rou)
        self.gravity = None if gravity is None else np.asarray(gravity, dtype=np.float32)
        self.cx = np.asarray(center, dtype=np.float32).astype(self.pd)
        self.angle = np.asarray(angle, dtype=np.float32).astype(self.pd)
        self.cu = np.zeros(3, self.pd)
        self.cw = np.zeros(3, self.pd)
        self.cF = np.zeros(3, self.pd)
        self.cT = np.zeros(3, self.pd)
        self.lx = np.zeros((M, 3), np.float32)
        self.lF = np.zeros((M, 3), np.float32)
        self.lu = np.zeros((M, 3), np.float32)
        self.lm = np.full((M, 1), np.float32(self.area) / np.float32(M), dtype=np.float32)

    def update_from_lar(self, dt, gl, rouf=1.0):
        """src/funlbm/particle/base/base.py:120-151."""
        if dt <= 0:
            raise ValueError("Time step must be positive")
        if rouf <= 0:
            raise ValueError("Fluid density must be positive")
        pd = self.pd
        g = self.gravity if self.gravity is not None else np.array([gl, 0, 0], dtype=np.float32)
        mass = pd.type(self.mass)
        tmp = (pd.type(1) - pd.type(self.rou / rouf)) * mass * g.astype(pd)
        lF = self.lF.astype(pd)
        lm = self.lm.astype(pd)
        self.cF = np.sum(-lF * lm, axis=0) + tmp
        self.cu = self.cu + self.cF / mass * pd.type(dt)
        self.cx = self.cx + self.cu * pd.type(dt)
        self.cT = -np.sum(np.cross(self.lx.astype(pd) - self.cx, lF) * lm, axis=0)
        self.cw = self.cw + pd.type(0.1) * self.cT * pd.type(dt) / self.I.astype(pd)

    def update(self):
        """angle += cw; lx = float32(R lag) + float32(cx)
        (src/funlbm/particle/base/base.py:153-156, coord.py:73-108)."""
        self.angle = self.angle + self.cw
        R = Rotation.from_rotvec(np.asarray(self.angle, dtype=np.float64))
        self.lx = R.apply(self.lag.astype(np.float64)).astype(np.float32) + self.cx.astype(np.float32)


def region_bounds(lx, size):
    """src/funlbm/lbm/lbm3d.py:32-44."""
    rl = np.floor(lx) - 1
    rl = np.maximum(rl, 0)
    rr = rl + 4
    rr = np.minimum(rr, np.asarray(size, dtype=np.float32))
    return rl.astype(np.int32), rr.astype(np.int32)


def peskin4(r):
    """Peskin's 4-point kernel (Acta Numerica 11 (2002), eq. 6.27), float32,
    for |r| <= 2 (the caller masks the rest).  Not a reference function: the
    option the build adds beside the reference's cosine kernel."""
    a = np.abs(np.asarray(r, dtype=np.float32))
    one, two, three, four, five, seven, eight, twelve = (np.float32(v) for v in (1, 2, 3, 4, 5, 7, 8, 12))
    inner = (three - two * a + np.sqrt(np.maximum(one + four * a - four * a * a, 0))) / eight
    outer = (five - two * a - np.sqrt(np.maximum(-seven + twelve * a - four * a * a, 0))) / eight
    return np.where(a <= one, inner, outer).astype(np.float32)


def weight_function(r, delta="cosine"):
    """src/funlbm/lbm/lbm3d.py:46-64 (float32); delta="peskin4" swaps the
    per-axis cosine kernel for peskin4."""
    mask = np.all(np.abs(r) <= 2, axis=-1, keepdims=True)
    if delta == "peskin4":
        phi = peskin4(r)
    else:
        phi = (np.float32(1) + np.cos(np.abs(r * np.float32(np.pi) / np.float32(2)))) / np.float32(4)
    tmp = np.where(mask, phi, np.float32(0)).astype(np.float32)
    w = (tmp[..., 0] * tmp[..., 1] * tmp[..., 2])[..., None]
    s = np.sum(w)
    if s > 0:
        w = w / s
    return w.astype(np.float32)


def _grid(rl, rr):
    gx, gy, gz = np.meshgrid(np.arange(rl[0], rr[0]), np.arange(rl[1], rr[1]), np.arange(rl[2], rr[2]), indexing="ij")
    return np.stack([gx, gy, gz], axis=-1)


def flow_to_lagrange(st: lbo.FlowState, particles, size, dt, step, delta="cosine"):
    """src/funlbm/lbm/lbm3d.py:66-100; returns False when a weight sum fails
    the check after step 100."""
    ok = True
    for p in particles:
        rl, rr = region_bounds(p.lx, size)
        lu = np.zeros_like(p.lu)
        lrou = np.zeros((len(p.lx), 1), np.float32)
        for m, lar in enumerate(p.lx):
            sl = (slice(rl[m, 0], rr[m, 0]), slice(rl[m, 1], rr[m, 1]), slice(rl[m, 2], rr[m, 2]))
            x = _grid(rl[m], rr[m]).astype(np.float32)
            w = weight_function(x - lar, delta)
            if abs(np.sum(w) - 1) > 1e-3 and step > 100:
                ok = False
            lu[m, :] = np.sum(st.u[sl] * w)
            lrou[m, :] = np.sum(st.rou[sl] * w)
        pd = p.pd
        d = p.lx.astype(pd) - p.cx
        p.lu = (p.cu + np.cross(p.cw, d)).astype(np.float32)
        p.lF = (np.float32(2) * lrou * (p.lu - lu) / np.float32(dt)).astype(np.float32)
    return ok


def particle_to_wall(particles, size, dx=1.0):
    """src/funlbm/lbm/lbm3d.py:121-167."""
    normals = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.float32)
    bounds = np.concatenate([np.zeros(3, np.float32), np.asarray(size, dtype=np.float32)])
    for p in particles:
        ext = np.concatenate([np.argmin(p.lx, axis=0), np.argmax(p.lx, axis=0)])
        coords = p.lx[ext][[0, 1, 2, 3, 4, 5], [0, 1, 2, 0, 1, 2]]
        d = np.float32(300) * (np.float32(1) - np.abs(coords - bounds) / np.float32(2 * dx))
        d = np.maximum(d, np.float32(0)).astype(np.float32)
        if d.max() == 0:
            continue
        rows = p.lF[ext] + normals * d[:, None]
        for j in range(6):  # index_put without accumulate: the last duplicate wins
            p.lF[ext[j]] = rows[j]


def lagrange_to_flow(st: lbo.FlowState, particles, size, delta="cosine"):
    """src/funlbm/lbm/lbm3d.py:102-119: FOL := 0, then overwrite per marker."""
    st.FOL = np.zeros_like(st.FOL)
    for p in particles:
        rl, rr = region_bounds(p.lx, size)
        for m, lar in enumerate(p.lx):
            sl = (slice(rl[m, 0], rr[m, 0]), slice(rl[m, 1], rr[m, 1]), slice(rl[m, 2], rr[m, 2]))
            x = _grid(rl[m], rr[m]).astype(np.float32)
            w = weight_function(x - lar, delta)
            st.FOL[sl] = (w * p.lF[m, :] * p.lm[m]).astype(st.dt)


def run_step(st: lbo.FlowState, particles, size, dt=1.0, dx=1.0, gl=0.0, step=1, delta="cosine"):
    """LBMBase.run_step (src/funlbm/lbm/base.py:146-179); returns run_status."""
    lbo.cul_equ(st)
    lbo.f_stream(st)
    lbo.update_u_rou(st)
    ok = flow_to_lagrange(st, particles, size, dt, step, delta)
    particle_to_wall(particles, size, dx)
    for p in particles:
        p.update_from_lar(dt, gl)
    lagrange_to_flow(st, particles, size, delta)
    lbo.cul_equ2(st)
    lbo.update_u_rou(st)
    for p in particles:
        p.update()
    return ok

"""Surface markers of an ellipsoid (init-time geometry, host side).

Follows generate_uniform_points_on_ellipsoid / find_intersection
(src/funlbm/particle/base/ellipsoid.py:10-95): unit cubes of edge dx tile the
box [-a-2dx, a+2dx) x ... (np.arange), a cube whose 8 vertices are neither all
inside nor all outside (value < 0 is inside) yields the intersection points of
every vertex pair of differing side (all 28 pairs, in the reference's pair
order), and the centroid of those points is a marker when there are at least
3 of them.  Markers are in cube order (x outermost, z innermost).

The reference finds each intersection with scipy.optimize.fsolve started at
t = 0.5; on a segment whose end points lie on opposite sides of a convex
quadric the quadratic has exactly one root in (0, 1), which is computed here
in closed form.
"""

import numpy as np

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


def ellipsoid_markers(a: float, b: float, c: float, dx: float = 1.0) -> np.ndarray:
    """Marker positions [M, 3] (float64) in the body frame."""
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

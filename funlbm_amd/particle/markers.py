"""Surface markers of ellipsoids (init-time geometry), generated on the GPU.

Replaces generate_uniform_points_on_ellipsoid / find_intersection
(src/funlbm/particle/base/ellipsoid.py:10-95) as Sphere._init and
Ellipsoid._init call them (src/funlbm/particle/base/sphere.py:15-53,
ellipsoid.py:107-143): unit cubes of edge dx tile the box
[-a-2dx, a+2dx) x ... (numpy.arange), a cube whose 8 vertices are neither all
inside nor all outside yields the intersection points of every vertex pair of
differing side (all 28 pairs, in the reference's pair order), and the centroid
of those points is a marker when there are at least 3 of them.  Markers are
in cube order (x outermost, z innermost).

The kernel (funlbm_surface_markers) handles every shape of a call in one
launch, one workgroup per shape, in double; the results are cached per
(a, b, c, dx), so a shape is generated once per process.
"""

import math
from typing import Dict, List, Sequence, Tuple

import numpy as np
import torch

from .. import _native as nat

_cache: Dict[Tuple[float, float, float, float], np.ndarray] = {}


def _area_estimate(a: float, b: float, c: float) -> float:
    """Surface area of the ellipsoid (Thomsen's approximation, within ~1.1 %)."""
    p = 1.6075
    return 4.0 * math.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3.0) ** (1.0 / p)


def _device(device) -> torch.device:
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    if dev.type != "cuda":
        raise RuntimeError(f"the marker generator runs on the GPU, not on {dev}")
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device())
    return dev


def surface_markers(shapes: Sequence[Sequence[float]], dx: float = 1.0, device=None) -> List[np.ndarray]:
    """Body-frame markers (float32 [M, 3], read-only host arrays) of each
    ellipsoid (a, b, c) in `shapes`; a sphere of radius r is (r, r, r)."""
    dx = float(dx)
    if not (dx > 0.0 and math.isfinite(dx)):
        raise ValueError(f"marker spacing must be positive, got {dx}")
    keys = []
    for s in shapes:
        a, b, c = (float(v) for v in s)
        if not all(v > 0.0 and math.isfinite(v) for v in (a, b, c)):
            raise ValueError(f"semi-axes must be positive, got {(a, b, c)}")
        keys.append((a, b, c, dx))
    todo = list(dict.fromkeys(k for k in keys if k not in _cache))
    if todo:
        _generate(todo, dx, _device(device))
    return [_cache[k] for k in keys]


def _generate(todo, dx, dev):
    semi = torch.tensor([k[:3] for k in todo], dtype=torch.float64, device=dev)
    # the cut cubes number about 1.5 per unit of area; the second pass (exact
    # capacity) only runs if the estimate was short
    cap = max(int(3.0 * _area_estimate(*k[:3]) / (dx * dx)) + 64 for k in todo)
    with torch.cuda.device(dev):
        for _ in range(2):
            out = torch.empty((len(todo), cap, 3), dtype=torch.float32, device=dev)
            counts = torch.empty(len(todo), dtype=torch.int32, device=dev)
            nat.call("funlbm_surface_markers", len(todo), nat.ptr(semi), dx, cap, nat.ptr(out),
                     nat.ptr(counts), nat.stream_handle(dev))
            cnt = counts.cpu().numpy()
            if np.any(cnt < 0):
                bad = [todo[i][:3] for i in np.nonzero(cnt < 0)[0]]
                raise ValueError(f"shapes too large for the marker generator: {bad}")
            if int(cnt.max()) <= cap:
                break
            cap = int(cnt.max())
        assert int(cnt.max()) <= cap
        host = out.cpu().numpy()
    for i, k in enumerate(todo):
        arr = np.ascontiguousarray(host[i, : int(cnt[i])])
        arr.setflags(write=False)
        _cache[k] = arr


def fibonacci_markers(a: float, b: float, c: float, dx: float = 1.0) -> np.ndarray:
    """Body-frame markers on a Fibonacci (golden-angle) lattice, a
    non-default option beside the reference's cube-intersection markers:
    N = round(A / dx^2) points for surface area A (exact for a sphere,
    Thomsen's approximation for an ellipsoid), one per dx^2 of surface;
    nearest neighbours then lie about 0.95 dx apart.  Point i of the
    unit sphere is (rho cos t, rho sin t, z) with z = 1 - (2i + 1)/N,
    rho = sqrt(1 - z^2), t = i pi (3 - sqrt 5) (equal-area bands, golden
    angle); an ellipsoid scales it by (a, b, c).  float32 [N, 3], read-only."""
    a, b, c, dx = (float(v) for v in (a, b, c, dx))
    if not all(v > 0.0 and math.isfinite(v) for v in (a, b, c, dx)):
        raise ValueError(f"semi-axes and spacing must be positive, got {(a, b, c)}, {dx}")
    key = ("fibonacci", a, b, c, dx)
    if key not in _fib_cache:
        area = 4.0 * math.pi * a * a if a == b == c else _area_estimate(a, b, c)
        n = max(4, int(round(area / (dx * dx))))
        i = np.arange(n, dtype=np.float64)
        z = 1.0 - (2.0 * i + 1.0) / n
        rho = np.sqrt(np.maximum(1.0 - z * z, 0.0))
        t = i * (math.pi * (3.0 - math.sqrt(5.0)))
        pts = np.stack([a * rho * np.cos(t), b * rho * np.sin(t), c * z], axis=1).astype(np.float32)
        pts.setflags(write=False)
        _fib_cache[key] = pts
    return _fib_cache[key]


_fib_cache: Dict[tuple, np.ndarray] = {}


def markers(kind: str, a: float, b: float, c: float, dx: float = 1.0, device=None) -> np.ndarray:
    """Markers of one shape by method: "cubes" (the reference's
    cube-intersection generator, on the GPU) or "fibonacci"."""
    if kind == "fibonacci":
        return fibonacci_markers(a, b, c, dx)
    if kind not in (None, "cubes"):
        raise ValueError(f"unknown marker method {kind!r}; expected 'cubes' or 'fibonacci'")
    return surface_markers([(a, b, c)], dx, device=device)[0]

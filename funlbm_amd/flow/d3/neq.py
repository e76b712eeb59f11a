"""The inlet velocity profile of the NON_EQUILIBRIUM boundary.

init_u (src/funlbm/flow/d3/base.py:10-37) evaluates, at the node
coordinates y = linspace(-(a-1)/2, (a-1)/2, a) and z likewise over b, the
series solution of laminar flow through a rectangular duct of half-widths
a/2 and b/2 (100 terms), normalised so that the parabola 1 - (y/ha)^2 is its
leading part, rounds each value to 6 decimals and scales it by u_max.  The
profile below is that table for u_max = 1 (the fused kernels scale it by
the per-step u_max on the device)."""

import numpy as np


def duct_profile(a: int, b: int, n_terms: int = 100) -> np.ndarray:
    """[a, b] float64 profile, rounded to 6 decimals as the reference does."""
    ha, hb = a / 2.0, b / 2.0
    y = np.linspace(-(a - 1) / 2, (a - 1) / 2, a)[:, None, None]
    z = np.linspace(-(b - 1) / 2, (b - 1) / 2, b)[None, :, None]
    n = np.arange(n_terms)
    k = (2 * n + 1)[None, None, :]
    base = 32.0 * (-1.0) ** (n + 1) / ((2 * n + 1) ** 3 * np.pi ** 3)
    series = np.sum(base[None, None, :] * np.cos(k * np.pi * y / (2 * ha))
                    * (np.cosh(k * np.pi * z / (2 * ha)) / np.cosh(k * np.pi * hb / (2 * ha))), axis=-1)
    val = 1.0 - (y[..., 0] / ha) ** 2 + series
    return np.round(val, 6)

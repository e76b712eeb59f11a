"""Storage formats against the float64 run of the same configuration (needs
a GPU).  Relative L-inf velocity error after the given steps, per format; the
bounds are the measured errors with a margin of about 1.5x or more.  Float32
populations limit flows whose velocities are near 1e-5 (their float32 ulp is
~3e-8), where the 16-bit deviation states are the more accurate."""

import json

import pytest

from funlbm_amd import precision

pytestmark = pytest.mark.gpu


def test_lid_driven_cavity_storage_errors():
    res = precision.compare(precision.cavity(), steps=2000)
    print(json.dumps(res))
    assert res["max_abs_u_f64"] > 1e-2
    assert res["f32"]["u_rel_linf"] < 5e-5
    assert res["f16"]["u_rel_linf"] < 1e-3
    assert res["bf16"]["u_rel_linf"] < 2e-2
    assert res["f16-sr"]["u_rel_linf"] < 1e-3
    assert res["bf16-sr"]["u_rel_linf"] < 2e-2


def test_poiseuille_storage_errors():
    res = precision.compare(precision.poiseuille(), steps=2000)
    print(json.dumps(res))
    assert res["max_abs_u_f64"] > 5e-3
    assert res["f32"]["u_rel_linf"] < 1e-4
    assert res["f16"]["u_rel_linf"] < 5e-3
    assert res["bf16"]["u_rel_linf"] < 0.25
    # stochastic stores keep the sub-ulp forcing increments that nearest
    # rounding drops
    assert res["f16-sr"]["u_rel_linf"] < 1e-3
    assert res["f16-sr"]["u_rel_linf"] < 0.5 * res["f16"]["u_rel_linf"]
    assert res["bf16-sr"]["u_rel_linf"] < 0.5 * res["bf16"]["u_rel_linf"]


def test_settling_sphere_centroid_errors():
    """A slowly settling sphere (velocities ~2e-5): float32 populations
    (ulp ~3e-8 near w) resolve u only to about 1e-3 of it and drift 2e-2 off
    float64 in 300 steps; the 16-bit states hold the small deviations and
    carry rho - 1 apart from rho, and stay within 1e-3 (centroids 1e-7 lu)."""
    cfg = precision.settling_sphere(size=(31, 31, 47), r=4, center=(16, 16, 30), g=4e-5)
    res = precision.compare(cfg, steps=300)
    print(json.dumps(res))
    assert res["f32"]["centroid_abs"] < 1e-4
    for st in ("f16", "f16-sr"):
        assert res[st]["u_rel_linf"] < 1e-3, st
        assert res[st]["centroid_abs"] < 1e-6, st
    for st in ("bf16", "bf16-sr"):
        assert res[st]["u_rel_linf"] < 5e-3, st
        assert res[st]["centroid_abs"] < 1e-6, st

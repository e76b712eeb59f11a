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


def test_dense_suspension_storage_errors():
    """The dense-suspension study at reduced size (D3Q27 48^3 periodic, three
    (4, 4, 8) ellipsoids, 500 steps).  Here the 16-bit states miss 1e-3: the
    reference zeroes u on every face after each moment update, which on
    periodic faces launches an acoustic field whose density fluctuation is
    10-20 times the decaying velocity, and a 16-bit deviation resolves u only
    to about 2^-11 of that density term.  Float32 populations do worse still
    (their ulp near w is above these velocities).  The bounds guard the
    measured errors with a margin; the 1e-3 bar is not met at this size."""
    cfg = precision.suspension(n=48, n_particles=3, mu=0.05 / 3)
    res = precision.compare(cfg, steps=500)
    print(json.dumps(res))
    assert res["max_abs_u_f64"] > 1e-6
    for st in ("f16", "f16-sr"):
        assert res[st]["u_rel_linf"] < 0.05 * res["f32"]["u_rel_linf"], st
        assert res[st]["u_rel_linf"] < 5e-3, st
        assert res[st]["centroid_abs"] < 1e-5, st
    for st in ("bf16", "bf16-sr"):
        assert res[st]["u_rel_linf"] < 4e-2, st

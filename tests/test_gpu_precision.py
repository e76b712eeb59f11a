"""Storage formats against the float64 run of the same configuration (needs
a GPU): float32 and f16 deviation storage stay within 1e-3 relative velocity
error; bf16 is reported."""

import json

import pytest

from funlbm_amd import precision

pytestmark = pytest.mark.gpu


def test_poiseuille_storage_errors():
    res = precision.compare(precision.poiseuille(), steps=1000)
    print(json.dumps(res))
    assert res["f32"]["u_rel_linf"] < 1e-5
    assert res["f16"]["u_rel_linf"] < 1e-3
    assert res["bf16"]["u_rel_linf"] < 1e-1


def test_settling_sphere_storage_errors():
    cfg = precision.settling_sphere(size=(31, 31, 47), r=4, center=(16, 16, 30), g=4e-5)
    res = precision.compare(cfg, steps=300)
    print(json.dumps(res))
    assert res["f32"]["u_rel_linf"] < 1e-4 and res["f32"]["centroid_abs"] < 1e-5
    assert res["f16"]["u_rel_linf"] < 1e-3 and res["f16"]["centroid_abs"] < 1e-3

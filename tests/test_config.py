"""Host-side configuration logic (CPU): the reference's JSON schema, its
`or`-defaults and its boundary-code lookup."""

import json

import numpy as np
import pytest

from funlbm_amd.config import BaseConfig, BoundaryCondition, deep_get
from funlbm_amd.flow import FlowConfig
from funlbm_amd.lbm import Config
from funlbm_amd.particle import ParticleConfig

# the free-fall example of the reference (example/freefall/config.json), retyped
FREEFALL = {
    "dx": 1,
    "dt": 1,
    "device": "cpu",
    "flow": {
        "param_type": "D3Q19",
        "param": {"Re": 0.45, "rho": 1.0, "mu": 0.0385, "gl": -0.0098},
        "size": [160, 100, 100],
        "boundary": {f: {"code": "PERIODICAL"} for f in ("input", "output", "back", "forward", "top", "bottom")},
    },
    "file": {"cache_dir": "./data", "per_steps": 100},
    "particles": [{"rou": 1.155, "coord": {"center": [40, 50, 50]}, "r": 7.5}],
}


def test_boundary_code_lookup():
    assert BoundaryCondition.find("PERIODICAL") is BoundaryCondition.PERIODICAL
    assert BoundaryCondition.find(1201) is BoundaryCondition.WALL_WITH_SPEED
    assert BoundaryCondition.find("NOPE") is BoundaryCondition.WALL  # unknown -> WALL
    assert BoundaryCondition.find(7) is BoundaryCondition.WALL
    assert BoundaryCondition.PERIODICAL.value == 11000 and BoundaryCondition.WALL.value == 1200


def test_deep_get_and_or_defaults():
    d = {"a": {"b": 0, "c": 3}, "x": ""}
    assert deep_get(d, "a.c") == 3 and deep_get(d, "a.z") is None
    cfg = BaseConfig().from_json(d)
    # falsy values fall back to the default, as `or` does in the reference
    assert cfg.get("a.b", 7) == 7 and cfg.get("x", "dflt") == "dflt"


def test_freefall_config_parses_like_the_reference(tmp_path):
    p = tmp_path / "config.json"
    p.write_text(json.dumps(FREEFALL))
    c = Config().from_file(str(p))
    fc = c.flow_config
    assert c.dt == 1 and c.dx == 1 and c.max_step == 10000
    assert list(fc.size) == [160, 100, 100] and fc.param_type == "D3Q19"
    assert fc.mu == 0.0385 and fc.gl == -0.0098 and fc.Re == 0.45
    assert all(b.is_condition(BoundaryCondition.PERIODICAL) for b in fc.boundary.faces())
    assert c.file_config.per_steps == 100
    (pc,) = c.particles
    # no "type": the reference builds an ellipsoid with a = b = c = 10
    assert pc.type == "ellipsoid" and pc.get("r") == 7.5 and pc.get("rou", 1.0) == 1.155
    assert pc.coord_config.center == [40, 50, 50]
    assert abs(pc.coord_config.alpha - np.pi / 2) < 1e-15 and pc.coord_config.beta == 0
    # the build's extension keys keep the reference's behaviour by default
    assert fc.storage == "f32" and fc.uzero_faces is None and fc.body_force is None and fc.gravity is None
    assert c.fused is True and c.log_every == 1


def test_defaults_without_keys():
    fc = FlowConfig().from_json({})
    assert list(fc.size) == [100, 100, 100] and fc.param_type == "D3Q19"
    assert all(b.is_condition(BoundaryCondition.WALL) for b in fc.boundary.faces())
    pc = ParticleConfig().from_json({"coord": {"alpha": 0}})
    assert abs(pc.coord_config.alpha - np.pi / 2) < 1e-15  # alpha: 0 cannot be set (or-default)


def test_unknown_lattice_raises_value_error():
    from funlbm_amd.flow import create_flow

    fc = FlowConfig().from_json({"param_type": "D3Q9"})
    with pytest.raises(ValueError, match="Unknown parameter type"):
        create_flow(fc)


def test_extension_keys():
    fc = FlowConfig().from_json({"storage": "f16", "dev_scale": 512, "uzero_faces": ["input", "output"],
                                 "body_force": [0, 0, 1e-6], "param": {"gravity": [0, 0, -1e-5]}})
    assert fc.storage == "f16" and fc.dev_scale == 512.0
    assert fc.uzero_faces == ["input", "output"] and fc.body_force == [0, 0, 1e-6]
    assert fc.gravity == [0, 0, -1e-5]
    from funlbm_amd.flow.d3.base import _uzero_mask

    assert _uzero_mask(None) == 0x3F and _uzero_mask(["input", "output"]) == 0b11
    assert _uzero_mask(["bottom", "top"]) == 0b110000


def test_cpu_device_is_refused():
    from funlbm_amd.base.work import device_detect

    with pytest.raises(RuntimeError, match="GPU"):
        device_detect("cpu")


def test_case_config_parses_to_the_same_case():
    """lbm_config_for writes a batched case (spheres and ellipsoids, walls,
    gravity, tau) in the reference's schema; parsed back, the solver's
    configuration holds the same lattice, faces and particles."""
    import math

    from funlbm_amd.batch.engine import CaseSpec, ParticleSpec
    from funlbm_amd.batch.workloads import lbm_config_for
    from funlbm_amd.lbm import Config

    spec = CaseSpec(nodes=(40, 36, 44), q=27, tau=0.7, steps=10, gravity=(0.0, 0.0, 3e-5),
                    particles=[ParticleSpec(kind="sphere", r=4.0, rou=1.4, center=(12.0, 11.0, 25.0)),
                               ParticleSpec(kind="ellipsoid", abc=(3.0, 3.0, 5.0), rou=2.1, center=(27.0, 24.0, 26.5),
                                            angle=(0.0, 0.4, -1.1), markers="fibonacci")],
                    uzero_mask=0x3F)
    cfg = Config().from_json(lbm_config_for(spec, "f16", "nearest"))
    fc = cfg.flow_config
    assert fc.param_type == "D3Q27"
    assert list(fc.size) == [39, 35, 43]
    assert math.isclose(fc.mu, 0.2 / 3)
    assert fc.storage == "f16" and fc.rounding == "nearest"
    assert list(fc.gravity) == [0.0, 0.0, 3e-5]
    from funlbm_amd.config import BoundaryCondition

    assert all(getattr(fc.boundary, f).is_condition(BoundaryCondition.WALL)
               for f in ("input", "output", "back", "forward", "bottom", "top"))
    ps = cfg.particles
    assert [p.type for p in ps] == ["sphere", "ellipsoid"]
    assert ps[1].markers == "fibonacci"
    assert ps[1].coord_config.alpha == 1e-9 and ps[1].coord_config.beta == 0.4 and ps[1].coord_config.gamma == -1.1

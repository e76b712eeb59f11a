from .base import Particle, ParticleConfig
from .ellipsoid import Ellipsoid
from .sphere import Sphere

__all__ = ["Particle", "ParticleConfig", "create_particle", "Ellipsoid", "Sphere"]


def create_particle(config: ParticleConfig, device=None) -> Particle:
    """"ellipsoid" (the default type) or anything else = sphere
    (src/funlbm/particle/base/__init__.py:8-12)."""
    if config.type == "ellipsoid":
        return Ellipsoid(config=config, device=device)
    return Sphere(config=config, device=device)

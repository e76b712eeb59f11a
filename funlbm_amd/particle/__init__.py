from .base import Ellipsoid, Particle, ParticleConfig, Sphere, create_particle
from .coord import CoordConfig, Coordinate
from .swarm import ParticleSwarm, create_particle_swarm

__all__ = [
    "Particle",
    "Ellipsoid",
    "Sphere",
    "ParticleConfig",
    "ParticleSwarm",
    "CoordConfig",
    "Coordinate",
    "create_particle_swarm",
    "create_particle",
]

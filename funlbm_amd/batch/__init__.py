from .alloc import RangeAllocator
from .engine import BatchEngine, CaseSpec, ParticleSpec

__all__ = ["BatchEngine", "CaseSpec", "ParticleSpec", "RangeAllocator"]

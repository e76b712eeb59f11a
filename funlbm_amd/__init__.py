"""MI355X-native (gfx950) per-timestep path of the funlbm lattice-Boltzmann
solver: D3Qn BGK collide-stream, bounce-back and periodic faces, and
immersed-boundary rigid particles, behind the reference's Python
flow/particle/solver API.  The arithmetic runs in hand-written HIP kernels
(libfunlbm_hip.so, C ABI in include/funlbm.h)."""

__version__ = "0.1.0"

"""ctypes binding of libfunlbm_hip.so (the C ABI declared in include/funlbm.h).

The library is required: importing a compute entry point without it raises
RuntimeError, there is no CPU fallback.
"""

import contextlib
import ctypes
import os
import threading

import torch

from .build import LIB_PATH

c_int32 = ctypes.c_int32
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_int64 = ctypes.c_int64

BC_PERIODICAL = 11000
BC_WALL = 1200
BC_WALL_WITH_SPEED = 1201
BC_NON_EQUILIBRIUM = 1400

ABI_VERSION = 4

DELTA_COSINE = 0
DELTA_PESKIN4 = 1
DELTA_CODES = {"cosine": DELTA_COSINE, "peskin4": DELTA_PESKIN4}

ROUND_NEAREST = 0
ROUND_STOCHASTIC = 1
ROUNDING_CODES = {"nearest": ROUND_NEAREST, "stochastic": ROUND_STOCHASTIC}

STORE_F32 = 0
STORE_F16 = 1
STORE_BF16 = 2
STORE_F64 = 3

STORAGE_CODES = {"f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16, "f64": STORE_F64}
STORAGE_DTYPES = {
    STORE_F32: torch.float32,
    STORE_F16: torch.float16,
    STORE_BF16: torch.bfloat16,
    STORE_F64: torch.float64,
}

# particle state layout (double[P][PSTRIDE])
P_CX, P_CU, P_CW, P_CF, P_CT, P_ANGLE, P_ROT = 0, 3, 6, 9, 12, 15, 18
P_MASS, P_INERTIA, P_RHO, P_GRAV = 27, 28, 31, 32
PSTRIDE = 40

EINVAL = -1
EUNSUPPORTED_BC = -2
EUNSUPPORTED_LATTICE = -3
ETOOBIG = -4
EFAILED = -5


class Lattice(ctypes.Structure):
    """funlbm_lattice of include/funlbm.h."""

    _fields_ = [
        ("q", c_int32),
        ("nx", c_int32),
        ("ny", c_int32),
        ("nz", c_int32),
        ("storage", c_int32),
        ("bc", c_int32 * 6),
        ("uzero_mask", c_int32),
        ("tau", c_float),
        ("rho0", c_float),
        ("dev_scale", c_float),
        ("body_force", c_float * 3),
        ("uw", (c_float * 3) * 6),
        ("rounding", c_int32),
        ("seed", ctypes.c_uint32),
        ("delta", c_int32),
    ]


P = c_void_p
LATP = ctypes.POINTER(Lattice)

# name -> (restype, argtypes)
SIGNATURES = {
    "funlbm_abi_version": (c_int32, []),
    "funlbm_last_error": (ctypes.c_char_p, []),
    "funlbm_state_bytes": (c_int32, [LATP, ctypes.POINTER(c_int64)]),
    "funlbm_collide": (c_int32, [LATP, P, P, P, P, P]),
    "funlbm_stream": (c_int32, [LATP, P, P, P, P, P]),
    "funlbm_moments": (c_int32, [LATP, P, P, P, P, P, P]),
    "funlbm_neq_moments": (c_int32, [LATP, P, P, P, ctypes.c_float, P, P]),
    "funlbm_forcing": (c_int32, [LATP, P, P, P, P]),
    "funlbm_pack": (c_int32, [LATP, P, P, P]),
    "funlbm_equilibrium_state": (c_int32, [LATP, P, P, P, P]),
    "funlbm_init_state": (c_int32, [LATP, P, P]),
    "funlbm_fused_step": (c_int32, [LATP, P, P, P, P, P, P, P, P, P]),
    "funlbm_fused_step_stats": (c_int32, [LATP, P, P, P, P, P, P, P, P, P, c_int32, P, P]),
    "funlbm_fused_step_stats_noalloc": (c_int32, [LATP, P, P, P, P, P, P, P, P, P, P]),
    "funlbm_materialize": (c_int32, [LATP, P, P, P, P, P, P, P, P, P, P, P]),
    "funlbm_copy_calibrate": (c_int32, [P, P, c_int64, P]),
    "funlbm_ibm_interp": (
        c_int32,
        [LATP, P, P, P, P, P, c_int32, P, P, P, c_float, c_int32, P, P, P, P],
    ),
    "funlbm_ibm_interp_boxed": (
        c_int32,
        [LATP, P, P, P, c_int32, P, P, c_int32, P, P, c_int32, P, P, c_int32, P, c_float, c_int32,
         P, P, P, P],
    ),
    "funlbm_wall_spring": (c_int32, [LATP, c_int32, P, P, P, c_float, P]),
    "funlbm_particle_forces": (c_int32, [c_int32, P, P, P, P, P, ctypes.c_double, ctypes.c_double, P]),
    "funlbm_particle_pose": (c_int32, [c_int32, P, P, P, P, c_int32, P]),
    "funlbm_ibm_spread": (c_int32, [LATP, c_int32, P, P, P, P, P, P, P, P, P, P]),
    "funlbm_ibm_spread_boxed": (c_int32, [LATP, c_int32, P, P, P, P, P, P, P, P, P, c_int32, P, P, P, P, c_int32,
                                          P]),
    "funlbm_surface_markers": (c_int32, [c_int32, P, ctypes.c_double, c_int32, P, P, P]),
    "funlbm_flow_stats": (c_int32, [LATP, P, P, P, P, P, P, P, c_int32, P, P]),
    "funlbm_flow_conservation": (c_int32, [LATP, P, P, P, P, P, P, P, c_int32, P, P]),
    "funlbm_case_bytes": (c_int32, []),
    "funlbm_case_pack": (c_int32, [LATP, c_int64, c_int64, c_int32, P]),
    "funlbm_batch_step": (c_int32, [c_int32] + [P] * 30 + [c_int32] * 4 + [P, P, P, c_int32, P, c_int32, P, P]),
}

_lib = None
_lock = threading.Lock()


def library_path() -> str:
    return LIB_PATH


def lib() -> ctypes.CDLL:
    """Loads the HIP library once; raises RuntimeError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build the HIP extension first "
                    "(python -c 'import __graft_entry__ as g; g.build()')"
                )
            h = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(h, name)
                fn.restype = res
                fn.argtypes = args
            if h.funlbm_abi_version() != ABI_VERSION:
                raise RuntimeError("libfunlbm_hip.so has an unexpected ABI version")
            _lib = h
    return _lib


def last_error() -> str:
    msg = lib().funlbm_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    """Maps a C status to the reference's exception types."""
    if rc == 0:
        return
    msg = f"{what}: {last_error()}"
    if rc == EUNSUPPORTED_BC:
        raise NotImplementedError(msg)
    if rc in (EUNSUPPORTED_LATTICE, EINVAL, ETOOBIG):
        raise ValueError(msg)
    raise RuntimeError(f"{msg} (status {rc})")


def ptr(t):
    """Raw device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return c_void_p(t.data_ptr())


def stream_handle(device=None):
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


def step_seed(base: int, steps_done: int) -> int:
    """The dither seed of the step after `steps_done` steps from `base`
    (step_seed in the kernels' common header)."""
    return (int(base) + int(steps_done) * 0x9E3779B9) & 0xFFFFFFFF


def state_plane(lat: Lattice) -> int:
    """Elements per direction plane of a fused-state buffer of `lat`
    (funlbm_state_bytes / (q * element size)); at least the node count."""
    nbytes = c_int64()
    check(lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)), "funlbm_state_bytes")
    esz = torch.tensor([], dtype=STORAGE_DTYPES[lat.storage]).element_size()
    return int(nbytes.value) // (int(lat.q) * esz)


_launch_state = threading.local()


@contextlib.contextmanager
def launches_off():
    """Within the block, call() launches nothing: the host-side bookkeeping
    of a step that a HIP graph replays runs without its kernels."""
    _launch_state.off = True
    try:
        yield
    finally:
        _launch_state.off = False


def call(name: str, *args) -> None:
    if getattr(_launch_state, "off", False):
        return
    check(getattr(lib(), name)(*args), name)

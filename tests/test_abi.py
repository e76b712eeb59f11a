"""The C-ABI library loads and exports every entry point include/funlbm.h
declares (no compute call; runs without a GPU)."""

import ctypes
import os
import re

import pytest

from funlbm_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "funlbm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(funlbm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "funlbm_fused_step" in names and "funlbm_ibm_spread" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(nat.library_path())
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert missing == []


def test_bindings_cover_header():
    assert sorted(nat.SIGNATURES) == declared_functions()


def test_abi_version_and_struct_size():
    lib = nat.lib()
    assert lib.funlbm_abi_version() == 4
    # 5 ints + 6 ints + 1 int + 3 floats + 3 floats + 18 floats + rounding + seed + delta
    assert ctypes.sizeof(nat.Lattice) == 4 * (5 + 6 + 1 + 3 + 3 + 18 + 3)


def test_rounding_modes_are_checked_on_the_host():
    lat = nat.Lattice()
    lat.q = 19
    lat.nx = lat.ny = lat.nz = 8
    lat.tau = 0.7
    lat.rho0 = 1.0
    lat.dev_scale = 1024.0
    for i in range(6):
        lat.bc[i] = nat.BC_WALL
    nbytes = ctypes.c_int64()
    lat.storage = nat.STORE_F32
    lat.rounding = nat.ROUND_STOCHASTIC
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == nat.EINVAL
    assert "F16/BF16" in nat.last_error()
    lat.rounding = 7
    lat.storage = nat.STORE_F16
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == nat.EINVAL
    for storage in (nat.STORE_F16, nat.STORE_BF16):
        lat.storage = storage
        lat.rounding = nat.ROUND_STOCHASTIC
        assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == 0


def test_step_seed_matches_the_kernels_rule():
    assert nat.step_seed(0, 0) == 0
    assert nat.step_seed(5, 1) == 5 + 0x9E3779B9
    assert nat.step_seed(0xFFFFFFFF, 2) == (0xFFFFFFFF + 2 * 0x9E3779B9) % (1 << 32)


def test_argument_errors_need_no_device():
    lat = nat.Lattice()
    lat.q = 9
    lat.nx = lat.ny = lat.nz = 4
    lat.tau = 0.8
    rc = nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(ctypes.c_int64()))
    assert rc == nat.EUNSUPPORTED_LATTICE
    assert "D3Q9" in nat.last_error()
    with pytest.raises(ValueError):
        nat.check(rc, "state_bytes")
    lat.q = 19
    for i in range(6):
        lat.bc[i] = 1400  # NON_EQUILIBRIUM: the reference defines it on the x faces only
    rc = nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(ctypes.c_int64()))
    assert rc == nat.EUNSUPPORTED_BC
    with pytest.raises(NotImplementedError):
        nat.check(rc, "state_bytes")
    for i in range(6):
        lat.bc[i] = nat.BC_WALL
    nbytes = ctypes.c_int64()
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == 0
    # one 256-node tile of 19 planes, plus a slack tile on either side
    assert nbytes.value == 19 * 3 * 256 * 4


def test_marker_generator_argument_errors_need_no_device():
    L = nat.lib()
    p = ctypes.c_void_p(16)  # never dereferenced: each call fails its host checks
    assert L.funlbm_surface_markers(0, None, 1.0, 0, None, None, None) == 0
    assert L.funlbm_surface_markers(-1, p, 1.0, 8, p, p, None) == nat.EINVAL
    assert L.funlbm_surface_markers(1, p, 1.0, -1, p, p, None) == nat.EINVAL
    assert L.funlbm_surface_markers(1, p, 0.0, 8, p, p, None) == nat.EINVAL
    assert L.funlbm_surface_markers(1, p, float("nan"), 8, p, p, None) == nat.EINVAL
    assert L.funlbm_surface_markers(1, None, 1.0, 8, p, p, None) == nat.EINVAL
    assert L.funlbm_surface_markers(1, p, 1.0, 8, None, p, None) == nat.EINVAL
    assert "NULL" in nat.last_error()


def test_marker_wrapper_rejects_bad_shapes_before_any_device_call():
    from funlbm_amd.particle.markers import surface_markers

    with pytest.raises(ValueError):
        surface_markers([(1.0, -2.0, 1.0)])
    with pytest.raises(ValueError):
        surface_markers([(1.0, 1.0, float("inf"))])
    with pytest.raises(ValueError):
        surface_markers([(1.0, 1.0, 1.0)], dx=0.0)
    with pytest.raises(RuntimeError):
        surface_markers([(1.0, 1.0, 1.0)], device="cpu")


@pytest.mark.parametrize("dims,storage", [((4, 4, 4), 0), ((7, 9, 11), 1), ((25, 25, 26), 1), ((16, 32, 32), 0),
                                          ((64, 64, 64), 0), ((256, 256, 256), 1), ((255, 255, 255), 0)])
def test_state_plane_has_slack_and_is_staggered(dims, storage):
    """A state buffer holds whole tiles of the node count plus two slack
    tiles (the z-face corner loads overreach by one node): more than n
    elements per direction, 4096-node tiles on large grids."""
    lat = nat.Lattice()
    lat.q = 27
    lat.nx, lat.ny, lat.nz = dims
    lat.tau = 0.8
    lat.storage = storage
    lat.dev_scale = 1024.0
    lat.rho0 = 1.0
    for i in range(6):
        lat.bc[i] = nat.BC_PERIODICAL
    n = dims[0] * dims[1] * dims[2]
    plane = nat.state_plane(lat)
    assert plane > n
    tile = 4096 if n >= 1 << 14 else 256
    assert plane == (-(-n // tile) + 2) * tile


def test_deviation_storage_needs_unit_rho0():
    """The 16-bit states keep f - w_k (rho0 = 1); another rho0 is refused,
    while float32/float64 ignore the field."""
    lat = nat.Lattice()
    lat.q = 19
    lat.nx = lat.ny = lat.nz = 8
    lat.tau = 0.8
    lat.dev_scale = 1024.0
    for i in range(6):
        lat.bc[i] = nat.BC_PERIODICAL
    nbytes = ctypes.c_int64()
    lat.storage, lat.rho0 = nat.STORE_F16, 1.5
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == nat.EINVAL
    assert "rho0" in nat.last_error()
    lat.rho0 = 1.0
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == 0
    lat.storage, lat.rho0 = nat.STORE_F32, 0.0
    assert nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(nbytes)) == 0


def test_boxed_interp_argument_errors_need_no_device():
    L = nat.lib()
    lat = nat.Lattice()
    lat.q = 19
    lat.nx = lat.ny = lat.nz = 8
    lat.tau = 0.8
    lat.rho0 = 1.0
    lat.dev_scale = 1024.0
    for i in range(6):
        lat.bc[i] = nat.BC_PERIODICAL
    p = ctypes.c_void_p(4096)
    lp = ctypes.byref(lat)
    args = lambda **kw: [kw.get(k, v) for k, v in (  # noqa: E731
        ("lat", lp), ("state", p), ("fol", p), ("rface", None), ("nm", 10), ("lx", p), ("pid", p),
        ("np", 1), ("ps", p), ("box", p), ("nch", 4), ("chunks", p), ("cache", p), ("cap", 1024),
        ("miss", p), ("dt", 1.0), ("check", 0), ("lu", p), ("lF", p), ("ctl", p), ("stream", None))]
    f = L.funlbm_ibm_interp_boxed
    assert f(*args(nm=0)) == 0  # nothing to do: no launch
    assert f(*args(nm=-1)) == nat.EINVAL
    assert f(*args(nch=-1)) == nat.EINVAL
    assert f(*args(cap=-1)) == nat.EINVAL
    assert f(*args(state=None)) == nat.EINVAL
    assert f(*args(box=None)) == nat.EINVAL
    assert f(*args(chunks=None)) == nat.EINVAL
    assert f(*args(miss=None)) == nat.EINVAL
    assert f(*args(cache=ctypes.c_void_p(4096 + 4))) == nat.EINVAL
    assert "aligned" in nat.last_error()
    assert f(*args(dt=0.0)) == nat.EINVAL


def test_noalloc_step_rejects_lattices_beyond_its_scratch():
    """The allocation-free logged step holds partials for 2^24 nodes; a
    larger lattice is refused on the host before any device call."""
    lat = nat.Lattice()
    lat.q = 19
    lat.nx = lat.ny = lat.nz = 257
    lat.tau = 0.6
    lat.rho0 = 1.0
    lat.dev_scale = 1024.0
    for i in range(6):
        lat.bc[i] = nat.BC_PERIODICAL
    rc = nat.lib().funlbm_fused_step_stats_noalloc(ctypes.byref(lat), None, None, None, None, None, None, None,
                                                   None, None, None)
    assert rc == nat.ETOOBIG
    assert "maximum" in nat.last_error()


@pytest.mark.parametrize("q,storage,dims,ok", [
    (19, 0, (1291, 1291, 1291), False),  # 2 151 685 171 nodes: past 2^31 - 1
    (19, 0, (1290, 1290, 1290), True),   # 2 146 689 000 nodes: below it
    (27, 3, (3, 3300, 3300), False),     # float64 D3Q27: a pull offset of 11e6 nodes x 216 B exceeds 2^31
    (27, 0, (3, 3300, 3300), True),      # the same grid in float32: 11e6 x 108 B fits
])
def test_size_limits_are_refused_on_the_host(q, storage, dims, ok):
    """The node count must stay below 2^31 and every pull source offset
    within 32 bits; both are checked on the host (funlbm_state_bytes needs no
    device) with FUNLBM_ETOOBIG."""
    lat = nat.Lattice()
    lat.q = q
    lat.nx, lat.ny, lat.nz = dims
    lat.tau = 0.8
    lat.storage = storage
    lat.dev_scale = 1024.0
    lat.rho0 = 1.0
    for i in range(6):
        lat.bc[i] = nat.BC_PERIODICAL
    out = ctypes.c_int64(0)
    rc = nat.lib().funlbm_state_bytes(ctypes.byref(lat), ctypes.byref(out))
    if ok:
        assert rc == 0 and out.value > 0
    else:
        assert rc == nat.ETOOBIG, nat.last_error()

"""Parity of the HIP fluid kernels with the CPU oracle (needs a GPU).

Streaming and bounce-back are permutations and must be bit-exact; the float
arithmetic must agree with the oracle within 1e-5 relative L-inf after the
step counts below (the reference's own criterion for 1000 steps)."""

import ctypes

import numpy as np
import pytest
import torch

from oracle import lbm_oracle as orc
from tests.helpers import deviation_round, deviation_round_sr, flow_json, random_state, rel_linf, step_seed

pytestmark = pytest.mark.gpu

P, W, MW = orc.PERIODICAL, orc.WALL, orc.WALL_WITH_SPEED
BC_SETS = {
    "periodic": [P] * 6,
    "wall": [W] * 6,
    "channel": [W, W, P, P, P, P],
    "mixed": [P, W, W, P, W, P],
}


def make_flow(q, size, mu, bcs, uw=None, **ext):
    from funlbm_amd.flow import FlowConfig, create_flow

    cfg = FlowConfig().from_json(flow_json(q, size, mu, bcs, uw=uw, **ext))
    fl = create_flow(cfg, device="cuda")
    fl.init()
    return fl


def oracle_state(q, size, mu, bcs, uw=None, dtype=np.float32, **kw):
    nodes = tuple(int(s) + 1 for s in size)
    return orc.init_flow(q, nodes, mu=mu, bc=bcs, dtype=dtype, uw=uw, **kw)


@pytest.mark.parametrize("q", [13, 15, 19, 27])
@pytest.mark.parametrize("bcname", list(BC_SETS))
def test_stream_bit_exact(q, bcname):
    size = (4, 5, 6)
    bcs = BC_SETS[bcname]
    fl = make_flow(q, size, 0.1, bcs)
    st = oracle_state(q, size, 0.1, bcs)
    f0 = random_state(q, st.nodes, seed=q * 7 + len(bcname))
    # unique values make every misplaced population visible
    f0 = f0 + np.arange(f0.size, dtype=np.float32).reshape(f0.shape) * np.float32(1e-3)
    fl.f = torch.from_numpy(f0)
    st.f = f0.copy()
    fl.f_stream()
    orc.f_stream(st)
    got = fl.f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), st.f.view(np.uint32))


@pytest.mark.parametrize("q", [15, 19])
def test_stream_moving_wall(q):
    size = (5, 4, 6)
    uw = [None, None, None, None, None, [0.05, 0.01, 0.0]]
    bcs = [W, W, P, P, W, MW]
    fl = make_flow(q, size, 0.1, bcs, uw=uw)
    st = oracle_state(q, size, 0.1, bcs, uw=uw)
    f0 = random_state(q, st.nodes, seed=3)
    fl.f = torch.from_numpy(f0)
    st.f = f0.copy()
    fl.update_u_rou()
    orc.update_u_rou(st)
    fl.f_stream()
    orc.f_stream(st)
    assert rel_linf(fl.f.cpu().numpy(), st.f) < 1e-6


@pytest.mark.parametrize("q", [13, 15, 19, 27])
def test_collide_moments_forcing(q):
    size = (5, 6, 7)
    bcs = BC_SETS["mixed"]
    fl = make_flow(q, size, 0.05, bcs)
    st = oracle_state(q, size, 0.05, bcs)
    f0 = random_state(q, st.nodes, seed=11 + q)
    rng = np.random.default_rng(q)
    fol = (1e-3 * rng.standard_normal(st.nodes + (3,))).astype(np.float32)
    fl.f = torch.from_numpy(f0)
    fl.FOL = torch.from_numpy(fol)
    st.f = f0.copy()
    st.FOL = fol.copy()
    fl.update_u_rou()
    orc.update_u_rou(st)
    assert rel_linf(fl.u.cpu().numpy(), st.u) < 1e-6
    assert rel_linf(fl.rou.cpu().numpy(), st.rou) < 1e-6
    fl.cul_equ()
    orc.cul_equ(st)
    assert rel_linf(fl.feq.cpu().numpy(), st.feq) < 1e-6
    assert rel_linf(fl.f.cpu().numpy(), st.f) < 1e-6
    fl.cul_equ2()
    orc.cul_equ2(st)
    assert rel_linf(fl.f.cpu().numpy(), st.f) < 1e-6


def run_solver(q, size, mu, bcs, steps, f0=None, fused=True, storage="f32", **flow_ext):
    from funlbm_amd.lbm import LBMD3, Config

    cfg = Config().from_json({"flow": flow_json(q, size, mu, bcs, storage=storage, **flow_ext),
                              "log_every": 0, "fused": fused})
    lbm = LBMD3(config=cfg)
    lbm.init()
    if f0 is not None:
        lbm.flow.f = torch.from_numpy(f0)
        lbm.flow.update_u_rou()
    for s in range(1, steps + 1):
        lbm.step = s
        lbm.run_step()
    return lbm


def run_oracle(q, size, mu, bcs, steps, f0=None, body_force=None):
    nodes = tuple(int(s) + 1 for s in size)
    st = orc.init_flow(q, nodes, mu=mu, bc=bcs, body_force=body_force)
    orc.init_solver(st)
    if f0 is not None:
        st.f = f0.copy()
        orc.update_u_rou(st)
    for _ in range(steps):
        orc.fluid_step(st)
    return st


@pytest.mark.parametrize("q", [13, 15, 19, 27])
@pytest.mark.parametrize("bcname", ["periodic", "channel", "wall"])
def test_fused_steps_match_oracle(q, bcname):
    size = (6, 7, 9)
    bcs = BC_SETS[bcname]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=5)
    lbm = run_solver(q, size, 0.05, bcs, 50, f0=f0)
    st = run_oracle(q, size, 0.05, bcs, 50, f0=f0)
    ef = rel_linf(lbm.flow.f.cpu().numpy(), st.f)
    er = rel_linf(lbm.flow.rou.cpu().numpy(), st.rou)
    # u is a difference of populations that has decayed to ~1e-4 here: bound
    # it by the float32 rounding of f (the oracle's own float32 and float64
    # runs differ by ~1e-3 of max|u| on this state)
    eu = rel_linf(lbm.flow.u.cpu().numpy(), st.u)
    assert ef < 1e-5 and er < 1e-5 and eu < 1e-2, (ef, er, eu)


@pytest.mark.parametrize("q", [19, 27])
def test_fused_equals_explicit_bitwise(q):
    size = (5, 6, 7)
    bcs = BC_SETS["mixed"]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=9)
    a = run_solver(q, size, 0.05, bcs, 20, f0=f0, fused=True)
    b = run_solver(q, size, 0.05, bcs, 20, f0=f0, fused=False)
    assert torch.equal(a.flow.f, b.flow.f)
    assert torch.equal(a.flow.u, b.flow.u)


def test_poiseuille_config1_1000_steps():
    """Channel of the Poiseuille case: 32 x 32 x 64 nodes, walls at x = 0 and
    x = 31, periodic y and z, tau = 0.8, body force g_z = 1e-6 (FOL extension)."""
    size = (31, 31, 63)
    bcs = BC_SETS["channel"]
    mu = 0.1
    g = [0.0, 0.0, 1e-6]
    lbm = run_solver(19, size, mu, bcs, 1000, body_force=g)
    st = run_oracle(19, size, mu, bcs, 1000, body_force=g)
    ef = rel_linf(lbm.flow.f.cpu().numpy(), st.f)
    eu = rel_linf(lbm.flow.u.cpu().numpy(), st.u)
    assert ef < 1e-5 and eu < 1e-4, (ef, eu)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_flow_stats_from_state(storage):
    """min/mean/max read straight from the fused state equal those of the
    materialised fields (the run loop's per-step statistics)."""
    size = (9, 11, 13)
    bcs = BC_SETS["channel"]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(19, nodes, seed=2)
    lbm = run_solver(19, size, 0.05, bcs, 7, f0=f0, storage=storage, body_force=[0.0, 1e-5, 0.0])
    fast = lbm.flow.stats().cpu().numpy()
    want = []
    for t in (lbm.flow.f, lbm.flow.u, lbm.flow.rou):
        want += [t.min().item(), t.double().mean().item(), t.max().item()]
    want = np.array(want)
    assert np.array_equal(fast[[0, 2, 3, 5, 6, 8]], want[[0, 2, 3, 5, 6, 8]])
    assert np.allclose(fast[[4, 7]], want[[4, 7]], rtol=1e-9, atol=1e-15)
    # the mean of f is the mean density over Q (a node's populations sum to
    # its density up to the float32 rounding of that sum)
    assert abs(fast[1] - want[7] / lbm.flow.q) <= 1e-12 * abs(fast[1])
    assert abs(fast[1] - want[1]) <= 1e-6 * abs(want[1])


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_conservation_from_state(storage):
    """Global mass and momentum read from the fused state in double equal the
    float64 sums over the materialised rho and u (1e-12: the two sum in
    different orders)."""
    size = (9, 11, 13)
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(19, nodes, seed=4)
    lbm = run_solver(19, size, 0.05, BC_SETS["channel"], 7, f0=f0, storage=storage,
                     body_force=[0.0, 1e-5, 0.0])
    fast = lbm.flow.conservation().cpu().numpy()
    rho = lbm.flow.rou.double().cpu().numpy().reshape(-1)
    u = lbm.flow.u.double().cpu().numpy().reshape(-1, 3)
    want = np.array([rho.sum()] + [(rho * u[:, d]).sum() for d in range(3)])
    assert abs(fast[0] - want[0]) <= 1e-12 * abs(want[0])
    scale = (rho * np.abs(u).max(axis=1)).sum()
    assert np.all(np.abs(fast[1:] - want[1:]) <= 1e-12 * scale)
    # the materialised fields give the same values through the fallback
    again = lbm.flow.conservation().cpu().numpy()
    assert np.allclose(again, want, rtol=1e-12, atol=1e-12 * scale)


@pytest.mark.parametrize("q", [19, 27])
def test_conservation_tracks_the_oracle(q):
    """Over 50 steps of a random state in a box walled on every face, the
    global mass and momentum read from the fused state follow the oracle's
    (its rho and u in float64, 1e-12: the float32 states are bit-identical
    and only the summation order differs).  The mass stays within 1e-5 of
    its value after the first step (1.1e-6 was measured for D3Q19 over these
    steps, less for D3Q27)."""
    size = (15, 13, 17)
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=5)
    bcs = BC_SETS["wall"]
    lbm = run_solver(q, size, 0.05, bcs, 0, f0=f0)
    c0 = None
    for s in range(1, 51):
        lbm.step = s
        lbm.run_step()
        c = lbm.flow.conservation().cpu().numpy()
        c0 = c if c0 is None else c0
    ref = run_oracle(q, size, 0.05, bcs, 50, f0=f0)
    rho = ref.rou.astype(np.float64).reshape(-1)
    u = ref.u.astype(np.float64).reshape(-1, 3)
    want = np.array([rho.sum()] + [(rho * u[:, d]).sum() for d in range(3)])
    scale = (rho * np.abs(u).max(axis=1)).sum()
    assert abs(c[0] - want[0]) <= 1e-12 * want[0]
    assert np.all(np.abs(c[1:] - want[1:]) <= 1e-12 * scale)
    assert abs(c[0] - c0[0]) <= 1e-5 * c0[0]


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("q", [19, 27])
def test_step_statistics_equal_separate_pass(q, storage):
    """The statistics that a logged fused step reduces in its own pass
    (funlbm_fused_step_stats) equal those of the separate pass over the state
    (funlbm_flow_stats): min and max exactly, means to 1e-12 (the two group
    the nodes differently); the state the step writes is the plain step's."""
    size = (9, 11, 13)
    bcs = BC_SETS["channel"]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=5)
    a = run_solver(q, size, 0.05, bcs, 5, f0=f0, storage=storage, body_force=[0.0, 1e-5, 0.0])
    b = run_solver(q, size, 0.05, bcs, 5, f0=f0, storage=storage, body_force=[0.0, 1e-5, 0.0])
    a.flow.stats_next = True
    for lbm in (a, b):
        lbm.step = 6
        lbm.run_step()
    fused = a.flow.stats().cpu().numpy().copy()
    a.flow._stats_valid = False
    sep = a.flow.stats().cpu().numpy()
    assert np.array_equal(fused[[0, 2, 3, 5, 6, 8]], sep[[0, 2, 3, 5, 6, 8]])
    assert np.allclose(fused[[1, 4, 7]], sep[[1, 4, 7]], rtol=1e-12, atol=1e-15)
    assert torch.equal(a.flow.f, b.flow.f)


@pytest.mark.parametrize("size", [(11, 9, 13), (9, 10, 15)])
@pytest.mark.parametrize("q", [13, 15, 19, 27])
def test_fused_bitwise_equal_to_oracle(q, size):
    """Every quotient of the kernels is the correctly rounded one (three-
    operation division by a rounded reciprocal) and every sum runs in the
    oracle's order, so the float32 state is bit-identical to the oracle's
    after many steps, body force and walls included.  With 16 nodes along z
    the interior runs through the four-nodes-per-thread vector path."""
    bcs = BC_SETS["channel"]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=4)
    g = [1e-5, 0.0, 2e-5]
    lbm = run_solver(q, size, 0.07, bcs, 40, f0=f0, body_force=g)
    st = run_oracle(q, size, 0.07, bcs, 40, f0=f0, body_force=g)
    got = lbm.flow.f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), st.f.view(np.uint32)), rel_linf(got, st.f)


@pytest.mark.parametrize("nodes", [(1, 4, 5), (4, 1, 5), (4, 5, 1), (2, 2, 2), (1, 1, 1), (3, 3, 2), (5, 3, 1)])
@pytest.mark.parametrize("bcname", ["periodic", "wall", "mixed"])
@pytest.mark.parametrize("q", [13, 15, 19, 27])
def test_fused_degenerate_lattices_bitwise(q, bcname, nodes):
    """One or two nodes along an axis, down to a single node: both faces of
    such an axis act on the same node (each for the directions entering
    through it), rows of one node wrap onto themselves, and fewer nodes than
    a wavefront hold the state.  The float32 fused step stays bit-identical
    to the oracle's np.roll and face overrides."""
    bcs = BC_SETS[bcname]
    size = tuple(n - 1 for n in nodes)
    f0 = random_state(q, nodes, seed=9)
    g = [1e-5, -2e-5, 3e-5]
    lbm = run_solver(q, size, 0.07, bcs, 12, f0=f0, body_force=g)
    st = run_oracle(q, size, 0.07, bcs, 12, f0=f0, body_force=g)
    got = lbm.flow.f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), st.f.view(np.uint32)), rel_linf(got, st.f)


@pytest.mark.parametrize("nodes", [(4, 5, 1), (2, 2, 2), (1, 1, 1)])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_deviation_storage_degenerate_lattices(storage, nodes):
    """The 16-bit fused step on lattices smaller than a wavefront (the
    partial-wave stores) against the oracle with the storage's round trip,
    within two storage ulps as in the test below."""
    q, bcs = 19, BC_SETS["mixed"]
    size = tuple(n - 1 for n in nodes)
    f0 = random_state(q, nodes, seed=10)
    g = [2e-5, 0.0, -1e-5]
    lbm = run_solver(q, size, 0.07, bcs, 20, f0=f0, storage=storage, body_force=g, rounding="nearest")
    st = run_oracle_rounded(q, size, 0.07, bcs, 20, f0, storage, body_force=g)
    got = lbm.flow.f.cpu().numpy()
    bits = {"f16": 11, "bf16": 8}[storage]
    dev = float(np.max(np.abs(f0 - st.lat.w.astype(np.float32))))
    tol = 2.0 * 2.0 ** -bits * dev / float(np.max(np.abs(st.f)))
    assert rel_linf(got, st.f) < tol, (rel_linf(got, st.f), tol)


def run_oracle_rounded(q, size, mu, bcs, steps, f0, storage, body_force=None, scale=1024.0,
                       rounding="nearest", seed=0):
    """The oracle's steps with the state storage's rounding applied to the
    post-collision populations, where the fused path stores them: the first
    collision's store (the packing of the explicit state) rounds to nearest,
    the store of step i + 1's collision by the solver's i-th fused step
    (i >= 1) rounds as `rounding` with the dither of step_seed(seed, i - 1)."""
    nodes = tuple(int(s) + 1 for s in size)
    st = orc.init_flow(q, nodes, mu=mu, bc=bcs, body_force=body_force)
    orc.init_solver(st)
    st.f = f0.copy()
    orc.update_u_rou(st)
    for i in range(steps):
        orc.cul_equ(st)
        if rounding == "stochastic" and i > 0:
            st.f = deviation_round_sr(st.f, st.lat.w, storage, step_seed(seed, i - 1), scale)
        else:
            st.f = deviation_round(st.f, st.lat.w, storage, scale)
        orc.f_stream(st)
        orc.update_u_rou(st)
        st.FOL = np.zeros_like(st.FOL)
        orc.cul_equ2(st)
        orc.update_u_rou(st)
    return st


@pytest.mark.parametrize("rounding", ["nearest", "stochastic"])
@pytest.mark.parametrize("size", [(9, 10, 15), (11, 9, 13), (11, 9, 149)])
@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("q", [19, 27])
def test_deviation_storage_matches_rounded_oracle(q, storage, size, rounding):
    """f16/bf16 state: the fused path against the oracle with the storage's
    round trip applied after every collision (to nearest, or stochastically
    with the solver's dither words).  The 16-bit paths compute on the scaled
    deviations with reciprocals and FMAs (float32 rounding, far below the
    storage's 2^-11 or 2^-8), so a rounding of the state can land on the other
    side now and then: after 30 steps f stays within two storage ulps of the
    largest deviation |f - w| of the run, the initial one (relative to
    max|f|).  The 12 x 10 x 150 lattice takes 4096-node state tiles, whose
    inner waves pull through LDS (the smaller ones straddle a tile edge in
    every wave)."""
    bcs = BC_SETS["channel"]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=6)
    g = [2e-5, 0.0, -1e-5]
    lbm = run_solver(q, size, 0.07, bcs, 30, f0=f0, storage=storage, body_force=g, rounding=rounding, seed=77)
    st = run_oracle_rounded(q, size, 0.07, bcs, 30, f0, storage, body_force=g, rounding=rounding, seed=77)
    got = lbm.flow.f.cpu().numpy()
    err = rel_linf(got, st.f)
    bits = {"f16": 11, "bf16": 8}[storage]
    dev = float(np.max(np.abs(f0 - st.lat.w.astype(np.float32))))
    tol = 2.0 * 2.0 ** -bits * dev / float(np.max(np.abs(st.f)))
    print(storage, q, size, "rel L-inf", err, "tolerance", tol)
    assert err < tol, (err, tol)


@pytest.mark.parametrize("q", [19, 27])
def test_neq_faces_match_oracle_bitwise(q):
    """NON_EQUILIBRIUM inlet and outlet (explicit path, chosen by LBMD3 for
    such faces): the duct-profile inlet, the copied outlet and the
    feq + (f - feq) streaming rule, bit for bit with the oracle over 25 steps."""
    N = orc.NON_EQUILIBRIUM
    size = (8, 7, 9)
    bcs = [N, N, W, W, P, P]
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=12)
    from funlbm_amd.lbm import LBMD3, Config

    flow = flow_json(q, size, 0.05, bcs)
    flow["param"]["Re"] = 40.0
    lbm = LBMD3(config=Config().from_json({"flow": flow, "log_every": 0}))
    lbm.init()
    lbm.flow.f = torch.from_numpy(f0)
    lbm.flow.update_u_rou()
    st = orc.init_flow(q, nodes, mu=0.05, bc=bcs, Re=40.0)
    orc.init_solver(st)
    st.f = f0.copy()
    orc.update_u_rou(st)
    for s in range(1, 26):
        lbm.step = s
        lbm.run_step()
        orc.fluid_step(st)
    got = lbm.flow.f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), st.f.view(np.uint32)), rel_linf(got, st.f)
    u = lbm.flow.u.cpu().numpy()
    assert np.array_equal(u.view(np.uint32), st.u.view(np.uint32)), rel_linf(u, st.u)
    assert float(np.abs(u[0, :, :, 0]).max()) > 0  # the inlet profile is there


def test_neq_faces_refused_by_fused_entry_points():
    from funlbm_amd import _native as nat

    N = orc.NON_EQUILIBRIUM
    fl = make_flow(19, (6, 5, 7), 0.05, [N, N, W, W, P, P])
    with pytest.raises(NotImplementedError):
        fl.begin_step()
    rc = nat.lib().funlbm_fused_step(ctypes.byref(fl.lat), None, None, None, None, None, None, None, None, None)
    assert rc == nat.EUNSUPPORTED_BC


MOVING_SETS = {
    # moving walls on the output (x) face and the top (z) face: the x = X,
    # z = Z edge takes both rules, the top one winning for e_z = -1
    "output_top": ([W, MW, P, P, W, MW], [None, [0.0, 0.02, -0.01], None, None, None, [0.05, 0.01, 0.0]]),
    # moving back (y) and bottom (z) faces, walls opposite
    "back_bottom": ([P, P, MW, W, MW, W], [None, None, [0.03, 0.0, 0.01], None, [-0.02, 0.04, 0.0], None]),
}


@pytest.mark.parametrize("faces", list(MOVING_SETS))
@pytest.mark.parametrize("q", [15, 19, 27])
def test_fused_moving_wall_bitwise_equal_to_oracle(q, faces):
    """WALL_WITH_SPEED on the fused path (the default run_step): the x/y-face
    rule through the per-node source, the z-face fix-up and the face density
    of the previous step that both read, bit for bit with the oracle's
    f_stream (src/funlbm/flow/d3/base.py:204-213,305-321) after 120 steps,
    the corner nodes of two moving faces included."""
    bcs, uw = MOVING_SETS[faces]
    size = (9, 10, 13)
    nodes = tuple(s + 1 for s in size)
    f0 = random_state(q, nodes, seed=21)
    lbm = run_solver(q, size, 0.06, bcs, 120, f0=f0, uw=uw)
    assert lbm.flow.has_moving_wall
    st = orc.init_flow(q, nodes, mu=0.06, bc=bcs, uw=uw)
    orc.init_solver(st)
    st.f = f0.copy()
    orc.update_u_rou(st)
    for _ in range(120):
        orc.fluid_step(st)
    got = lbm.flow.f.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), st.f.view(np.uint32)), rel_linf(got, st.f)
    rho = lbm.flow.rou.cpu().numpy()
    assert np.array_equal(rho.view(np.uint32), st.rou.view(np.uint32))
    # the walls drive a flow: the test is not passing on a state at rest
    assert float(np.abs(st.u).max()) > 1e-3

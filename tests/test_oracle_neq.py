"""NON_EQUILIBRIUM inlet/outlet faces of the oracle (CPU), pinned by the duct
flow's mathematics and by per-node brute force; the product's duct profile
is checked against the oracle's."""

import numpy as np
import pytest

from oracle import lbm_oracle as orc
from tests.helpers import random_state


def test_duct_series_is_the_duct_flow():
    """The series solves the laminar duct problem: zero on the four walls,
    Laplacian -2/ha^2 inside (so the parabola 1 - (y/ha)^2 carries the
    pressure gradient), and the centre of a square duct at 0.5894 of the
    plane-channel maximum (the known 0.2947 G a^2/mu with a the half side)."""
    ha, hb = 3.0, 5.0
    for t in np.linspace(-0.9, 0.9, 7):
        assert abs(orc.duct_series(ha, t * hb, ha, hb)) < 1e-6
        assert abs(orc.duct_series(-ha, t * hb, ha, hb)) < 1e-6
        assert abs(orc.duct_series(t * ha, hb, ha, hb)) < 1e-5
        assert abs(orc.duct_series(t * ha, -hb, ha, hb)) < 1e-5
    h = 1e-3
    for y, z in [(0.0, 0.0), (1.2, -2.0), (-2.1, 3.3)]:
        lap = (orc.duct_series(y + h, z, ha, hb) + orc.duct_series(y - h, z, ha, hb)
               + orc.duct_series(y, z + h, ha, hb) + orc.duct_series(y, z - h, ha, hb)
               - 4 * orc.duct_series(y, z, ha, hb)) / h ** 2
        assert abs(lap - (-2.0 / ha ** 2)) < 1e-3 * (2.0 / ha ** 2), (y, z, lap)
    assert abs(orc.duct_series(0.0, 0.0, 1.0, 1.0) - 0.5894) < 1e-4


@pytest.mark.parametrize("a,b", [(5, 7), (16, 16), (9, 9), (32, 20)])
def test_product_profile_equals_oracle(a, b):
    from funlbm_amd.flow.d3.neq import duct_profile

    assert np.array_equal(duct_profile(a, b), orc.duct_profile(a, b))
    # the node grid is symmetric, so the rounded profile is too
    p = orc.duct_profile(a, b)
    assert np.array_equal(p, p[::-1, :]) and np.array_equal(p, p[:, ::-1])


@pytest.mark.parametrize("q", [15, 19, 27])
def test_neq_stream_rule_brute_force(q):
    """Every direction of the x = 0 (x = X) plane becomes feq + (streamed f -
    feq) of the plane x = 1 (x = X-1), except where a y or z face rule takes
    over; the streamed value there is the periodic roll."""
    nodes = (5, 4, 6)
    bc = [orc.NON_EQUILIBRIUM, orc.NON_EQUILIBRIUM, orc.PERIODICAL, orc.PERIODICAL, orc.WALL, orc.WALL]
    st = orc.init_flow(q, nodes, mu=0.05, bc=bc)
    st.f = random_state(q, nodes, seed=3)
    rng = np.random.default_rng(1)
    st.feq = (st.f * (1 + 0.01 * rng.standard_normal(st.f.shape))).astype(np.float32)
    fcopy = st.f.copy()
    orc.f_stream(st)
    X, Y, Z = nodes
    e = st.lat.e
    checked = 0
    for x0, xi in ((0, 1), (X - 1, X - 2)):
        for y in range(Y):
            for z in range(Z):
                for k in range(q):
                    ey, ez = e[k, 1], e[k, 2]
                    if (y == 0 and ey == 1) or (y == Y - 1 and ey == -1) or (z == 0 and ez == 1) or \
                            (z == Z - 1 and ez == -1):
                        continue
                    src = fcopy[(xi - e[k, 0]) % X, (y - ey) % Y, (z - ez) % Z, k]
                    want = st.feq[x0, y, z, k] + (src - st.feq[xi, y, z, k])
                    assert st.f[x0, y, z, k] == np.float32(want), (x0, y, z, k)
                    checked += 1
    assert checked > 0
    # the interior is the plain roll
    for k in range(q):
        assert st.f[2, 1, 2, k] == fcopy[(2 - e[k, 0]) % X, (1 - e[k, 1]) % Y, (2 - e[k, 2]) % Z, k]


def test_neq_moment_overrides():
    """Inlet: u_x = uw * profile with uw = float32(Re mu) / max(rho) / min(Y, Z)
    in float32, u_y = u_z = 0, rho copied from x = 1; outlet: rho and u copied
    from x = X - 2 (after the face zeroing)."""
    q, nodes = 19, (6, 5, 7)
    bc = [orc.NON_EQUILIBRIUM, orc.NON_EQUILIBRIUM, orc.WALL, orc.WALL, orc.WALL, orc.WALL]
    st = orc.init_flow(q, nodes, mu=0.04, bc=bc, Re=25.0)
    st.f = random_state(q, nodes, seed=8, amp=3e-2)
    rho, u = orc.moments(st.lat, st.f, st.fol_eff(), orc.face_mask(st.nodes, st.uzero_mask))
    orc.update_u_rou(st)
    uw = np.float32(np.float32(25.0 * 0.04) / rho.max()) / np.float32(5)
    prof = orc.duct_profile(5, 7)
    assert np.array_equal(st.u[0, :, :, 0], (np.float64(uw) * prof).astype(np.float32))
    assert np.all(st.u[0, :, :, 1:] == 0)
    assert np.array_equal(st.rou[0], rho[1])
    assert np.array_equal(st.rou[-1], rho[-2])
    assert np.array_equal(st.u[-1], u[-2])
    # p keeps rho / 3 of the moments, before the face copies
    assert np.array_equal(st.p[0], rho[0] / np.float32(3))

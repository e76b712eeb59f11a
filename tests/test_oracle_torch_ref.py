"""The torch-CPU restatement of the reference's fluid step (the timed CPU
baseline) against the numpy oracle: same state after several steps, to
float32 rounding (torch's products and norm sum in their own order)."""

import numpy as np
import pytest
import torch

from oracle import lbm_oracle as orc
from oracle.torch_ref import TorchFlow
from tests.helpers import random_state

P, W = orc.PERIODICAL, orc.WALL


@pytest.mark.parametrize("bcs", [[P] * 6, [W] * 6, [P, W, W, P, W, P]])
@pytest.mark.parametrize("q", [13, 15, 19, 27])
def test_torch_restatement_matches_oracle(q, bcs):
    nodes = (6, 7, 8)
    f0 = random_state(q, nodes, seed=q + len(set(bcs)))
    st = orc.init_flow(q, nodes, mu=0.05, bc=bcs)
    st.f = f0.copy()
    orc.update_u_rou(st)
    tf = TorchFlow(q, nodes, tau=np.float32(3 * 0.05 + 0.5), bc=bcs)
    tf.f = torch.from_numpy(f0.copy())
    tf.update_u_rou()
    for _ in range(6):
        orc.fluid_step(st)
        tf.step()
    f = tf.f.numpy()
    assert np.max(np.abs(f - st.f)) / np.max(np.abs(st.f)) < 2e-6
    u = tf.u.numpy()
    assert np.max(np.abs(u - st.u)) < 1e-6 * np.max(np.abs(st.f))
    assert np.array_equal(u[0], np.zeros_like(u[0])) and np.array_equal(u[:, :, -1], np.zeros_like(u[:, :, -1]))


def test_torch_restatement_streams_like_the_oracle():
    """Streaming alone is a permutation: bit-exact with the oracle."""
    for q in (15, 27):
        nodes = (4, 5, 6)
        bcs = [P, W, W, P, W, P]
        f0 = random_state(q, nodes, seed=1) + np.arange(np.prod(nodes) * q, dtype=np.float32).reshape(
            nodes + (q,)) * np.float32(1e-3)
        st = orc.init_flow(q, nodes, mu=0.05, bc=bcs)
        st.f = f0.copy()
        tf = TorchFlow(q, nodes, tau=0.65, bc=bcs)
        tf.f = torch.from_numpy(f0.copy())
        orc.f_stream(st)
        tf.f_stream()
        assert np.array_equal(tf.f.numpy().view(np.uint32), st.f.view(np.uint32))

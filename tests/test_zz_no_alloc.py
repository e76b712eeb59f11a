"""The allocation-free logged fused step (needs a GPU): its statistics and
state equal those of funlbm_fused_step_stats with caller-provided partials,
bit for bit, on two calls in a row, the second on a smaller lattice."""

import ctypes

import numpy as np
import pytest
import torch

from funlbm_amd import _native as nat
from tests.helpers import flow_json, random_state

pytestmark = pytest.mark.gpu

W, P = nat.BC_WALL, nat.BC_PERIODICAL


def stepped_flow(q, size, storage, steps):
    from funlbm_amd.lbm import LBMD3, Config

    cfg = Config().from_json({"flow": flow_json(q, size, 0.05, [W, W, P, P, P, P], storage=storage,
                                                body_force=[0.0, 1e-5, 2e-5]), "log_every": 0})
    lbm = LBMD3(config=cfg)
    lbm.init()
    lbm.flow.f = torch.from_numpy(random_state(q, tuple(s + 1 for s in size), seed=q))
    lbm.flow.update_u_rou()
    for s in range(1, steps + 1):
        lbm.step = s
        lbm.run_step()
    return lbm.flow


def both(fl):
    """One logged step from the flow's current state, through each entry
    point, into fresh destination buffers: (statistics, states)."""
    lib = nat.lib()
    src = fl._state[fl._cur]
    nb = -(-fl.n // 256)
    nb += min(nb, 256)
    partials = torch.empty(9 * nb, device=fl.device, dtype=torch.float64)
    outs, dsts = [], []
    for noalloc in (False, True):
        dst = torch.full_like(src, float("nan"))
        out = torch.full((9,), float("nan"), device=fl.device, dtype=torch.float64)
        args = [ctypes.byref(fl.lat), nat.ptr(src), nat.ptr(dst), fl._fol_ptr(fl._fc), None, None, None, None, None]
        if noalloc:
            rc = lib.funlbm_fused_step_stats_noalloc(*args, nat.ptr(out), nat.stream_handle(fl.device))
        else:
            rc = lib.funlbm_fused_step_stats(*args, nat.ptr(partials), nb, nat.ptr(out),
                                             nat.stream_handle(fl.device))
        assert rc == 0, nat.last_error()
        outs.append(out)
        dsts.append(dst)
    torch.cuda.synchronize()
    return outs, dsts


def test_noalloc_logged_step_equals_the_allocating_one():
    big = stepped_flow(19, (40, 35, 47), "f32", 3)  # 59 904 nodes: 234 partials + 234
    small = stepped_flow(27, (12, 9, 14), "f16", 3)  # 1 950 nodes: 8 + 8
    for fl in (big, small):
        (want, got), (d_want, d_got) = both(fl)
        assert torch.isfinite(want).all()
        assert torch.equal(want, got), (want, got)
        n = fl.n
        # the state's nodes (the slack tiles hold no data)
        from tests.helpers import state_index
        idx = torch.from_numpy(state_index(fl.q, n)).to(fl.device)
        a = d_want.view(-1)[idx.view(-1)]
        b = d_got.view(-1)[idx.view(-1)]
        assert torch.equal(a, b)

"""The test suite's model of the f16/bf16 state storage, checked against
torch's own float32 -> half / bfloat16 conversions (round to nearest even)."""

import numpy as np
import torch

from oracle import lbm_oracle as orc
from tests.helpers import deviation_round


def test_deviation_round_matches_torch_conversions():
    rng = np.random.default_rng(0)
    lat = orc.Lattice(19, np.float32)
    f = (lat.w * (1 + 1e-2 * rng.standard_normal((64, 19)))).astype(np.float32)
    # ties and near-ties of the 16-bit grids as well
    f[0] = lat.w + np.float32(2.0 ** -21)
    f[1] = lat.w - np.float32(3 * 2.0 ** -22)
    wr = lat.w.astype(np.float32)
    s = np.float32(1024.0)
    d = torch.from_numpy(((f - wr) * s).astype(np.float32))
    for storage, dt in (("f16", torch.float16), ("bf16", torch.bfloat16)):
        want = (d.to(dt).float().numpy() * (np.float32(1) / s) + wr).astype(np.float32)
        got = deviation_round(f, lat.w, storage, 1024.0)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), storage
    # the round trip moves values, so a checker that skipped it would differ
    assert not np.array_equal(deviation_round(f, lat.w, "bf16"), f)

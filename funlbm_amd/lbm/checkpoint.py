"""Checkpoints in the reference's layout (src/funlbm/lbm/base.py:198-227,
src/funlbm/flow/base.py:208-217, src/funlbm/particle/base/base.py:190-214,
src/funlbm/particle/swarm.py:34-41).

The reference writes an HDF5 file with the datasets `step`, `flow/f`
([X, Y, Z, Q]) and `particle/particle_{i:06d}/{cu, cx, cF, cw, lx, lF, lu,
lrou}`.  This package writes the same datasets, under the same paths as keys,
into a compressed NumPy archive (h5py is not a dependency), plus
`particle/particle_{i:06d}/state`: the particle's float64 state row
(centre, velocities, force, torque, rotation vector and matrix, mass, inertia,
density, gravity), which restores its rotation exactly, `flow/FOL`, the
force density of the last spreading, which the next step's first moment
update reads, and `flow/steps_done`, the fused steps taken, from which the
stochastic rounding of a 16-bit state draws its next dither.  Loading sets FOL (zero if the archive has none, as the
reference's load leaves it), then f, recomputes the moments and restores the
particle datasets, so a resumed run equals the uninterrupted one.
"""

import os

import numpy as np
import torch

PARTICLE_KEYS = ("cu", "cx", "cF", "cw", "lx", "lF", "lu", "lrou")


def _particle_group(i: int) -> str:
    return f"particle/particle_{str(i).zfill(6)}"


def dump(lbm, path: str) -> str:
    """Writes the checkpoint of `lbm` to `path` (.npz); returns the path."""
    data = {"step": np.array([lbm.step], dtype=np.int64),
            "flow/f": lbm.flow.f.detach().cpu().numpy()}
    if hasattr(lbm.flow, "fol_array"):
        data["flow/FOL"] = lbm.flow.fol_array().detach().cpu().numpy()
    if hasattr(lbm.flow, "steps_done"):
        data["flow/steps_done"] = np.array([lbm.flow.steps_done], dtype=np.int64)
    sw = lbm.particle_swarm
    for i, p in enumerate(sw.particles):
        g = _particle_group(i)
        for key in PARTICLE_KEYS:
            t = getattr(p, key)
            data[f"{g}/{key}"] = t.detach().cpu().numpy()
        data[f"{g}/state"] = sw.pstate[i].detach().cpu().numpy()
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez_compressed(path, **data)
    return path


def load(lbm, path: str) -> None:
    """Restores a checkpoint written by dump (no pickled objects are read)."""
    with np.load(path, allow_pickle=False) as z:
        lbm.step = int(z["step"][0])
        fl = lbm.flow
        f = z["flow/f"]
        if tuple(f.shape) != tuple(fl.f.shape):
            raise ValueError(f"checkpoint f has shape {f.shape}, the flow {tuple(fl.f.shape)}")
        if "flow/FOL" in z.files:
            fl.FOL = torch.from_numpy(np.ascontiguousarray(z["flow/FOL"]))
        else:
            fl.FOL = torch.zeros(tuple(fl.f.shape[:3]) + (3,))
        fl.f = torch.from_numpy(np.ascontiguousarray(f))
        fl.update_u_rou()
        if "flow/steps_done" in z.files and hasattr(fl, "steps_done"):
            fl.steps_done = int(z["flow/steps_done"][0])
        sw = lbm.particle_swarm
        for i, p in enumerate(sw.particles):
            g = _particle_group(i)
            if f"{g}/state" in z.files:
                sw.pstate[i].copy_(torch.from_numpy(z[f"{g}/state"]).to(sw.pstate.device))
            for key in PARTICLE_KEYS:
                name = f"{g}/{key}"
                if name in z.files:
                    dst = getattr(p, key)
                    dst.copy_(torch.from_numpy(z[name]).to(dst.device, dst.dtype).view_as(dst))

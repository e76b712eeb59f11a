"""Moment boxes of the particles for the interpolation from the fused state.

A particle whose body-frame markers lie within R of its centre has every
marker stencil inside the cube of side 2R + 4 from floor(centre) - R - 1.  The
boxed interpolation (funlbm_ibm_interp_boxed, and the batched step) computes
(rho, u) once for every node of that cube, then interpolates every marker
from those values instead of pulling Q populations per covering marker; the
result is the same bit for bit (flow_to_lagrange,
src/funlbm/lbm/lbm3d.py:46-100).  A box is worth it while its node count is
small against the markers' 64 stencil pulls each; beyond that, and for
particles without markers, R = -1 (no box: its markers pull per node).
"""

import math
from typing import List, Sequence, Tuple

import numpy as np

CHUNK = 256  # nodes per chunk of the box kernel (one block pass)
MAX_R = 512  # largest R the kernels accept (S^3 stays in int32)
# a box is used while its nodes are at most this many per marker; each
# marker pulls 64 stencil nodes without a box
NODES_PER_MARKER = 32


def box_radius(lag: np.ndarray) -> int:
    """R for body-frame markers `lag` ([M, 3]): the ceiling of their largest
    distance from the centre (a marker that float rounding puts a hair
    beyond it misses its box only while a centre coordinate lies within that
    hair of an integer, and is then interpolated per node); -1 if the box is
    not worth computing."""
    lag = np.asarray(lag, dtype=np.float64).reshape(-1, 3)
    if len(lag) == 0:
        return -1
    r = float(np.sqrt((lag * lag).sum(1)).max())
    R = max(int(math.ceil(r - 1e-4)), 0)
    if R > MAX_R or box_nodes(R) > NODES_PER_MARKER * len(lag):
        return -1
    return R


# boxes whose claims a block forms in shared memory (at most 16384 nodes)
CLAIM_MAX_R = 10


def claims_in_box(R: int) -> bool:
    """Whether a particle of box radius R claims its spreading nodes through
    its box in shared memory (otherwise per marker)."""
    return 0 <= R <= CLAIM_MAX_R


def box_nodes(R: int) -> int:
    """Nodes of the cube of a box of radius R (0 for no box)."""
    return (2 * R + 4) ** 3 if R >= 0 else 0


def box_chunks(R: int) -> int:
    return -(-box_nodes(R) // CHUNK)


def layout(radii: Sequence[int], offsets: Sequence[int] = None) -> Tuple[np.ndarray, np.ndarray, int]:
    """Box rows int32[P, 3] = (cache offset, R, first chunk) and the chunk
    list (particle index of each chunk) for particles of radii `radii`.
    Without `offsets` the boxes are packed from cache offset 0; the third
    value is the cache nodes they need."""
    P = len(radii)
    rows = np.zeros((P, 3), dtype=np.int32)
    chunks: List[np.ndarray] = []
    nodes = 0
    first = 0
    for i, R in enumerate(radii):
        R = int(R)
        off = nodes if offsets is None else int(offsets[i])
        rows[i] = (off if R >= 0 else 0, R, first)
        if R >= 0:
            nc = box_chunks(R)
            chunks.append(np.full(nc, i, dtype=np.int32))
            first += nc
            if offsets is None:
                nodes += box_nodes(R)
    chunk = np.concatenate(chunks) if chunks else np.zeros(0, dtype=np.int32)
    if offsets is not None:
        nodes = max((int(offsets[i]) + box_nodes(int(R)) for i, R in enumerate(radii) if R >= 0), default=0)
    return rows, chunk, nodes

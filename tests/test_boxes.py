"""Moment boxes of the boxed interpolation (host logic, CPU only): every
marker stencil lies in its particle's box, and the box tables are laid out
as the kernels index them."""

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from funlbm_amd.particle import boxes
from oracle import ibm_oracle as ibo


def stencil_nodes(x, size):
    """Stencil node coordinates of marker x per axis (src/funlbm/lbm/lbm3d.py:32-44):
    [max(floor(x) - 1, 0), min(rl + 4, size))."""
    out = []
    for d in range(3):
        rl = max(np.floor(np.float32(x[d])) - 1, 0)
        rl = min(rl, size[d])
        rr = min(rl + 4, size[d])
        out.append(np.arange(int(rl), int(rr)))
    return out


@pytest.mark.parametrize("semi", [(4, 4, 4), (3, 3, 5), (4, 4, 8), (2.5, 2.5, 2.5), (7, 7, 7)])
def test_every_stencil_lies_in_the_box(semi):
    """Markers placed as the pose kernel places them (float32 rotation result
    plus the float32 centre) at random centres and orientations: each
    interior stencil node falls inside the cube [floor(c) - R - 1,
    floor(c) + R + 3).  Centres on and next to integers, where float rounding
    may put a stencil one node beyond (the kernels then interpolate that
    marker per node), lose at most a few markers."""
    lag = ibo.surface_points_closed_form(*semi).astype(np.float32)
    R = boxes.box_radius(lag)
    assert R == int(np.ceil(max(semi)))
    S = 2 * R + 4
    rng = np.random.default_rng(7)
    size = (600, 600, 600)
    generic = list(rng.uniform(20, 580, size=(40, 3)))
    near = [np.array([100.0, 200.0, 300.0]), np.array([100 - 1e-9, 200 + 1e-9, 300 - 3e-7]),
            np.array([511.9999999, 64.00000001, 99.5])]
    for c in generic + near:
        rot = Rotation.from_rotvec(rng.normal(size=3)).as_matrix()
        v = (lag.astype(np.float64) @ rot.T).astype(np.float32)
        lx = v + c.astype(np.float32)
        o = np.floor(c).astype(np.int64) - (R + 1)
        out = 0
        for x in lx:
            inside = True
            for d, nodes in enumerate(stencil_nodes(x, size)):
                local = nodes - o[d]
                inside &= bool(local.min() >= 0 and local.max() < S)
            out += not inside
        frac = np.abs(c - np.round(c))
        if frac.min() > 1e-3:
            assert out == 0, (semi, c)
        else:
            assert out <= 0.05 * len(lx), (semi, c, out)


def test_radius_rules():
    assert boxes.box_radius(np.zeros((0, 3))) == -1
    # a margin above the largest distance, integer
    lag = np.zeros((100, 3))
    lag[0] = (4.0, 0, 0)
    assert boxes.box_radius(lag) == 4
    lag[0] = (4.01, 0, 0)
    assert boxes.box_radius(lag) == 5
    # one marker does not pay for a box of (2*4 + 4)^3 nodes
    assert boxes.box_radius(np.array([[3.2, 0, 0]])) == -1
    # a box is used only while it holds at most 32 nodes per marker
    many = ibo.surface_points_closed_form(4, 4, 4)
    R = boxes.box_radius(many)
    assert R >= 0 and boxes.box_nodes(R) <= boxes.NODES_PER_MARKER * len(many)
    assert boxes.box_radius(many[:10]) == -1
    assert boxes.box_nodes(-1) == 0 and boxes.box_nodes(2) == 512


def test_layout_packs_boxes_and_chunks():
    radii = [3, -1, 5, 0]
    rows, chunks, nodes = boxes.layout(radii)
    assert nodes == sum(boxes.box_nodes(r) for r in radii)
    assert rows[1].tolist() == [0, -1, rows[2, 2]]
    off = 0
    for i, r in enumerate(radii):
        if r < 0:
            continue
        assert rows[i, 0] == off and rows[i, 1] == r
        n = boxes.box_chunks(r)
        # the chunks of box i are consecutive from its first chunk
        assert (chunks[rows[i, 2] : rows[i, 2] + n] == i).all()
        off += boxes.box_nodes(r)
    assert len(chunks) == sum(boxes.box_chunks(r) for r in radii if r >= 0)
    # given offsets (the batch engine's allocator), the boxes stay where they are
    rows2, chunks2, nodes2 = boxes.layout([2, 4], [1000, 0])
    assert rows2[:, 0].tolist() == [1000, 0]
    assert nodes2 == max(1000 + boxes.box_nodes(2), boxes.box_nodes(4))
    assert chunks2.tolist() == [0] * boxes.box_chunks(2) + [1] * boxes.box_chunks(4)

"""Same-level neighbour table (dgr_block_neighbours, host only): the links
the ghost refill and the gathering RK4 stages copy ghosts through.  A link
pairs blocks whose corners are sz-7 points apart on a common lattice; blocks
off that lattice by a fraction of a spacing overlap their lattice neighbours
and must get no link at all."""
import numpy as np

from pydgr import layout
from pydgr.native import block_neighbours


def _block(corner, h=1.0 / 64, sz=13, group=0, off=0):
    return {"offset": off, "sz": [sz] * 3, "bflag": 0, "group": group,
            "pmin": [c - 3 * h for c in corner], "pmax": [c + (sz - 4) * h for c in corner]}


def slot(dx, dy, dz):
    return (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)


def test_aligned_blocks_link_both_ways():
    h = 1.0 / 64
    e = 6 * h  # one element: 13^3 blocks are 6 spacings wide
    b = [_block((0.0, 0.0, 0.0)), _block((e, 0.0, 0.0), off=13 ** 3), _block((e, e, e), off=2 * 13 ** 3)]
    n = block_neighbours(b)
    assert n[0, slot(1, 0, 0)] == 1 and n[1, slot(-1, 0, 0)] == 0
    assert n[0, slot(1, 1, 1)] == 2 and n[2, slot(-1, -1, -1)] == 0
    assert n[1, slot(0, 1, 1)] == 2
    assert (n >= 0).sum() == 6


def test_offset_blocks_get_no_link():
    h = 1.0 / 64
    e = 6 * h
    # block 1 sits half a spacing off the lattice of block 0: it overlaps its
    # would-be neighbour position and must not be linked either way
    b = [_block((0.0, 0.0, 0.0)), _block((e + 0.5 * h, 0.0, 0.0), off=13 ** 3),
         _block((0.0, e, 0.0), off=2 * 13 ** 3)]
    n = block_neighbours(b)
    assert (n[1] == -1).all()
    assert 1 not in n
    assert n[0, slot(0, 1, 0)] == 2 and n[2, slot(0, -1, 0)] == 0


def test_lattice_origin_need_not_be_zero():
    h = 0.37 / 6
    e = 6 * h
    o = (0.1234, -0.777, 3.3)
    b = [_block(o, h=h), _block((o[0], o[1], o[2] + e), h=h, off=13 ** 3)]
    n = block_neighbours(b)
    assert n[0, slot(0, 0, 1)] == 1 and n[1, slot(0, 0, -1)] == 0


def test_groups_and_sizes_never_link():
    h = 1.0 / 64
    e = 6 * h
    b = [_block((0.0, 0.0, 0.0)), _block((e, 0.0, 0.0), group=1, off=13 ** 3),
         _block((0.0, e, 0.0), sz=19, off=2 * 13 ** 3)]
    assert (block_neighbours(b) == -1).all()


def test_uniform_grid_is_fully_linked():
    blocks, _ = layout.uniform_grid(4, dx=1.0 / 64)
    n = block_neighbours(blocks)
    # every block has as many neighbours as lattice positions inside the cube
    for b, row in enumerate(n):
        i, j, k = b % 4, (b // 4) % 4, b // 16
        want = sum(1 for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)
                   if (dx, dy, dz) != (0, 0, 0) and 0 <= i + dx < 4 and 0 <= j + dy < 4 and 0 <= k + dz < 4)
        assert (row >= 0).sum() == want


def test_mixed_batch_links_are_consistent():
    """On the uneven 3000-block batch every link is symmetric and joins
    blocks whose corners differ by exactly (sz-7) h along linked axes."""
    blocks, _ = layout.mixed_blocks(3000, seed=0)
    n = block_neighbours(blocks)
    for b, row in enumerate(n):
        for s in np.nonzero(row >= 0)[0]:
            dx, dy, dz = s % 3 - 1, (s // 3) % 3 - 1, s // 9 - 1
            o = row[s]
            assert n[o, slot(-dx, -dy, -dz)] == b
            B, O = blocks[b], blocks[o]
            for a, d in enumerate((dx, dy, dz)):
                h = (B["pmax"][a] - B["pmin"][a]) / (B["sz"][a] - 1)
                assert abs(O["pmin"][a] - B["pmin"][a] - d * (B["sz"][a] - 7) * h) < 1e-9 * max(1.0, abs(B["pmin"][a]))

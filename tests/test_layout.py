"""Block-table generators: the adaptive octree tiles its cube exactly, is
2:1 balanced across faces, edges and corners, refines toward the
punctures, and gives every block consistent sizes, spacing and flags."""
import collections

import numpy as np

from pydgr import layout


def _octants(blocks, lev, half_width):
    out = []
    for b, l in zip(blocks, lev):
        e = 2.0 * half_width / (1 << l)
        dx = (b["pmax"][0] - b["pmin"][0]) / (b["sz"][0] - 1)
        assert abs(dx - e / 6) < 1e-12 * e
        corner = [b["pmin"][d] + 3 * dx for d in range(3)]
        idx = tuple(int(round((corner[d] + half_width) / e)) for d in range(3))
        out.append((l,) + idx)
    return out


def test_adaptive_octree_tiles_balanced():
    hw = 8.0
    blocks, dof, lev = layout.adaptive_octree(half_width=hw, base_level=1, levels=4, centers=((1.3, 0.2, -0.4),),
                                              factor=2.0)
    assert dof == len(blocks) * 13 ** 3 and blocks[-1]["offset"] + 13 ** 3 == dof
    octs = _octants(blocks, lev, hw)
    vol = sum((2 * hw / (1 << o[0])) ** 3 for o in octs)
    assert abs(vol - (2 * hw) ** 3) < 1e-9
    assert len(set(octs)) == len(octs)
    leaves = set(octs)
    lo = min(lev)

    def leaf_level_at(l, i, j, k):
        n = 1 << l
        if not all(0 <= c < n for c in (i, j, k)):
            return None
        while l >= lo:
            if (l, i, j, k) in leaves:
                return l
            l, i, j, k = l - 1, i >> 1, j >> 1, k >> 1
        return None

    for (l, i, j, k) in octs:
        for d in range(27):
            nl = leaf_level_at(l, i + d % 3 - 1, j + (d // 3) % 3 - 1, k + d // 9 - 1)
            assert nl is None or nl >= l - 1
    assert max(lev) == 4 and collections.Counter(lev)[4] > 0


def test_adaptive_octree_refines_toward_centres_and_flags_faces():
    hw = 8.0
    c = np.array([2.0, -1.0, 0.5])
    blocks, _, lev = layout.adaptive_octree(half_width=hw, base_level=1, levels=4, centers=(tuple(c),), factor=2.0)
    finest = [b for b, l in zip(blocks, lev) if l == 4]
    assert finest
    e4 = 2 * hw / 16
    for b in finest:
        # the parent (edge 2 e4) was split for centre distance < factor * 2 e4
        mid = 0.5 * (np.array(b["pmin"]) + np.array(b["pmax"]))
        assert np.linalg.norm(mid - c) < 2.0 * 2 * e4 + 0.5 * np.sqrt(3) * e4
    for b in blocks:
        for d in range(3):
            lo_face = abs(b["pmin"][d] + 3 * (b["pmax"][d] - b["pmin"][d]) / 12 + hw) < 1e-9
            hi_face = abs(b["pmax"][d] - 3 * (b["pmax"][d] - b["pmin"][d]) / 12 - hw) < 1e-9
            assert bool(b["bflag"] >> (2 * d) & 1) == lo_face
            assert bool(b["bflag"] >> (2 * d + 1) & 1) == hi_face

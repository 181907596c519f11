"""Block tables for the benchmark workloads.

A block is a dict with the arguments bssnrhs takes per block
(BSSN_GR/include/rhs.h:47-50): offset into the unzipped arrays, padded size
sz[3], boundary flag bflag (OCT_DIR_* bits, include/dendro.h:84-89) and the
padded extent pmin/pmax (corner -/+ 3 dx, BSSN_GR/src/rhs.cpp:51-62).
Offsets are consecutive (DENDRO_BLOCK_ALIGN_FACTOR = 1, include/dendro.h:173).
"""
import numpy as np

GHOST = 3


def block_dict(offset, sz, bflag, corner, dx, ele_extent):
    """Block whose interior spans corner .. corner + ele_extent*dx per axis."""
    pmin = [corner[d] - GHOST * dx for d in range(3)]
    pmax = [corner[d] + ele_extent[d] * dx + GHOST * dx for d in range(3)]
    return {"offset": int(offset), "sz": [int(s) for s in sz], "bflag": int(bflag), "pmin": pmin, "pmax": pmax}


def uniform_grid(nb, ele_order=6, dx=1.0 / 64, center=True, outer_bflags=True):
    """nb^3 blocks of one order-`ele_order` element each (sz = ele_order+7)
    tiling a cube; blocks on the outer layer carry the boundary flags of
    their outer faces."""
    sz = ele_order + 1 + 2 * GHOST
    n = sz ** 3
    L = nb * ele_order * dx
    o = -0.5 * L if center else 0.0
    blocks = []
    off = 0
    for bz in range(nb):
        for by in range(nb):
            for bx in range(nb):
                bf = 0
                if outer_bflags:
                    for d, bc in enumerate((bx, by, bz)):
                        if bc == 0:
                            bf |= 1 << (2 * d)
                        if bc == nb - 1:
                            bf |= 1 << (2 * d + 1)
                corner = [o + bx * ele_order * dx, o + by * ele_order * dx, o + bz * ele_order * dx]
                blocks.append(block_dict(off, [sz] * 3, bf, corner, dx, [ele_order] * 3))
                off += n
    return blocks, off


def block_grid(nblocks, ele_order=6, dx=1.0 / 64, center=True):
    """Exactly `nblocks` blocks of one element each: the first nblocks cells
    of an nb^3 cube (nb = ceil(nblocks^(1/3))) in x-fastest order (full
    z-layers, then full rows of the last layer, then a partial row), like an
    SFC-contiguous share of a cube.  Every face without a neighbouring block
    carries its boundary flag (Sommerfeld terms, as on the cube's outer
    faces).  Where the last layer or row ends, a block can have ghost points
    beyond no flagged face and with no block to copy from (a concave
    corner); they keep the values they are given (the initial data).  A
    perfect cube gives uniform_grid(nb)."""
    nb = 1
    while nb ** 3 < nblocks:
        nb += 1
    sz = ele_order + 1 + 2 * GHOST
    L = nb * ele_order * dx
    o = -0.5 * L if center else 0.0
    cells = [(bx, by, bz) for bz in range(nb) for by in range(nb) for bx in range(nb)][:nblocks]
    have = set(cells)
    blocks = []
    off = 0
    for (bx, by, bz) in cells:
        bf = 0
        for d in range(3):
            for s, bit in ((-1, 2 * d), (1, 2 * d + 1)):
                q = [bx, by, bz]
                q[d] += s
                if tuple(q) not in have:
                    bf |= 1 << bit
        corner = [o + bx * ele_order * dx, o + by * ele_order * dx, o + bz * ele_order * dx]
        blocks.append(block_dict(off, [sz] * 3, bf, corner, dx, [ele_order] * 3))
        off += sz ** 3
    return blocks, off


def mixed_blocks(nblocks, seed=0, dist=((13, 0.60), (19, 0.25), (25, 0.10), (37, 0.05)), dx=1.0 / 64,
                 box=1.0):
    """Blocks of uneven padded sizes drawn from `dist` (seeded), placed at
    random corners in [-box/2, box/2]^3, all interior (bflag 0)."""
    rng = np.random.default_rng(seed)
    sizes = np.array([s for s, _ in dist])
    probs = np.array([p for _, p in dist], dtype=float)
    probs /= probs.sum()
    pick = rng.choice(len(sizes), size=nblocks, p=probs)
    corners = rng.uniform(-0.5 * box, 0.5 * box, size=(nblocks, 3))
    blocks = []
    off = 0
    for b in range(nblocks):
        sz = int(sizes[pick[b]])
        blocks.append(block_dict(off, [sz] * 3, 0, corners[b].tolist(), dx, [sz - 1 - 2 * GHOST] * 3))
        off += sz ** 3
    return blocks, off


def mixed_lattices(dx=1.0 / 64, counts=((13, 3), (19, 2), (25, 2), (37, 1))):
    """Blocks of uneven padded sizes, each size on a cube lattice of its own
    (n^3 blocks of that size, neighbours sz-7 points apart) and the lattices
    side by side along x, from x = 2 dx so that no point of a flagged face
    sits at the origin, where the Sommerfeld term divides by r: a mixed
    batch in which every size has same-level neighbours.  Boundary flags on
    every outer face of each lattice."""
    blocks = []
    off = 0
    x0 = 2 * dx
    for sz, n in counts:
        e = (sz - 1 - 2 * GHOST) * dx
        for bz in range(n):
            for by in range(n):
                for bx in range(n):
                    bf = 0
                    for d, bc in enumerate((bx, by, bz)):
                        if bc == 0:
                            bf |= 1 << (2 * d)
                        if bc == n - 1:
                            bf |= 1 << (2 * d + 1)
                    corner = [x0 + bx * e, -0.5 * n * e + by * e, -0.5 * n * e + bz * e]
                    blocks.append(block_dict(off, [sz] * 3, bf, corner, dx, [sz - 1 - 2 * GHOST] * 3))
                    off += sz ** 3
        x0 += n * e + 8 * dx
    return blocks, off


def interior_points(blocks):
    return int(sum((b["sz"][0] - 6) * (b["sz"][1] - 6) * (b["sz"][2] - 6) for b in blocks))


def padded_points(blocks):
    return int(sum(b["sz"][0] * b["sz"][1] * b["sz"][2] for b in blocks))


def interior_index(blocks):
    """Flat indices of every interior point, block by block (int64 array)."""
    out = []
    for b in blocks:
        s0, s1, s2 = b["sz"]
        k, j, i = np.meshgrid(np.arange(3, s2 - 3), np.arange(3, s1 - 3), np.arange(3, s0 - 3), indexing="ij")
        out.append((b["offset"] + i + s0 * (j + s1 * k)).ravel())
    return np.concatenate(out).astype(np.int64)


def adaptive_octree(half_width=32.0, base_level=3, levels=6, centers=((4.0, 0.0, 0.0), (-4.0, 0.0, 0.0)),
                    factor=2.0, ele_order=6):
    """2:1-balanced octree of blocks refined toward `centers` (punctures).

    The cube [-half_width, half_width]^3 is cut into octants of level
    base_level; an octant of level l < base_level + levels - 1 is split
    while the distance from its centre to the nearest centre is below
    `factor` times its edge, then the leaves are 2:1 balanced across faces,
    edges and corners (a leaf is split while any neighbouring leaf is more
    than one level finer).  Each leaf becomes one block of one order-
    `ele_order` element (sz = ele_order + 7) with dx = edge / ele_order;
    blocks on the domain boundary carry the flags of their outer faces.
    Returns (blocks, dof, levels_of_blocks)."""
    top = base_level + levels - 1
    cen = np.asarray(centers, dtype=float)

    def edge(l):
        return 2.0 * half_width / (1 << l)

    def centre(l, i, j, k):
        e = edge(l)
        return np.array([-half_width + (i + 0.5) * e, -half_width + (j + 0.5) * e, -half_width + (k + 0.5) * e])

    n0 = 1 << base_level
    leaves = {(base_level, i, j, k) for i in range(n0) for j in range(n0) for k in range(n0)}
    for l in range(base_level, top):
        split = [o for o in leaves if o[0] == l and
                 np.min(np.linalg.norm(cen - centre(*o), axis=1)) < factor * edge(l)]
        for o in split:
            leaves.remove(o)
            _, i, j, k = o
            for c in range(8):
                leaves.add((l + 1, 2 * i + (c & 1), 2 * j + ((c >> 1) & 1), 2 * k + (c >> 2)))

    def leaf_at(l, i, j, k):
        """Leaf containing octant (l, i, j, k) if it is that octant or an
        ancestor; None if the region is refined further or outside."""
        n = 1 << l
        if not (0 <= i < n and 0 <= j < n and 0 <= k < n):
            return None
        while l >= base_level:
            if (l, i, j, k) in leaves:
                return (l, i, j, k)
            l, i, j, k = l - 1, i >> 1, j >> 1, k >> 1
        return None

    changed = True
    while changed:
        changed = False
        for o in sorted(leaves, key=lambda t: -t[0]):
            if o not in leaves:
                continue
            l, i, j, k = o
            for d in range(27):
                if d == 13:
                    continue
                nb = leaf_at(l, i + d % 3 - 1, j + (d // 3) % 3 - 1, k + d // 9 - 1)
                if nb is not None and nb[0] < l - 1:
                    leaves.remove(nb)
                    bl, bi, bj, bk = nb
                    for c in range(8):
                        leaves.add((bl + 1, 2 * bi + (c & 1), 2 * bj + ((c >> 1) & 1), 2 * bk + (c >> 2)))
                    changed = True

    sz = ele_order + 1 + 2 * GHOST
    blocks, lev = [], []
    off = 0
    for o in sorted(leaves):
        l, i, j, k = o
        e = edge(l)
        n = 1 << l
        bf = 0
        for d, c in enumerate((i, j, k)):
            if c == 0:
                bf |= 1 << (2 * d)
            if c == n - 1:
                bf |= 1 << (2 * d + 1)
        corner = [-half_width + i * e, -half_width + j * e, -half_width + k * e]
        blocks.append(block_dict(off, [sz] * 3, bf, corner, e / ele_order, [ele_order] * 3))
        lev.append(l)
        off += sz ** 3
    return blocks, off, lev

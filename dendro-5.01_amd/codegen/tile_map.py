"""Emit csrc/tile_map.h: the thread -> point order of a 7x7x7 tile.

Every LDS stencil read of a thread is its point's index in the 13x13x13
staging buffer plus an offset shared by all threads, so LDS bank conflicts
are set by the point indices alone.  With 64 four-byte banks, a 64-bit read
by 32 lanes is conflict-free when the 32 indices are distinct mod 32.  The
points are grouped by index mod 32 and dealt out round-robin, so each run of
32 consecutive threads holds 32 distinct residues.

Face maps (one per axis and side) serve tiles of full extent that touch one
flagged face: the points of the three slabs nearest the face, where the
closure stencils apply, come first (nearest slab first), so they fill the
first three waves and the other waves take no closure branch.  Each run of
32 threads is filled greedily in that order with points of distinct
residues, which keeps the runs conflict-free.

    python tile_map.py > ../csrc/tile_map.h
"""
from collections import defaultdict

T, S = 7, 13


def order():
    cls = defaultdict(list)
    for tk in range(T):
        for tj in range(T):
            for ti in range(T):
                c = (ti + 3) + S * ((tj + 3) + S * (tk + 3))
                cls[c % 32].append((ti, tj, tk))
    out = []
    k = 0
    while True:
        g = [v[k] for _, v in sorted(cls.items()) if len(v) > k]
        if not g:
            break
        out.extend(g)
        k += 1
    return out


def lds_index(p):
    a, b, c = p
    return (a + 3) + S * ((b + 3) + S * (c + 3))


def face_order(axis, side):
    """Points by slab distance from the face (axis, side 0 low / 1 high),
    dealt into runs of 32 with distinct residues."""
    near = (0, 1, 2) if side == 0 else (T - 1, T - 2, T - 3)

    def key(p):
        return (near.index(p[axis]) if p[axis] in near else len(near), lds_index(p))

    rem = sorted(((a, b, c) for c in range(T) for b in range(T) for a in range(T)), key=key)
    out = []
    while rem:
        used, run, rest = set(), [], []
        for p in rem:
            r = lds_index(p) % 32
            if len(run) < 32 and r not in used:
                used.add(r)
                run.append(p)
            else:
                rest.append(p)
        out.extend(run)
        rem = rest
    return out


def check(pts):
    assert len(pts) == T ** 3 and len(set(pts)) == T ** 3
    for h in range(0, len(pts), 32):
        rs = [lds_index(p) % 32 for p in pts[h:h + 32]]
        assert len(rs) == len(set(rs))
    return [a | (b << 3) | (c << 6) for a, b, c in pts]


def main():
    vals = check(order())
    print("// generated by dendro-5.01_amd/codegen/tile_map.py; do not edit.")
    print("// Thread t of a tile owns point (m & 7, (m >> 3) & 7, m >> 6), m = c_tile_map[t];")
    print("// each run of 32 threads has LDS indices that are distinct mod 32.")
    print("#pragma once")
    print("__constant__ unsigned short c_tile_map[%d] = {" % len(vals))
    for i in range(0, len(vals), 16):
        print("    " + ", ".join(str(v) for v in vals[i:i + 16]) + ",")
    print("};")
    print("// Face maps, index 2 * axis + side (side 0 = low face): the same form, the")
    print("// points of the three slabs nearest that face first.")
    print("__constant__ unsigned short c_face_map[6][%d] = {" % len(vals))
    for axis in range(3):
        for side in range(2):
            fv = check(face_order(axis, side))
            print("  {")
            for i in range(0, len(fv), 16):
                print("    " + ", ".join(str(v) for v in fv[i:i + 16]) + ",")
            print("  },")
    print("};")


if __name__ == "__main__":
    main()

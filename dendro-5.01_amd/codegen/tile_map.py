"""Emit csrc/tile_map.h: the thread -> point order of a 7x7x7 tile.

Every LDS stencil read of a thread is its point's index in the 13x13x13
staging buffer plus an offset shared by all threads, so LDS bank conflicts
are set by the point indices alone.  With 64 four-byte banks, a 64-bit read
by 32 lanes is conflict-free when the 32 indices are distinct mod 32.  The
points are grouped by index mod 32 and dealt out round-robin, so each run of
32 consecutive threads holds 32 distinct residues.

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


def main():
    pts = order()
    assert len(pts) == T ** 3 and len(set(pts)) == T ** 3
    for h in range(0, len(pts), 32):
        rs = [((a + 3) + S * ((b + 3) + S * (c + 3))) % 32 for a, b, c in pts[h:h + 32]]
        assert len(rs) == len(set(rs))
    vals = [a | (b << 3) | (c << 6) for a, b, c in pts]
    print("// generated by dendro-5.01_amd/codegen/tile_map.py; do not edit.")
    print("// Thread t of a tile owns point (m & 7, (m >> 3) & 7, m >> 6), m = c_tile_map[t];")
    print("// each run of 32 threads has LDS indices that are distinct mod 32.")
    print("#pragma once")
    print("__constant__ unsigned short c_tile_map[%d] = {" % len(vals))
    for i in range(0, len(vals), 16):
        print("    " + ", ".join(str(v) for v in vals[i:i + 16]) + ",")
    print("};")


if __name__ == "__main__":
    main()

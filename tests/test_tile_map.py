"""Thread -> point maps of a 7x7x7 tile (csrc/tile_map.h): each map is a
permutation of the tile's points, every run of 32 threads reads LDS
indices distinct mod 32 (conflict-free 64-bit reads over 64 four-byte
banks), the header is what the generator prints, and in each face map the
closure points (the three slabs nearest the face) lie in the first three
waves only."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "dendro-5.01_amd", "csrc", "tile_map.h")
GEN = os.path.join(ROOT, "dendro-5.01_amd", "codegen", "tile_map.py")
T, S = 7, 13


def read_maps():
    text = re.sub(r"//[^\n]*", "", open(HDR).read())
    head, _, faces = text.partition("c_face_map")
    default = [int(v) for v in re.findall(r"\b\d+\b", head.split("{", 1)[1].split("}", 1)[0])]
    blocks = re.findall(r"\{([^{}]*)\}", faces.split("=", 1)[1])
    face = [[int(v) for v in re.findall(r"\b\d+\b", b)] for b in blocks]
    return default, face


def point(m):
    return m & 7, (m >> 3) & 7, m >> 6


def lds(p):
    a, b, c = p
    return (a + 3) + S * ((b + 3) + S * (c + 3))


def test_header_matches_generator():
    out = subprocess.run([sys.executable, GEN], check=True, capture_output=True, text=True).stdout
    assert out == open(HDR).read()


def test_maps_are_conflict_free_permutations():
    default, face = read_maps()
    assert len(face) == 26
    for m in [default] + face:
        pts = [point(v) for v in m]
        assert len(pts) == T ** 3 and len(set(pts)) == T ** 3
        assert all(max(p) < T for p in pts)
        for h in range(0, len(pts), 32):
            rs = [lds(p) % 32 for p in pts[h:h + 32]]
            assert len(rs) == len(set(rs)), h


FACE_SETS = [(sx, sy, sz) for sz in range(3) for sy in range(3) for sx in range(3) if sx or sy or sz]


def face_distance(p, sides):
    d = T
    for a, sd in enumerate(sides):
        if sd == 1:
            d = min(d, p[a])
        elif sd == 2:
            d = min(d, T - 1 - p[a])
    return d


def test_face_map_count_and_index():
    _, face = read_maps()
    assert len(face) == 26
    # index sx + 3 sy + 9 sz - 1
    assert [sx + 3 * sy + 9 * sz - 1 for sx, sy, sz in FACE_SETS] == list(range(26))


@pytest.mark.parametrize("sides", FACE_SETS)
def test_face_map_groups_closure_points(sides):
    _, face = read_maps()
    pts = [point(v) for v in face[sides[0] + 3 * sides[1] + 9 * sides[2] - 1]]
    pos = [t for t, p in enumerate(pts) if face_distance(p, sides) <= 2]
    nfaces = sum(1 for x in sides if x)
    assert len(pos) == T ** 3 - (T - 3) ** nfaces * T ** (3 - nfaces)
    waves = -(-len(pos) // 64)
    if nfaces == 1:
        # wave w holds the 49 points at distance w (one closure row for the
        # wave's closure lanes) and 15 points at distance >= 3
        for w in range(3):
            d = [face_distance(pts[t], sides) for t in range(64 * w, 64 * w + 64)]
            assert d.count(w) == 49 and sum(1 for x in d if x >= 3) == 15
        assert max(pos) < 64 * 3
        return
    # closure points first: at most one extra wave where residues force interleaving
    assert max(pos) < 64 * (waves + 1)
    # the nearest points (one-sided rows) fill wave 0
    assert all(face_distance(pts[t], sides) <= 1 for t in range(64))

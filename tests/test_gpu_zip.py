"""Zip to unique nodal values (dgr_zip, Mesh::zip of include/mesh.tcc:3775-
3822) and the constraint norms over the zipped nodes (dgr_physcon_norms,
computeConstraintL2Norm of BSSN_GR/include/grUtils.tcc:40-88), through the C
ABI, against a CPU count over unique node coordinates.

The CPU side knows nothing of the library's ownership rule: it takes every
interior point of every block as an integer key on the finest lattice,
removes duplicates, and removes hanging nodes (keys inside a coarser block's
closed interior box but off that block's lattice, which Mesh::zip never
writes).  Every input value is an exact integer function of the key, so all
copies of a node agree bit for bit and each zipped value names its node."""
import numpy as np
import pytest
import torch

from pydgr import Plan, layout, make_params

pytestmark = pytest.mark.gpu


def lattice(blocks):
    """Per block: integer keys (on the finest spacing, from a common origin)
    of its interior points in padded-index order, their padded indices, and
    the block's interior box and spacing in keys."""
    hmin = min((b["pmax"][0] - b["pmin"][0]) / (b["sz"][0] - 1) for b in blocks)
    org = np.min([np.asarray(b["pmin"]) for b in blocks], axis=0)
    out = []
    for b in blocks:
        s0, s1, s2 = b["sz"]
        h = (b["pmax"][0] - b["pmin"][0]) / (s0 - 1)
        step = int(round(h / hmin))
        lo = np.rint((np.asarray(b["pmin"]) + 3 * h - org) / hmin).astype(np.int64)
        k, j, i = np.meshgrid(np.arange(s2 - 6), np.arange(s1 - 6), np.arange(s0 - 6), indexing="ij")
        key = np.stack([lo[0] + step * i.ravel(), lo[1] + step * j.ravel(), lo[2] + step * k.ravel()], axis=1)
        idx = b["offset"] + (i + 3).ravel() + s0 * ((j + 3).ravel() + s1 * (k + 3).ravel())
        hi = lo + step * (np.asarray([s0, s1, s2]) - 7)
        out.append((key, idx, lo, hi, step))
    return out


def unique_nodes(lat):
    """Sorted unique non-hanging keys (N, 3)."""
    keys = np.unique(np.concatenate([k for k, _, _, _, _ in lat]), axis=0)
    hanging = np.zeros(len(keys), bool)
    for _, _, lo, hi, step in lat:
        inside = np.all((keys >= lo) & (keys <= hi), axis=1)
        off = np.any((keys - lo) % step != 0, axis=1)
        hanging |= inside & off
    return keys[~hanging]


def code(key):
    """Exact integer value naming a key."""
    key = np.asarray(key, dtype=np.int64)
    return (key[..., 0] + 4096 * key[..., 1] + 4096 * 4096 * key[..., 2]).astype(np.float64)


def fill(blocks, dof, lat, fn):
    U = np.zeros(dof)
    for key, idx, _, _, _ in lat:
        U[idx] = fn(key)
    return U


def check_zip(cuda, blocks, dof):
    lat = lattice(blocks)
    ref = unique_nodes(lat)
    plan = Plan(blocks, make_params())
    n = plan.zip_size()
    print("blocks %d, unique nodes: library %d, CPU count %d" % (len(blocks), n, len(ref)), flush=True)
    assert n == len(ref)
    U = torch.as_tensor(fill(blocks, dof, lat, code), device=cuda)
    Z = torch.full((n,), -1.0, dtype=torch.float64, device=cuda)
    plan.zip(Z, U)
    torch.cuda.synchronize()
    z = Z.cpu().numpy()
    # every unique node exactly once
    np.testing.assert_array_equal(np.sort(z), np.sort(code(ref)))
    # norms over the zipped nodes: 6 integer-valued arrays, mask >= threshold
    fns = [lambda k, q=q: ((k[:, 0] * (q + 3) + k[:, 1] * 7 + k[:, 2] * (2 * q + 1)) % 41 - 20).astype(np.float64)
           for q in range(6)]
    con = torch.as_tensor(np.stack([fill(blocks, dof, lat, f) for f in fns]), device=cuda)
    mask_fn = lambda k: ((k[:, 0] + k[:, 1] + k[:, 2]) % 3).astype(np.float64)
    mask = torch.as_tensor(fill(blocks, dof, lat, mask_fn), device=cuda)
    norms = plan.physcon_norms(con, mask, threshold=1.0)
    sel = mask_fn(ref) >= 1.0
    assert norms["points"] == int(sel.sum())
    for q, name in enumerate(norms["l2"]):
        v = fns[q](ref)[sel]
        l2 = np.sqrt(np.mean(v * v))
        assert abs(norms["l2"][name] - l2) <= 1e-13 * l2, (name, norms["l2"][name], l2)
        assert norms["max"][name] == np.abs(v).max()
    plan.close()
    return n


def test_zip_uniform_grid(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    assert check_zip(cuda, blocks, dof) == 25 ** 3  # 4 blocks of 7 nodes per axis share faces


def test_zip_mixed_lattices(cuda):
    blocks, dof = layout.mixed_lattices()
    check_zip(cuda, blocks, dof)


def test_zip_adaptive_octree(cuda):
    blocks, dof, lev = layout.adaptive_octree(half_width=4.0, base_level=1, levels=4, centers=((0.7, 0.1, -0.2),),
                                              factor=1.5)
    assert len(set(lev)) >= 3
    n = check_zip(cuda, blocks, dof)
    # hanging nodes and shared coarse/fine nodes are both present here:
    # fewer unique nodes than interior points minus same-level sharing alone
    assert n < sum((b["sz"][0] - 6) ** 3 for b in blocks)

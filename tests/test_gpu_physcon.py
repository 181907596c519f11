"""GPU constraint monitor (dgr_physcon, dgr_physcon_norms) through the C ABI
against the oracle's restatement of physical_constraints on the same seeded
inputs.  Bar: max|d| / max|ref| <= 1e-11 per output over all interior
points; padding untouched; psi4 exactly 0 on the z axis; the masked norms
equal the ones computed from the oracle's outputs with the same node
ownership."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import Plan, data, layout, make_params

from helpers import TOL, rel_err_per_var

SENTINEL = -777.0

pytestmark = pytest.mark.gpu


def run_case(cuda, blocks, dof, U, deriv_order=42):
    ref = po.physcon_blocks(U.numpy(), blocks, fill=SENTINEL)
    Ud = U.to(cuda)
    Cd = torch.full((6, dof), SENTINEL, dtype=torch.float64, device=cuda)
    plan = Plan(blocks, make_params(deriv_order=deriv_order))
    plan.physcon(Cd, Ud)
    torch.cuda.synchronize()
    C = Cd.cpu().numpy()
    ii = layout.interior_index(blocks)
    mask = np.ones(dof, bool)
    mask[ii] = False
    assert np.all(C[:, mask] == SENTINEL), "padding written"
    assert np.all(np.isfinite(C[:, ii]))
    rel = rel_err_per_var(C, ref, ii)
    print("max rel per output", np.array2string(rel, precision=2), flush=True)
    assert rel.max() <= TOL, rel
    return plan, Cd, Ud, C, ref


def test_flat_bumps_with_faces(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=3))


def test_gauge_wave_interior_blocks(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=False)
    run_case(cuda, blocks, dof, data.gauge_wave(blocks, dof, seed=5))


def test_brill_lindquist_and_z_axis(cuda):
    """Blocks around a puncture; the grid holds points on the z axis."""
    blocks, dof = layout.uniform_grid(4, dx=0.05, outer_bflags=True)
    U = data.brill_lindquist(blocks, dof, m1=0.5, m2=0.5, sep=1.0)
    U += 1e-4 * data.flat_bumps(blocks, dof, seed=6, amp=1.0, width=0.3) - 1e-4 * data.flat_bumps(blocks, dof, amp=0.0)
    plan, Cd, Ud, C, ref = run_case(cuda, blocks, dof, U)
    X, Y, Z, idx = [c.numpy() for c in data.coords(blocks)]
    pos = np.empty(dof, dtype=np.int64)
    pos[idx] = np.arange(dof)
    ii = layout.interior_index(blocks)
    ax = ii[(np.abs(X[pos][ii]) <= 1e-7) & (np.abs(Y[pos][ii]) <= 1e-7)]
    assert ax.size > 0
    assert np.all(C[4:6, ax] == 0.0)


def test_mixed_sizes_random_faces_and_order64_plan(cuda):
    """Uneven blocks, random face flags; a deriv64 plan still monitors with
    deriv42 (physcon.h)."""
    blocks, dof = layout.mixed_blocks(10, seed=8, dist=((13, 0.4), (19, 0.3), (25, 0.3)))
    rng = np.random.default_rng(8)
    for b in blocks:
        b["bflag"] = int(rng.integers(0, 64))
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=8, amp=1e-2), deriv_order=64)


def owned_index(blocks):
    """Interior points, each node once: a high-face point is dropped when a
    same-level block adjoins that face (uniform grids)."""
    out = []
    key = {tuple(np.round(np.asarray(b["pmin"]) / 1e-9).astype(np.int64)): n for n, b in enumerate(blocks)}
    for b in blocks:
        s0, s1, s2 = b["sz"]
        h = [(b["pmax"][d] - b["pmin"][d]) / (b["sz"][d] - 1) for d in range(3)]
        has = []
        for d in range(3):
            q = np.asarray(b["pmin"], dtype=np.float64).copy()
            q[d] += (b["sz"][d] - 7) * h[d]
            has.append(tuple(np.round(q / 1e-9).astype(np.int64)) in key)
        k, j, i = np.meshgrid(np.arange(3, s2 - 3), np.arange(3, s1 - 3), np.arange(3, s0 - 3), indexing="ij")
        keep = ~((has[0] & (i == s0 - 4)) | (has[1] & (j == s1 - 4)) | (has[2] & (k == s2 - 4)))
        out.append(b["offset"] + (i + s0 * (j + s1 * k))[keep])
    return np.sort(np.concatenate(out))


def test_masked_norms(cuda):
    blocks, dof = layout.uniform_grid(4, dx=0.05, outer_bflags=True)
    U = data.brill_lindquist(blocks, dof, m1=0.5, m2=0.5, sep=1.0)
    plan, Cd, Ud, C, ref = run_case(cuda, blocks, dof, U)
    thr = 0.2
    norms = plan.physcon_norms(Cd, Ud[0], threshold=thr)
    own = owned_index(blocks)
    assert own.size == 25 ** 3  # 4 blocks of 7 nodes per axis, neighbours share a face: 4*6+1 nodes
    sel = own[U[0].numpy()[own] >= thr]
    assert norms["points"] == sel.size and 0 < sel.size < own.size
    for q, name in enumerate(po.PHYSCON_NAMES):
        l2 = np.sqrt(np.mean(C[q, sel] ** 2))
        assert abs(norms["l2"][name] - l2) <= 1e-12 * max(l2, 1e-300), (name, norms["l2"][name], l2)
        assert norms["max"][name] == np.abs(C[q, sel]).max(), name
    again = plan.physcon_norms(Cd, Ud[0], threshold=thr)
    assert again == norms  # fixed reduction order

"""GPU parity: libdgr_bssn.so through its C ABI against the CPU oracle on the
same seeded inputs.  Bar: max|dRHS| / max|RHS| <= 1e-11 per variable over
all interior points; padding must stay untouched."""
import numpy as np
import pytest
import torch

import pyoracle as po
from pydgr import Plan, data, layout, make_params

from helpers import TOL, rel_err_per_var

SENTINEL = -12345.0

pytestmark = pytest.mark.gpu


def run_case(cuda, blocks, dof, U, prm_kw=None, per_block=False):
    prm_kw = prm_kw or {}
    ref = po.rhs_blocks(U.numpy(), blocks, po.params(**prm_kw), fill=SENTINEL)
    Ud = U.to(cuda)
    Rd = torch.full_like(Ud, SENTINEL)
    plan = Plan(blocks, make_params(**prm_kw))
    plan.rhs(Rd, Ud, per_block=per_block)
    torch.cuda.synchronize()
    R = Rd.cpu().numpy()
    ii = layout.interior_index(blocks)
    rel = rel_err_per_var(R, ref, ii)
    mask = np.ones(dof, bool)
    mask[ii] = False
    assert np.all(R[:, mask] == SENTINEL), "padding written"
    assert np.all(np.isfinite(R[:, ii]))
    print("max rel per var %.3e" % rel.max(), flush=True)
    assert rel.max() <= TOL, rel
    return R, ref


def test_flat_bumps_64_blocks_interior(cuda):
    """64 blocks of 13^3, bflag 0, flat space + 8 Gaussian bumps per variable."""
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=False)
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=0))


def test_flat_bumps_with_boundary_faces(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=1))


def test_flat_space_exactly_zero(cuda):
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, amp=0.0)
    Ud = U.to(cuda)
    Rd = torch.full_like(Ud, SENTINEL)
    Plan(blocks, make_params()).rhs(Rd, Ud)
    torch.cuda.synchronize()
    ii = torch.as_tensor(layout.interior_index(blocks), device=cuda)
    assert torch.count_nonzero(Rd[:, ii]).item() == 0


def test_gauge_wave_with_noise(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    run_case(cuda, blocks, dof, data.gauge_wave(blocks, dof, seed=2))


def test_brill_lindquist_strong_field(cuda):
    """Blocks around one puncture (chi down to ~1e-3)."""
    blocks, dof = layout.uniform_grid(4, dx=0.05, outer_bflags=True)
    for b in blocks:
        for d, c in enumerate((4.0, 0.0, 0.0)):
            b["pmin"][d] += c
            b["pmax"][d] += c
    run_case(cuda, blocks, dof, data.brill_lindquist(blocks, dof))


def test_mixed_block_sizes(cuda):
    """Uneven padded sizes 13..37 in one launch, random boundary flags."""
    blocks, dof = layout.mixed_blocks(12, seed=4, dist=((13, 0.25), (19, 0.25), (25, 0.25), (37, 0.25)))
    rng = np.random.default_rng(4)
    for b in blocks:
        b["bflag"] = int(rng.integers(0, 64))
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=5, amp=1e-2, width=0.3))


def test_nonstandard_parameters(cuda):
    """Other gauge parameters and eta damping active (r >= eta_r0)."""
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 32, outer_bflags=True)
    kw = dict(lambda_=(1, 0, 1, 0), lambda_f=(0.5, 0.75), eta_const=1.5, eta_r0=0.2, eta_damping_exp=2.0,
              ko_sigma=0.1)
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=6, amp=1e-2, width=0.2), kw)


def test_per_block_launch_matches_batched_bitwise(cuda):
    blocks, dof = layout.mixed_blocks(6, seed=7, dist=((13, 0.5), (19, 0.5)))
    U = data.flat_bumps(blocks, dof, seed=8).to(cuda)
    plan = Plan(blocks, make_params())
    R1 = torch.full_like(U, SENTINEL)
    R2 = torch.full_like(U, SENTINEL)
    plan.rhs(R1, U)
    plan.rhs(R2, U, per_block=True)
    torch.cuda.synchronize()
    assert torch.equal(R1, R2)


def test_repeat_calls_deterministic(cuda):
    blocks, dof = layout.uniform_grid(3, dx=1.0 / 64)
    U = data.flat_bumps(blocks, dof, seed=9).to(cuda)
    plan = Plan(blocks, make_params())
    outs = []
    for _ in range(3):
        R = torch.full_like(U, SENTINEL)
        plan.rhs(R, U)
        outs.append(R)
    torch.cuda.synchronize()
    assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])


def test_rhs_ordered_on_the_callers_stream(cuda):
    """The boundary-tile kernel runs on the plan's own stream; the call must
    still complete in order on the caller's stream: work queued after it on
    that stream (a copy) sees every tile, with no device-wide sync."""
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, seed=10).to(cuda)
    plan = Plan(blocks, make_params())
    ref = torch.full_like(U, SENTINEL)
    plan.rhs(ref, U)
    torch.cuda.synchronize()
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        for _ in range(3):
            R = torch.full_like(U, SENTINEL)
            plan.rhs(R, U, st)
            C = R.clone()
            st.synchronize()
            assert torch.equal(C, ref)
    plan.close()


# 6th-order option: deriv64 first / second / advective derivatives with
# ko_deriv42 (QuadGrav/include/derivs.h:80-96), same bar.

def test_sixth_order_flat_bumps_with_faces(cuda):
    blocks, dof = layout.uniform_grid(4, dx=1.0 / 64, outer_bflags=True)
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=11), {"deriv_order": 64})


def test_sixth_order_mixed_sizes_random_faces(cuda):
    blocks, dof = layout.mixed_blocks(12, seed=12, dist=((13, 0.25), (19, 0.25), (25, 0.25), (37, 0.25)))
    rng = np.random.default_rng(12)
    for b in blocks:
        b["bflag"] = int(rng.integers(0, 64))
    run_case(cuda, blocks, dof, data.flat_bumps(blocks, dof, seed=13, amp=1e-2, width=0.3), {"deriv_order": 64})


def test_sixth_order_brill_lindquist(cuda):
    blocks, dof = layout.uniform_grid(4, dx=0.05, outer_bflags=True)
    for b in blocks:
        b["pmin"][0] += 4.0
        b["pmax"][0] += 4.0
    run_case(cuda, blocks, dof, data.brill_lindquist(blocks, dof), {"deriv_order": 64})


def test_sixth_order_differs_from_fourth(cuda):
    """The option really switches operators (not the same kernel twice)."""
    blocks, dof = layout.uniform_grid(2, dx=1.0 / 64, outer_bflags=True)
    U = data.flat_bumps(blocks, dof, seed=14, amp=1e-2, width=0.1).to(cuda)
    out = {}
    for order in (42, 64):
        R = torch.full_like(U, SENTINEL)
        plan = Plan(blocks, make_params(deriv_order=order))
        plan.rhs(R, U)
        R2 = torch.full_like(U, SENTINEL)
        plan.rhs(R2, U, per_block=True)
        torch.cuda.synchronize()
        assert torch.equal(R, R2)
        out[order] = R
    assert not torch.equal(out[42], out[64])


def test_adaptive_octree_brill_lindquist(cuda):
    """Blocks of 4 refinement levels (2:1 balanced) around a puncture, with
    domain-face flags, in one launch; both stencil orders."""
    blocks, dof, _ = layout.adaptive_octree(half_width=4.0, base_level=1, levels=4,
                                            centers=((0.7, 0.1, -0.2),), factor=1.5)
    U = data.brill_lindquist(blocks, dof, m1=0.5, m2=0.0, sep=1.4)
    for order in (42, 64):
        run_case(cuda, blocks, dof, U, {"deriv_order": order})


def test_small_blocks_order4_elements(cuda):
    """Blocks of order-4 elements (sz = 11, 5^3 interior points, the
    reference's own test bed: BSSN_GR/src/rhsTestBed.cpp:63) without flagged
    faces, and sz = 12 blocks (the smallest the one-sided closures allow)
    with every face flagged, in one batch."""
    from pydgr.layout import block_dict
    blocks, off = [], 0
    dx = 1.0 / 48
    for q, (sz, bf) in enumerate([(11, 0)] * 6 + [(12, 63)] * 3 + [(7, 0)]):
        corner = [0.05 + 0.23 * q, -0.1 + 0.01 * q, 0.02 * q]
        blocks.append(block_dict(off, [sz] * 3, bf, corner, dx, [sz - 7] * 3))
        off += sz ** 3
    run_case(cuda, blocks, off, data.flat_bumps(blocks, off, seed=11, amp=1e-2, width=0.3))

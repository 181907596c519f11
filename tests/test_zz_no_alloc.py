"""dgr_bssn_rhs_blocks_noalloc (the one-shot bssnRHS that allocates no
device memory; its tables live in a file-scope device array) against the
existing one-shot entry point dgr_bssn_rhs_blocks, through the C ABI: two
calls in a row, the second on a smaller input, each bitwise equal to
dgr_bssn_rhs_blocks on every point (interior written, padding untouched).
Its size limits are checked on the CPU (tests/test_abi.py)."""
import numpy as np
import pytest
import torch

from pydgr import data, layout, make_params, native

SENTINEL = -12345.0

pytestmark = pytest.mark.gpu


def both(cuda, blocks, dof, U, stream):
    """RHS by the existing one-shot entry point and by the no-allocation
    one, from the same input, into arrays prefilled with a sentinel."""
    Ud = U.to(cuda)
    Ra = torch.full_like(Ud, SENTINEL)
    Rb = torch.full_like(Ud, SENTINEL)
    prm = make_params()
    native.rhs_blocks(Ra, Ud, blocks, prm, stream=stream)
    native.rhs_blocks_noalloc(Rb, Ud, blocks, prm, stream=stream)
    stream.synchronize()
    return Ra.cpu().numpy(), Rb.cpu().numpy()


def check_equal(blocks, dof, Ra, Rb):
    ii = layout.interior_index(blocks)
    mask = np.ones(dof, bool)
    mask[ii] = False
    assert np.all(Rb[:, mask] == SENTINEL), "padding written"
    assert np.all(np.isfinite(Rb[:, ii]))
    assert np.array_equal(Ra, Rb), np.abs(Ra - Rb).max()


def test_noalloc_equals_one_shot_twice_second_smaller(cuda):
    native.check(native.lib().dgr_init(0), "dgr_init")
    stream = torch.cuda.Stream()
    # first call: 27 blocks of 13^3, the outer ones with flagged faces
    # (interior and boundary kernels), gauge wave with noise
    b1, dof1 = layout.uniform_grid(3, dx=1.0 / 64, outer_bflags=True)
    print("call 1: %d blocks 13^3, gauge wave" % len(b1), flush=True)
    Ra, Rb = both(cuda, b1, dof1, data.gauge_wave(b1, dof1, seed=3), stream)
    check_equal(b1, dof1, Ra, Rb)
    # second call, smaller: 8 blocks of mixed sizes 13-25 (several tiles per
    # block), no flagged faces, other data; the array keeps call 1's tables
    # beyond what call 2 rewrites
    b2, dof2 = layout.mixed_blocks(8, seed=5, dist=((13, 0.5), (19, 0.3), (25, 0.2)))
    print("call 2: %d blocks, sizes %s" % (len(b2), sorted({b["sz"][0] for b in b2})), flush=True)
    Ra2, Rb2 = both(cuda, b2, dof2, data.flat_bumps(b2, dof2, seed=4), stream)
    check_equal(b2, dof2, Ra2, Rb2)

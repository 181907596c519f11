"""The C-ABI library loads and exports every entry point include/dgr_bssn.h
declares (no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "dgr_bssn.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(dgr_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for s in ("dgr_bssn_rhs", "dgr_bssn_rhs_per_block", "dgr_plan_create", "dgr_init"):
        assert s in syms


def test_library_exports_all_declared_symbols():
    from pydgr import native
    L = ctypes.CDLL(native.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing


def test_error_strings():
    from pydgr import native
    L = native.lib()
    assert L.dgr_error_string(0) == b"ok"
    assert L.dgr_error_string(-1) == b"invalid argument"


def test_plan_rejects_bad_input_without_device_work():
    from pydgr import native
    L = native.lib()
    h = ctypes.c_void_p()
    p = native.make_params()
    assert L.dgr_plan_create(ctypes.byref(h), None, 0, ctypes.byref(p)) == -1
    b = native.blocks_array([{"offset": 0, "sz": [6, 13, 13], "bflag": 0, "pmin": [0, 0, 0], "pmax": [1, 1, 1]}])
    assert L.dgr_plan_create(ctypes.byref(h), b, 1, ctypes.byref(p)) == -1
    # flagged faces need 6 interior points for the one-sided closures
    b = native.blocks_array([{"offset": 0, "sz": [11, 13, 13], "bflag": 1, "pmin": [0, 0, 0], "pmax": [1, 1, 1]}])
    assert L.dgr_plan_create(ctypes.byref(h), b, 1, ctypes.byref(p)) == -1
    b = native.blocks_array([{"offset": 0, "sz": [13, 13, 13], "bflag": 0, "pmin": [0, 0, 0], "pmax": [1, 1, 1]}])
    bad = native.Schedule.from_dict({"tile_max": [8, 7, 7, 7], "order": 0, "grid_mult": 1})
    assert L.dgr_plan_create_ex(ctypes.byref(h), b, 1, ctypes.byref(p), ctypes.byref(bad)) == -1
    bad = native.Schedule.from_dict({"tile_max": [7, 7, 7, 7], "order": 4, "grid_mult": 1})
    assert L.dgr_plan_create_ex(ctypes.byref(h), b, 1, ctypes.byref(p), ctypes.byref(bad)) == -1
    p.deriv_order = 66  # only 42 and 64 exist
    assert L.dgr_plan_create(ctypes.byref(h), b, 1, ctypes.byref(p)) == -4


def test_noalloc_rejects_inputs_beyond_its_limits_without_device_work():
    """dgr_bssn_rhs_blocks_noalloc returns DGR_E_INVALID before any HIP call
    when the block or tile tables would not fit its device array (the array
    pointers are never dereferenced on that path)."""
    import numpy as np
    from pydgr import native
    L = native.lib()
    prm = native.make_params()
    fake = (ctypes.c_void_p * 24)(*[ctypes.c_void_p(0x1000 + 8 * v) for v in range(24)])
    pp = ctypes.cast(fake, ctypes.POINTER(ctypes.c_void_p))
    big = np.zeros(8193, dtype=native._block_dtype())  # DGR_NOALLOC_MAX_BLOCKS + 1
    big["sz"] = 13
    big["pmax"] = 1.0
    arr = native.blocks_array(big)
    assert L.dgr_bssn_rhs_blocks_noalloc(pp, pp, arr, len(big), ctypes.byref(prm), None) == -1
    # 8192 blocks of 61^3: 8^3 tiles each, beyond DGR_NOALLOC_MAX_TILES
    many = np.zeros(8192, dtype=native._block_dtype())
    many["sz"] = 61
    many["pmax"] = 1.0
    arr = native.blocks_array(many)
    assert L.dgr_bssn_rhs_blocks_noalloc(pp, pp, arr, len(many), ctypes.byref(prm), None) == -1
    # a block below the smallest size is rejected as by dgr_plan_create
    one = native.blocks_array([{"offset": 0, "sz": [6, 13, 13], "bflag": 0, "pmin": [0, 0, 0], "pmax": [1, 1, 1]}])
    assert L.dgr_bssn_rhs_blocks_noalloc(pp, pp, one, 1, ctypes.byref(prm), None) == -1

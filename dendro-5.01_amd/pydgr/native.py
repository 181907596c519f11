"""ctypes binding of libdgr_bssn.so (the C ABI declared in include/dgr_bssn.h).

Loading fails loudly when the library has not been built: there is no
CPU fallback for any entry point.
"""
import ctypes
import os

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("DGR_LIB_PATH") or os.path.join(_PKG, "lib", "libdgr_bssn.so")
NUM_VARS = 24

VAR_NAMES = ["alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1", "beta2", "B0", "B1", "B2",
             "gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "At0", "At1", "At2", "At3", "At4", "At5"]


class DgrError(RuntimeError):
    pass


class Block(ctypes.Structure):
    _fields_ = [("offset", ctypes.c_uint64), ("sz", ctypes.c_uint32 * 3), ("bflag", ctypes.c_uint32),
                ("pmin", ctypes.c_double * 3), ("pmax", ctypes.c_double * 3), ("group", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


BLOCK_DTYPE = None


def _block_dtype():
    import numpy as np
    global BLOCK_DTYPE
    if BLOCK_DTYPE is None:
        BLOCK_DTYPE = np.dtype([("offset", "<u8"), ("sz", "<u4", 3), ("bflag", "<u4"), ("pmin", "<f8", 3),
                                ("pmax", "<f8", 3), ("group", "<u4"), ("reserved", "<u4")], align=True)
        assert BLOCK_DTYPE.itemsize == ctypes.sizeof(Block)
    return BLOCK_DTYPE


class BSSNParams(ctypes.Structure):
    _fields_ = [("lambda_", ctypes.c_double * 4), ("lambda_f", ctypes.c_double * 2),
                ("eta_const", ctypes.c_double), ("eta_r0", ctypes.c_double),
                ("eta_damping_exp", ctypes.c_double), ("ko_sigma", ctypes.c_double),
                ("chi_floor", ctypes.c_double), ("deriv_order", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


SIZE_CLASSES = 4          # padded block extent <= 13, <= 19, <= 25, larger
ORDER_BLOCK, ORDER_XCD, ORDER_LARGEST, ORDER_SFC = 0, 1, 2, 3


class Schedule(ctypes.Structure):
    """dgr_schedule: tile extent per size class, tile order, grid size."""
    _fields_ = [("tile_max", ctypes.c_uint32 * SIZE_CLASSES), ("order", ctypes.c_uint32),
                ("grid_mult", ctypes.c_uint32)]

    def as_dict(self):
        return {"tile_max": list(self.tile_max), "order": int(self.order), "grid_mult": int(self.grid_mult)}

    @classmethod
    def from_dict(cls, d):
        s = cls()
        s.tile_max[:] = [int(x) for x in d["tile_max"]]
        s.order = int(d["order"])
        s.grid_mult = int(d["grid_mult"])
        return s


def default_schedule():
    s = Schedule()
    lib().dgr_schedule_default(ctypes.byref(s))
    return s


def size_class(sz):
    m = max(sz)
    return 0 if m <= 13 else 1 if m <= 19 else 2 if m <= 25 else 3


def make_params(lambda_=(1, 1, 1, 1), lambda_f=(1.0, 0.0), eta_const=2.0, eta_r0=30.0, eta_damping_exp=1.0,
                ko_sigma=0.4, chi_floor=0.1, deriv_order=42):
    p = BSSNParams()
    p.lambda_[:] = [float(x) for x in lambda_]
    p.lambda_f[:] = [float(x) for x in lambda_f]
    p.eta_const = eta_const
    p.eta_r0 = eta_r0
    p.eta_damping_exp = eta_damping_exp
    p.ko_sigma = ko_sigma
    p.chi_floor = chi_floor
    p.deriv_order = deriv_order
    return p


_lib = None


def lib():
    """The loaded library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DgrError("libdgr_bssn.so not built (%s); run python dendro-5.01_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        pp = ctypes.POINTER(ctypes.c_void_p)
        L.dgr_error_string.restype = ctypes.c_char_p
        L.dgr_error_string.argtypes = [ctypes.c_int]
        L.dgr_init.argtypes = [ctypes.c_int]
        L.dgr_plan_create.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Block), ctypes.c_uint32,
                                      ctypes.POINTER(BSSNParams)]
        L.dgr_plan_create_ex.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(Block), ctypes.c_uint32,
                                         ctypes.POINTER(BSSNParams), ctypes.POINTER(Schedule)]
        L.dgr_schedule_default.argtypes = [ctypes.POINTER(Schedule)]
        L.dgr_schedule_default.restype = None
        L.dgr_plan_destroy.argtypes = [vp]
        L.dgr_plan_points.restype = ctypes.c_uint64
        L.dgr_plan_points.argtypes = [vp]
        L.dgr_plan_tiles.restype = ctypes.c_uint32
        L.dgr_plan_tiles.argtypes = [vp]
        for f in ("dgr_bssn_rhs", "dgr_bssn_rhs_per_block"):
            getattr(L, f).argtypes = [vp, pp, pp, vp]
        L.dgr_bssn_rhs_blocks.argtypes = [pp, pp, ctypes.POINTER(Block), ctypes.c_uint32,
                                          ctypes.POINTER(BSSNParams), vp]
        L.dgr_bssn_rhs_blocks_noalloc.argtypes = [pp, pp, ctypes.POINTER(Block), ctypes.c_uint32,
                                                  ctypes.POINTER(BSSNParams), vp]
        L.dgr_ghost_refill.argtypes = [vp, pp, vp]
        L.dgr_plan_level_fills.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint64)]
        L.dgr_enforce_constraints.argtypes = [vp, pp, vp]
        L.dgr_physcon.argtypes = [vp, pp, pp, vp]
        L.dgr_init_puncture.argtypes = [vp, pp, ctypes.POINTER(Puncture), vp]
        L.dgr_init_bumps.argtypes = [vp, pp, vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, vp]
        L.dgr_tp_params_default.argtypes = [ctypes.POINTER(TPParams)]
        L.dgr_tp_params_default.restype = None
        L.dgr_tp_solve.argtypes = [ctypes.POINTER(vp), ctypes.POINTER(TPParams)]
        L.dgr_tp_info.argtypes = [vp, ctypes.POINTER(TPInfo)]
        L.dgr_tp_collocation.argtypes = [vp, vp, ctypes.c_uint64]
        L.dgr_tp_u.argtypes = [vp, vp, vp, ctypes.c_uint64]
        L.dgr_tp_evaluate.argtypes = [vp, vp, pp, vp]
        L.dgr_tp_point.argtypes = [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp]
        L.dgr_tp_destroy.argtypes = [vp]
        L.dgr_bssn_rhs_blkwise.argtypes = [vp, pp, pp, ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, vp]
        L.dgr_bssn_rhs_blk.argtypes = [vp, vp, vp, ctypes.c_uint32, ctypes.c_uint32, vp]
        L.dgr_ets_create.argtypes = [ctypes.POINTER(vp), vp, pp, ctypes.c_int]
        L.dgr_ets_create_tableau.argtypes = [ctypes.POINTER(vp), vp, pp, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                             ctypes.c_uint32]
        L.dgr_ets_evolve.argtypes = [vp, ctypes.c_double, vp]
        L.dgr_ets_time.argtypes = [vp]
        L.dgr_ets_time.restype = ctypes.c_double
        L.dgr_ets_status.argtypes = [vp, vp]
        L.dgr_ets_destroy.argtypes = [vp]
        L.dgr_physcon_norms.argtypes = [vp, pp, vp, ctypes.c_double, ctypes.POINTER(ctypes.c_double), vp]
        L.dgr_rk4_create.argtypes = [ctypes.POINTER(vp), vp, pp, ctypes.c_int]
        L.dgr_rk4_work_doubles.restype = ctypes.c_uint64
        L.dgr_rk4_work_doubles.argtypes = [vp]
        L.dgr_rk4_create_with_work.argtypes = [ctypes.POINTER(vp), vp, pp, vp, ctypes.c_int]
        L.dgr_rk4_set_group_dt.argtypes = [vp, ctypes.POINTER(ctypes.c_double), ctypes.c_uint32]
        L.dgr_rk4_reset_stage.argtypes = [vp, vp]
        L.dgr_rk4_run.argtypes = [vp, ctypes.c_int, vp]
        L.dgr_rk4_step.argtypes = [vp, ctypes.c_double, vp]
        L.dgr_rk4_steps.argtypes = [vp, ctypes.c_double, ctypes.c_int, vp]
        L.dgr_rk4_status.argtypes = [vp]
        L.dgr_rk4_destroy.argtypes = [vp]
        L.dgr_block_neighbours.argtypes = [ctypes.POINTER(Block), ctypes.c_uint32, vp]
        L.dgr_zip_size.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
        L.dgr_zip.argtypes = [vp, vp, vp, vp]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc != 0:
        msg = lib().dgr_error_string(rc).decode()
        raise DgrError("%s failed: %d (%s)" % (what, rc, msg))


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
    return ctypes.cast(arr, ctypes.POINTER(ctypes.c_void_p)), arr


class Puncture(ctypes.Structure):
    """dgr_puncture (include/dgr_bssn.h)."""
    _fields_ = [("mass", ctypes.c_double), ("pos", ctypes.c_double * 3), ("p", ctypes.c_double * 3),
                ("spin", ctypes.c_double), ("spin_theta", ctypes.c_double), ("spin_phi", ctypes.c_double)]


def punctures(holes):
    """Two hole dicts (mass, pos, optional p, spin, spin_theta, spin_phi)
    -> ctypes array of 2 dgr_puncture."""
    arr = (Puncture * 2)()
    for q, h in enumerate(holes):
        arr[q].mass = h["mass"]
        for d in range(3):
            arr[q].pos[d] = h["pos"][d]
            arr[q].p[d] = h.get("p", (0.0, 0.0, 0.0))[d]
        arr[q].spin = h.get("spin", 0.0)
        arr[q].spin_theta = h.get("spin_theta", 0.0)
        arr[q].spin_phi = h.get("spin_phi", 0.0)
    return arr


class TPParams(ctypes.Structure):
    """dgr_tp_params (include/dgr_bssn.h): the reference's TPID parameters."""
    _v3 = ctypes.c_double * 3
    _fields_ = [("par_b", ctypes.c_double), ("par_m_plus", ctypes.c_double), ("par_m_minus", ctypes.c_double),
                ("target_M_plus", ctypes.c_double), ("target_M_minus", ctypes.c_double),
                ("par_P_plus", _v3), ("par_P_minus", _v3), ("par_S_plus", _v3), ("par_S_minus", _v3),
                ("center_offset", _v3), ("give_bare_mass", ctypes.c_int), ("npoints_A", ctypes.c_int),
                ("npoints_B", ctypes.c_int), ("npoints_phi", ctypes.c_int), ("Newton_maxit", ctypes.c_int),
                ("initial_lapse", ctypes.c_int), ("Newton_tol", ctypes.c_double), ("adm_tol", ctypes.c_double),
                ("initial_lapse_psi_exponent", ctypes.c_double), ("TP_epsilon", ctypes.c_double),
                ("TP_Tiny", ctypes.c_double), ("TP_Extend_Radius", ctypes.c_double)]


class TPInfo(ctypes.Structure):
    """dgr_tp_info_t."""
    _fields_ = [("mp", ctypes.c_double), ("mm", ctypes.c_double), ("mp_adm", ctypes.c_double),
                ("mm_adm", ctypes.c_double), ("E", ctypes.c_double), ("J", ctypes.c_double * 3),
                ("newton_steps", ctypes.c_int), ("last_update", ctypes.c_double)]


TP_LAPSE = {"antisymmetric": 0, "averaged": 1, "psin": 2, "brownsville": 3, "unit": 4}


def tp_params(**kw):
    """dgr_tp_params: the reference's defaults (dgr_tp_params_default) with
    the given fields replaced (3-vectors as sequences; initial_lapse also as
    a name of TP_LAPSE)."""
    p = TPParams()
    lib().dgr_tp_params_default(ctypes.byref(p))
    for k, v in kw.items():
        if k == "initial_lapse" and isinstance(v, str):
            v = TP_LAPSE[v]
        cur = getattr(p, k)
        if isinstance(cur, ctypes.Array):
            for d in range(3):
                cur[d] = float(v[d])
        else:
            setattr(p, k, v)
    return p


class TwoPunctures:
    """TwoPunctures initial data (dgr_tp_*): the spectral solve on
    construction, then evaluate(plan, u) fills every padded point."""

    def __init__(self, params=None, **kw):
        self.params = params if params is not None else tp_params(**kw)
        h = ctypes.c_void_p()
        check(lib().dgr_tp_solve(ctypes.byref(h), ctypes.byref(self.params)), "dgr_tp_solve")
        self.handle = h

    def info(self):
        i = TPInfo()
        check(lib().dgr_tp_info(self.handle, ctypes.byref(i)), "dgr_tp_info")
        return {"mp": i.mp, "mm": i.mm, "mp_adm": i.mp_adm, "mm_adm": i.mm_adm, "E": i.E, "J": list(i.J),
                "newton_steps": i.newton_steps, "last_update": i.last_update}

    def collocation(self):
        import numpy as np
        p = self.params
        out = np.empty(p.npoints_A * p.npoints_B * p.npoints_phi)
        check(lib().dgr_tp_collocation(self.handle, out.ctypes.data_as(ctypes.c_void_p), out.size),
              "dgr_tp_collocation")
        return out

    def u_at(self, xyz):
        """u at absolute coordinates (n, 3), evaluated on the host."""
        import numpy as np
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty(len(xyz))
        check(lib().dgr_tp_u(self.handle, xyz.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
                             len(xyz)), "dgr_tp_u")
        return out

    def point(self, x, y, z):
        """The 24 variables at one point on the host (dgr_tp_point, the
        reference's per-point TwoPunctures call)."""
        import numpy as np
        out = np.empty(24)
        check(lib().dgr_tp_point(self.handle, float(x), float(y), float(z), out.ctypes.data_as(ctypes.c_void_p)),
              "dgr_tp_point")
        return out

    def evaluate(self, plan, u, stream=None):
        up, _u = ptr_array(var_ptrs(u))
        check(lib().dgr_tp_evaluate(self.handle, plan.handle, up, _stream_ptr(stream)), "dgr_tp_evaluate")

    def close(self):
        if self.handle:
            lib().dgr_tp_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NUM_CONSTRAINTS = 6
CONSTRAINT_NAMES = ("ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img")


def var_ptrs(t):
    """Device pointers of the 24 rows of a (24, dof) float64 tensor, or of a
    list of 24 tensors."""
    if isinstance(t, (list, tuple)):
        return [x.data_ptr() for x in t]
    assert t.dim() == 2 and t.shape[0] == NUM_VARS
    return [t[v].data_ptr() for v in range(NUM_VARS)]


def blocks_array(blocks):
    """Block dicts (offset, sz, bflag, pmin, pmax, optional group) or a
    numpy array of BLOCK_DTYPE -> ctypes array of dgr_block."""
    import numpy as np
    dt = _block_dtype()
    if isinstance(blocks, np.ndarray) and blocks.dtype == dt:
        a = np.ascontiguousarray(blocks)
    else:
        n = len(blocks)
        a = np.zeros(n, dtype=dt)
        a["offset"] = [int(b["offset"]) for b in blocks]
        a["sz"] = [b["sz"] for b in blocks]
        a["bflag"] = [int(b["bflag"]) for b in blocks]
        a["pmin"] = [b["pmin"] for b in blocks]
        a["pmax"] = [b["pmax"] for b in blocks]
        a["group"] = [int(b.get("group", 0)) for b in blocks]
    arr = (Block * len(a)).from_buffer_copy(a.tobytes())
    return arr


def block_neighbours(blocks):
    """Same-level neighbour table of a block list (dgr_block_neighbours,
    host only): int32 array (nblocks, 27), entry (dx+1) + 3(dy+1) + 9(dz+1)
    the neighbour's index or -1."""
    import numpy as np
    arr = blocks_array(blocks)
    out = np.empty((len(arr), 27), dtype=np.int32)
    check(lib().dgr_block_neighbours(arr, len(arr), out.ctypes.data_as(ctypes.c_void_p)), "dgr_block_neighbours")
    return out


def torch_float64():
    import torch
    return torch.float64


def _stream_ptr(stream):
    if stream is None:
        return None
    return ctypes.c_void_p(int(getattr(stream, "cuda_stream", stream)))


class Plan:
    """Device-resident block table + tile work table (dgr_plan_create)."""

    def __init__(self, blocks, params, device=0, schedule=None):
        L = lib()
        check(L.dgr_init(device), "dgr_init")
        self._barr = blocks_array(blocks)
        self.params = params
        self.schedule = schedule
        h = ctypes.c_void_p()
        sp = ctypes.byref(schedule) if schedule is not None else None
        check(L.dgr_plan_create_ex(ctypes.byref(h), self._barr, len(blocks), ctypes.byref(params), sp),
              "dgr_plan_create_ex")
        self.handle = h
        self.nblocks = len(blocks)

    @property
    def points(self):
        return int(lib().dgr_plan_points(self.handle))

    def level_fills(self):
        """(interpolation boxes, node copies) filling ghosts at level jumps."""
        b, c = ctypes.c_uint32(), ctypes.c_uint64()
        check(lib().dgr_plan_level_fills(self.handle, ctypes.byref(b), ctypes.byref(c)), "dgr_plan_level_fills")
        return int(b.value), int(c.value)

    @property
    def tiles(self):
        return int(lib().dgr_plan_tiles(self.handle))

    def rhs(self, rhs, u, stream=None, per_block=False):
        L = lib()
        rp, _r = ptr_array(var_ptrs(rhs))
        up, _u = ptr_array(var_ptrs(u))
        f = L.dgr_bssn_rhs_per_block if per_block else L.dgr_bssn_rhs
        check(f(self.handle, rp, up, _stream_ptr(stream)), f.__name__)

    def ghost_refill(self, u, stream=None):
        up, _u = ptr_array(var_ptrs(u))
        check(lib().dgr_ghost_refill(self.handle, up, _stream_ptr(stream)), "dgr_ghost_refill")

    def enforce_constraints(self, u, stream=None):
        """Returns True when every point passed the reference's checks."""
        up, _u = ptr_array(var_ptrs(u))
        rc = lib().dgr_enforce_constraints(self.handle, up, _stream_ptr(stream))
        if rc == -5:
            return False
        check(rc, "dgr_enforce_constraints")
        return True

    def rhs_blkwise(self, rhs, u, blk_ids, stream=None):
        """RHS of the blocks blk_ids only (BSSNCtx::rhs_blkwise)."""
        rp, _r = ptr_array(var_ptrs(rhs))
        up, _u = ptr_array(var_ptrs(u))
        ids = (ctypes.c_uint32 * len(blk_ids))(*[int(b) for b in blk_ids])
        check(lib().dgr_bssn_rhs_blkwise(self.handle, rp, up, ids, len(blk_ids), _stream_ptr(stream)),
              "dgr_bssn_rhs_blkwise")

    def rhs_blk(self, out, inp, blk_id, stream=None):
        """RHS of one block with block-local (24, NN) arrays (BSSNCtx::rhs_blk)."""
        assert out.is_contiguous() and inp.is_contiguous() and out.shape == inp.shape and out.shape[0] == NUM_VARS
        check(lib().dgr_bssn_rhs_blk(self.handle, ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(inp.data_ptr()),
                                     NUM_VARS, int(blk_id), _stream_ptr(stream)), "dgr_bssn_rhs_blk")

    def init_puncture(self, u, holes, stream=None):
        """punctureData initial data at every padded point of every block
        of the plan (holes: two dicts, see punctures())."""
        up, _u = ptr_array(var_ptrs(u))
        arr = punctures(holes)
        check(lib().dgr_init_puncture(self.handle, up, arr, _stream_ptr(stream)), "dgr_init_puncture")

    def init_bumps(self, u, centers, amp, width, stream=None):
        """Perturbed flat space at every padded point (dgr_init_bumps):
        centers is a device float64 tensor (24, nbumps, 3)."""
        assert centers.is_cuda and centers.dtype == torch_float64() and centers.dim() == 3 and centers.shape[0] == 24
        c = centers.contiguous()
        up, _u = ptr_array(var_ptrs(u))
        check(lib().dgr_init_bumps(self.handle, up, ctypes.c_void_p(c.data_ptr()), int(c.shape[1]), float(amp),
                                   float(width), _stream_ptr(stream)), "dgr_init_bumps")

    def physcon(self, con, u, stream=None):
        """Constraint monitor (physical_constraints): con is a (6, dof)
        float64 tensor (ham, mom0..2, psi4_real, psi4_img), written on the
        interior points of every block."""
        assert con.dim() == 2 and con.shape[0] == NUM_CONSTRAINTS
        cp, _c = ptr_array([con[q].data_ptr() for q in range(NUM_CONSTRAINTS)])
        up, _u = ptr_array(var_ptrs(u))
        check(lib().dgr_physcon(self.handle, cp, up, _stream_ptr(stream)), "dgr_physcon")

    def zip_size(self):
        """Number of unique nodes of the plan's blocks (dgr_zip_size)."""
        n = ctypes.c_uint64()
        check(lib().dgr_zip_size(self.handle, ctypes.byref(n)), "dgr_zip_size")
        return int(n.value)

    def zip(self, zipped, unzipped, stream=None):
        """Mesh::zip of one unzipped float64 device array into `zipped`
        (float64 device tensor of zip_size() elements)."""
        assert zipped.numel() >= self.zip_size()
        check(lib().dgr_zip(self.handle, ctypes.c_void_p(zipped.data_ptr()), ctypes.c_void_p(unzipped.data_ptr()),
                            _stream_ptr(stream)), "dgr_zip")

    def physcon_norms(self, con, mask, threshold=0.3, stream=None):
        """Masked L2 and max norms of the 6 constraint outputs over the
        nodes where mask >= threshold (alpha and 0.3 by default in the
        reference).  Returns a dict."""
        cp, _c = ptr_array([con[q].data_ptr() for q in range(NUM_CONSTRAINTS)])
        out = (ctypes.c_double * 13)()
        check(lib().dgr_physcon_norms(self.handle, cp, ctypes.c_void_p(mask.data_ptr()), float(threshold), out,
                                      _stream_ptr(stream)), "dgr_physcon_norms")
        return {"l2": dict(zip(CONSTRAINT_NAMES, out[0:6])), "max": dict(zip(CONSTRAINT_NAMES, out[6:12])),
                "points": int(out[12])}

    def close(self):
        if getattr(self, "handle", None):
            lib().dgr_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rhs_blocks(rhs, u, blocks, params, stream=None):
    """One-shot bssnRHS over device arrays (dgr_bssn_rhs_blocks)."""
    L = lib()
    barr = blocks_array(blocks)
    rp, _r = ptr_array(var_ptrs(rhs))
    up, _u = ptr_array(var_ptrs(u))
    check(L.dgr_bssn_rhs_blocks(rp, up, barr, len(blocks), ctypes.byref(params), _stream_ptr(stream)),
          "dgr_bssn_rhs_blocks")


def rhs_blocks_noalloc(rhs, u, blocks, params, stream=None):
    """One-shot bssnRHS that allocates no device memory
    (dgr_bssn_rhs_blocks_noalloc): asynchronous on `stream`; at most
    DGR_NOALLOC_MAX_BLOCKS blocks and DGR_NOALLOC_MAX_TILES tiles."""
    L = lib()
    barr = blocks_array(blocks)
    rp, _r = ptr_array(var_ptrs(rhs))
    up, _u = ptr_array(var_ptrs(u))
    check(L.dgr_bssn_rhs_blocks_noalloc(rp, up, barr, len(blocks), ctypes.byref(params), _stream_ptr(stream)),
          "dgr_bssn_rhs_blocks_noalloc")


class RK4:
    """Device-resident RK4 evolver of the (24, dof) state `u` (dgr_rk4_*).
    `work`: optional caller-owned float64 device tensor of at least
    Plan-dependent dgr_rk4_work_doubles elements."""

    def __init__(self, plan, u, use_graph=True, work=None):
        self.plan = plan
        self.u = u
        up, _u = ptr_array(var_ptrs(u))
        h = ctypes.c_void_p()
        L = lib()
        if work is None:
            check(L.dgr_rk4_create(ctypes.byref(h), plan.handle, up, 1 if use_graph else 0), "dgr_rk4_create")
        else:
            need = int(L.dgr_rk4_work_doubles(plan.handle))
            if work.numel() < need:
                raise DgrError("work buffer too small: %d < %d doubles" % (work.numel(), need))
            check(L.dgr_rk4_create_with_work(ctypes.byref(h), plan.handle, up, ctypes.c_void_p(work.data_ptr()),
                                             1 if use_graph else 0), "dgr_rk4_create_with_work")
        self.handle = h

    def step(self, dt, nsteps=1, stream=None):
        check(lib().dgr_rk4_steps(self.handle, float(dt), int(nsteps), _stream_ptr(stream)), "dgr_rk4_steps")

    def set_group_dt(self, dts):
        arr = (ctypes.c_double * len(dts))(*[float(x) for x in dts])
        check(lib().dgr_rk4_set_group_dt(self.handle, arr, len(dts)), "dgr_rk4_set_group_dt")

    def run(self, nsteps=1, stream=None):
        check(lib().dgr_rk4_run(self.handle, int(nsteps), _stream_ptr(stream)), "dgr_rk4_run")

    def reset_stage(self, stream=None):
        """Copy the state into the stage buffer (after writing a new state
        into this evolver: ghosts no neighbour fills keep the new state's
        values)."""
        check(lib().dgr_rk4_reset_stage(self.handle, _stream_ptr(stream)), "dgr_rk4_reset_stage")

    def status_ok(self):
        rc = lib().dgr_rk4_status(self.handle)
        if rc == -5:
            return False
        check(rc, "dgr_rk4_status")
        return True

    def close(self):
        if getattr(self, "handle", None):
            lib().dgr_rk4_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


ETS_RK3, ETS_RK4 = 0, 1


class ETS:
    """Explicit Runge-Kutta stepper with a Butcher tableau on a plan's
    device-resident state u (dgr_ets_*): ETS type (ETS_RK3 / ETS_RK4) or an
    explicit tableau (a, b, c)."""

    def __init__(self, plan, u, type=ETS_RK4, tableau=None):
        L = lib()
        self.plan = plan
        self._u = u
        up, self._ua = ptr_array(var_ptrs(u))
        h = ctypes.c_void_p()
        if tableau is None:
            check(L.dgr_ets_create(ctypes.byref(h), plan.handle, up, int(type)), "dgr_ets_create")
        else:
            a, b, c = (list(map(float, x)) for x in tableau)
            n = len(b)
            assert len(a) == n * n and len(c) == n
            arr = lambda x: (ctypes.c_double * len(x))(*x)  # noqa: E731
            check(L.dgr_ets_create_tableau(ctypes.byref(h), plan.handle, up, arr(a), arr(b), arr(c), n),
                  "dgr_ets_create_tableau")
        self.handle = h

    def evolve(self, dt, nsteps=1, stream=None):
        for _ in range(nsteps):
            check(lib().dgr_ets_evolve(self.handle, float(dt), _stream_ptr(stream)), "dgr_ets_evolve")

    @property
    def time(self):
        return lib().dgr_ets_time(self.handle)

    def status_ok(self, stream=None):
        rc = lib().dgr_ets_status(self.handle, _stream_ptr(stream))
        if rc == -5:
            return False
        check(rc, "dgr_ets_status")
        return True

    def close(self):
        if self.handle:
            lib().dgr_ets_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


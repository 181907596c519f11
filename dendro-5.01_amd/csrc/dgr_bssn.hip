/*
 * MI355X (gfx950) BSSN right-hand side on ghost-padded blocks.
 *
 * Reference behaviour: bssnrhs (BSSN_GR/src/rhs.cpp:86-392) per block:
 * deriv42 first / second / mixed derivatives (BSSN_GR/src/derivs.cpp),
 * upwinded deriv42adv, the generated BSSN equations, Sommerfeld terms on
 * flagged faces (rhs.cpp:838-982), then Kreiss-Oliger dissipation
 * (ko_deriv42, rhs.cpp:331-373) on every interior point.
 *
 * GPU design (one workgroup per tile of at most 7x7x7 interior points; a
 * 13^3 block is exactly one tile):
 *   - the 24 state values at each thread's point are read once from HBM;
 *   - the 24 padded variables are staged through LDS one at a time; the
 *     global loads of the next variable are issued into registers before
 *     the current one is differentiated, so HBM latency overlaps compute;
 *   - for the 11 variables with second derivatives, the x-derivative over
 *     the (y,z)-widened tile and the y-derivative over the z-widened tile
 *     are written to LDS and differentiated again, which is exactly the
 *     reference's mixed derivative by composition (bssnrhs_derivs.h);
 *   - every stencil is evaluated with FP contraction off and the
 *     reference's operation order, so derivative values are bit-identical
 *     to the reference's (this is where the O(1) background cancels);
 *   - each variable's derivatives are consumed at once by the staged
 *     equations (bssn_staged.h) so no derivative array ever exists;
 *   - Kreiss-Oliger and Sommerfeld terms are folded into one per-variable
 *     register that is added after the equations, as the reference does.
 * All arithmetic is FP64 on the vector ALUs.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "../../include/dgr_bssn.h"
#include "bssn_staged.h"
#include "stencil_forms.h"
#include "dgr_internal.h"
#include "tile_map.h"

namespace {

using namespace dgr_internal;

constexpr int TX = 7, TY = 7, TZ = 7;         // interior tile
constexpr int NPT = TX * TY * TZ;              // 343 points
// 8 waves: threads [0, 343) own a point each; all threads share the
// staging, prefetch and derivative-plane loops (8 waves per CU fit the
// 256-VGPR budget and hide more latency than 6)
constexpr int NT = 512;
constexpr int SX = TX + 6, SY = TY + 6, SZ = TZ + 6;  // staged region
constexpr int NSU = SX * SY * SZ;              // 2197
constexpr int NPF = (NSU + NT - 1) / NT;       // prefetch registers per thread
// Derivative planes for the mixed derivatives by composition: d_x u over
// the tile widened by GW points in y and z, d_y u widened by GW in z.  The
// reference computes the first derivatives on the widened ranges the second
// differentiation reads: +-2 for deriv42 (BSSN_GR/src/derivs.cpp:19-27,
// 96-104), +-3 for deriv64 (QuadGrav/src/derivs.cpp:2151-2158, 2293-2300).
// The 6th-order planes do not both fit next to the per-point slots, so they
// share one buffer and are built one after the other.
template <int ORD>
struct Planes {
  static constexpr int GW = ORD == 64 ? 3 : 2;
  static constexpr int GXY = TY + 2 * GW, GXZ = TZ + 2 * GW;
  static constexpr int NGX = TX * GXY * GXZ;   // 847 / 1183
  static constexpr int NGY = TX * TY * GXZ;    // 539 / 637
  static constexpr bool SHARED = ORD == 64;
  static constexpr int N = SHARED ? (NGX > NGY ? NGX : NGY) : NGX + NGY;
};

// ---------------------------------------------------------------------------
// Stencils.  The value at axis coordinate q' of the current line is
// A[o + q'*s]: `s` is the stride along the axis, q the coordinate of the
// output point.  Interior formulas and operation order follow
// BSSN_GR/src/derivs.cpp (deriv42_x :29-65, deriv42_xx :265-316,
// deriv42adv_x :513-622, ko_deriv42_x :937-1060).  In boundary tiles the
// interior formula is evaluated everywhere and the one-sided closures of
// points within three of a flagged face replace it from the coefficient
// rows of stencil_forms.h (one code path instead of a branch per case; the
// rows reproduce the closures bit for bit).
// ---------------------------------------------------------------------------
__constant__ dgr_forms::Form c_forms42[dgr_forms::F_NFORMS] = DGR_FORMS_TABLE;
__constant__ dgr_forms::Form c_forms64[dgr_forms::G_NFORMS] = DGR_FORMS64_TABLE;
constexpr int NFORMS_MAX =
    (int)dgr_forms::F_NFORMS > (int)dgr_forms::G_NFORMS ? (int)dgr_forms::F_NFORMS : (int)dgr_forms::G_NFORMS;

// The variable arrays are global memory; saying so turns generic (flat)
// accesses into global ones, which do not hold up waits on LDS reads.
typedef __attribute__((address_space(1))) double gdouble;
__device__ __forceinline__ const gdouble* gmem(const double* p) { return (const gdouble*)p; }
__device__ __forceinline__ gdouble* gmem(double* p) { return (gdouble*)p; }
__shared__ dgr_forms::Form s_forms[NFORMS_MAX];  // boundary kernels only: the order's table

// Closure row f over the 9-point window centred at A[c], scaled.  Only
// points with a nonzero coefficient are read (a closure never reaches past
// its face, where a derivative plane may end).
__device__ __forceinline__ double eval_form(int f, const double* A, int c, int s, const AxisC& x) {
  const dgr_forms::Form& F = s_forms[f];
  double w[dgr_forms::W];
#pragma unroll
  for (int m = 0; m < dgr_forms::W; ++m) w[m] = F.c[m] != 0.0 ? A[c + (m - 4) * s] : 0.0;
  const double sum = F.desc ? dgr_forms::form_sum_down(F.c, w) : dgr_forms::form_sum_up(F.c, w);
  return dgr_forms::form_scale(sum, F.sel, &x.ih2);
}

// Is q within reach of a flagged face's closures (three points)?
__device__ __forceinline__ bool near_face(int q, const AxisC& x) {
  return (x.lf && q <= 5) || (x.hf && q >= x.e - 3);
}

// Interior first / second derivatives over the window m3..p3 = A[q-3..q+3]:
// deriv42_x / deriv42_xx (BSSN_GR/src/derivs.cpp:29-65, 265-316) or
// deriv64_x / deriv64_xx (QuadGrav/src/derivs.cpp:2161-2170, 2585-2596).
template <int ORD>
__device__ __forceinline__ double d1_int(double m3, double m2, double m1, double p1, double p2, double p3,
                                         const AxisC& x) {
#pragma clang fp contract(off)
  if constexpr (ORD == 64) return (-m3 + 9.0 * m2 - 45.0 * m1 + 45.0 * p1 - 9.0 * p2 + p3) * x.ih60;
  return (m2 - 8.0 * m1 + 8.0 * p1 - p2) * x.ih12;
}
template <int ORD>
__device__ __forceinline__ double d2_int(double m3, double m2, double m1, double z0, double p1, double p2,
                                         double p3, const AxisC& x) {
#pragma clang fp contract(off)
  if constexpr (ORD == 64)
    return (2.0 * m3 - 27.0 * m2 + 270.0 * m1 - 490.0 * z0 + 270.0 * p1 - 27.0 * p2 + 2.0 * p3) * x.isq180;
  return (-m2 + 16.0 * m1 - 30.0 * z0 + 16.0 * p1 - p2) * x.isq12;
}

template <bool BND, int ORD>
__device__ __forceinline__ double d1_op(const double* A, int o, int s, int q, const AxisC& x) {
  const int c = o + q * s;
  const double m2 = A[c - 2 * s], m1 = A[c - s], p1 = A[c + s], p2 = A[c + 2 * s];
  const double m3 = ORD == 64 ? A[c - 3 * s] : 0.0, p3 = ORD == 64 ? A[c + 3 * s] : 0.0;
  double r = d1_int<ORD>(m3, m2, m1, p1, p2, p3, x);
  if (BND && near_face(q, x)) {
    if constexpr (ORD == 64) {
      const int f = dgr_forms::form64_d1(q, x.e, x.lf, x.hf);
      if (f != (int)dgr_forms::G_D1) r = eval_form(f, A, c, s, x);
    } else {
      // the closure rows read q-2 .. q+2 on the interior side of the face
      const int f = dgr_forms::form_d1(q, x.e, x.lf, x.hf);
      if (f != (int)dgr_forms::F_D1) r = dgr_forms::closure42_d1(f, m2, m1, A[c], p1, p2, &x.ih2);
    }
  }
  return r;
}

// (falloff, asymptotic value) of the Sommerfeld terms per variable
// (BSSN_GR/src/rhs.cpp:267-320), VAR order.
__constant__ double c_falloff[24] = {1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2};
__constant__ double c_asym[24] = {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0};

// Variables with second derivatives (bssnrhs_derivs.h).
__host__ __device__ constexpr bool has_d2(int v) {
  return v == dgr::V_ALPHA || v == dgr::V_CHI || (v >= dgr::V_BETA0 && v <= dgr::V_BETA2) ||
         (v >= dgr::V_SG0 && v <= dgr::V_SG5);
}
// Variables whose first derivatives enter the equations.
__host__ __device__ constexpr bool has_d1(int v) {
  return has_d2(v) || v == dgr::V_K || (v >= dgr::V_GT0 && v <= dgr::V_GT2);
}

// Processing order: the metric first (Christoffels), then chi, alpha (the
// Christoffels die here), Gt (closes the Ricci tensor), beta, K, then the
// variables that only need advection and dissipation.
constexpr int h_order[24] = {dgr::V_SG0, dgr::V_SG1, dgr::V_SG2, dgr::V_SG3, dgr::V_SG4, dgr::V_SG5,
                             dgr::V_CHI, dgr::V_ALPHA, dgr::V_GT0, dgr::V_GT1, dgr::V_GT2,
                             dgr::V_BETA0, dgr::V_BETA1, dgr::V_BETA2, dgr::V_K, dgr::V_B0,
                             dgr::V_B1, dgr::V_B2, dgr::V_SA0, dgr::V_SA1, dgr::V_SA2, dgr::V_SA3,
                             dgr::V_SA4, dgr::V_SA5};

struct TileGeom {
  uint64_t off;       // block offset
  int sz0, sz1, sz2;
  int i0, j0, k0, nx, ny, nz;
  int rx, ry, rz;     // staged region extent (n+6)
  int nreg;
};

// Source offset (from the block's first padded point) of every staging slot
// of this thread and a bit per slot that lies inside the tile's region.
// Computed once per tile; every variable's prefetch then costs one address
// add per slot.  Gather mode (RK4 stages): a ghost slot with a same-level
// neighbour reads the neighbour's interior, sd[direction] being the index
// distance to that source (0 for interior slots and for ghosts without a
// neighbour, which keep their own padding): the same-level refill done on
// the fly, so the RK4 step needs no refill kernel between stages.
struct Slots {
  int32_t off[NPF];
  uint32_t valid;
};

// In gather mode sd holds the plan's distance per ghost direction (27,
// slot 13 = the block itself).  With stage_in the input is an RK4 stage
// buffer: a nonzero distance is taken from the compact index the point
// would have in its own block (its source is an interior point, in the
// buffer's compact region), a zero one keeps the block's own padding.
template <bool GATHER>
__device__ __forceinline__ Slots make_slots(const TileGeom& T, int t, const long long* sd, bool stage_in) {
  Slots S;
  S.valid = 0;
#pragma unroll
  for (int m = 0; m < NPF; ++m) {
    const int e = t + m * NT;
    const int a = e % SX;
    const int b = (e / SX) % SY;
    const int c = e / (SX * SY);
    const int i = T.i0 - 3 + a, j = T.j0 - 3 + b, k = T.k0 - 3 + c;
    long long o = (long long)i + (long long)T.sz0 * (j + (long long)T.sz1 * k);
    if constexpr (GATHER) {
      const int dx = i < 3 ? 0 : (i >= T.sz0 - 3 ? 2 : 1);
      const int dy = j < 3 ? 0 : (j >= T.sz1 - 3 ? 2 : 1);
      const int dz = k < 3 ? 0 : (k >= T.sz2 - 3 ? 2 : 1);
      const long long d = sd[dx + 3 * dy + 9 * dz];
      if (stage_in && d > 0)
        o = (long long)(i - 3) + (long long)(T.sz0 - 6) * ((j - 3) + (long long)(T.sz1 - 6) * (k - 3)) + d;
      else
        o += d;
    }
    S.off[m] = (int32_t)o;
    if (e < NSU && a < T.rx && b < T.ry && c < T.rz) S.valid |= 1u << m;
  }
  return S;
}

// Issue the global loads of one variable's staged region into registers.
// The region is walked in the fixed SX x SY x SZ layout of the LDS buffer;
// slots outside a partial tile's region are skipped and their LDS slots are
// never read.
// Plain inputs are read once (non-temporal loads); in gather mode (CACHED)
// ghost slots read neighbours' interiors that the neighbours' own tiles,
// scheduled on the same L2, read too, so those loads are ordinary.
template <bool CACHED = false>
__device__ __forceinline__ void prefetch(double (&r)[NPF], const double* __restrict__ g, const TileGeom& T,
                                         const Slots& S) {
  const double* base = g + T.off;
#pragma unroll
  for (int m = 0; m < NPF; ++m)
    if ((S.valid >> m) & 1u) {
      if constexpr (CACHED) r[m] = gmem(base)[S.off[m]];
      else r[m] = __builtin_nontemporal_load(gmem(base) + S.off[m]);
    }
}

// zmask (boundary tiles): bit m set when slot m lies beyond a flagged face.
// Those padding values are read only by interior formulas that the face
// closures then replace; staging them as 0 keeps every value the closure
// sums touch finite, whatever the caller left in the padding.
__device__ __forceinline__ void stage(const double (&r)[NPF], double* su, const TileGeom& T, int t, uint32_t zmask) {
#pragma unroll
  for (int m = 0; m < NPF; ++m) {
    const int e = t + m * NT;
    if (e < NSU) su[e] = ((zmask >> m) & 1u) ? 0.0 : r[m];
  }
}

// Slots of this thread beyond a flagged face (boundary tiles).
__device__ __forceinline__ uint32_t beyond_face_mask(const TileGeom& T, unsigned bflag, int t) {
  uint32_t z = 0;
#pragma unroll
  for (int m = 0; m < NPF; ++m) {
    const int e = t + m * NT;
    const int i = T.i0 - 3 + e % SX, j = T.j0 - 3 + (e / SX) % SY, k = T.k0 - 3 + e / (SX * SY);
    const bool out = ((bflag & 1u) && i < 3) || ((bflag & 2u) && i >= T.sz0 - 3) || ((bflag & 4u) && j < 3) ||
                     ((bflag & 8u) && j >= T.sz1 - 3) || ((bflag & 16u) && k < 3) || ((bflag & 32u) && k >= T.sz2 - 3);
    z |= (uint32_t)out << m;
  }
  return z;
}

// Per-thread context of one tile.
struct Ctx {
  bool act, on_face;
  unsigned bflag;
  int ti, tj, tk, gi, gj, gk;
  double eta, sigma;
  uint64_t o;   // global index of the point
  uint64_t oo;  // index of its right-hand side (= o, or the compact slot in RK4 stages)
  const AxisC* ax;  // per-axis constants of the block (LDS)
  const double* geo;  // pmin[3], h[3] of the block (LDS)
  dgr::Params P;
};

// Coordinates of block point (gi, gj, gk) and its distance from the origin
// (rhs.cpp:208-213: x = pmin + i hx).  Recomputed where needed rather than
// kept live across the variable loop.
__device__ __forceinline__ double point_coords(const double* geo, int gi, int gj, int gk, double& x, double& y,
                                               double& z) {
#pragma clang fp contract(off)
  x = geo[0] + gi * geo[3];
  y = geo[1] + gj * geo[4];
  z = geo[2] + gk * geo[5];
  return sqrt(x * x + y * y + z * z);
}

// (ETA_R0 / r)^ETA_DAMPING_EXP of the damped eta (rhs.cpp:219-223).  The
// default exponent 1 needs no pow (pow(x, 1) is x exactly); any other goes
// through an out-of-line pow, so that its polynomial constants are not
// hoisted into registers across the tile loop (where they were spilled and
// reloaded behind the staging loads).
__device__ __noinline__ double eta_pow(double x, double e) { return pow(x, e); }
__device__ __forceinline__ double eta_damping(double x, double e) { return e == 1.0 ? x : eta_pow(x, e); }

// rhs = eq + (KO), or (Sommerfeld + KO) on flagged faces (rhs.cpp:265-373)
__device__ __forceinline__ double rhs_value(const Ctx& C, double eq, double ex) {
#pragma clang fp contract(off)
  return C.on_face ? ex : (eq + ex);
}
// Outputs that become final several variables later park their dissipation
// (+ Sommerfeld) term in the output array; the variable that completes them
// loads the parked terms back (process_var).

// Right-hand sides a variable completes, stored after the next variable is
// staged (run_vars): a store still in flight when the staging waits for its
// prefetched registers would be waited for as well, since loads and stores
// share one counter and the wait for an older load cannot count past a
// younger store.  Per processed variable V, the output arrays and values
// (NOUT of them); parked terms are plain stores, final values nontemporal.
template <int V>
struct Outs {
  static constexpr int N = V == dgr::V_BETA2 ? 8 : (V == dgr::V_K ? 4 : 1);
  static constexpr bool PARK = (V >= dgr::V_SG0 && V <= dgr::V_SG5) || V == dgr::V_CHI ||
                               (V >= dgr::V_GT0 && V <= dgr::V_GT2);
  static __device__ __forceinline__ constexpr int var(int k) {
    return k == 0 ? V : (V == dgr::V_BETA2 ? (k < 7 ? dgr::V_SG0 + (k - 1) : dgr::V_CHI) : dgr::V_GT0 + (k - 1));
  }
  double val[N];
};
// Output k of variable V: kept for store_outs (deriv42), or stored where it
// is computed (the 6th-order bodies, whose register allocation prefers it).
template <int ORD, int V>
__device__ __forceinline__ void put_out(const IOPtrs* __restrict__ io, const Ctx& C, Outs<V>& o, int k, double v) {
  o.val[k] = v;
  if constexpr (ORD != 42) {
    if constexpr (Outs<V>::PARK) gmem(io->r.v[V])[C.oo] = v;
    else __builtin_nontemporal_store(v, gmem(io->r.v[Outs<V>::var(k)]) + C.oo);
  }
}
template <int V>
__device__ __forceinline__ void store_outs(const IOPtrs* __restrict__ io, const Ctx& C, const Outs<V>& o) {
  if (!C.act) return;
  if constexpr (Outs<V>::PARK) {
    gmem(io->r.v[V])[C.oo] = o.val[0];
  } else {
#pragma unroll
    for (int k = 0; k < Outs<V>::N; ++k)
      __builtin_nontemporal_store(o.val[k], gmem(io->r.v[Outs<V>::var(k)]) + C.oo);
  }
}

// Per-thread LDS slots: element i of a thread's array sits at p[i * NPT]
// (one slot column per tile point).  The constraint monitor keeps each
// array in slots of its own:
constexpr int NSLOT = 48;  // C2 18, R 6, gt_rhs 6, At_rhs 6, small accumulators 12
template <int N>
struct LdsArr {
  double* p;
  __device__ __forceinline__ double& operator[](int i) { return p[i * NPT]; }
  __device__ __forceinline__ const double& operator[](int i) const { return p[i * NPT]; }
};
struct LdsStore {
  template <int N>
  using A = LdsArr<N>;
};
using KPoint = dgr::PointT<LdsStore>;

// The right-hand side packs the arrays by lifetime into 41 slots, which
// leaves room for a second staging buffer.  In the processing order below
// the Christoffels (C2, slots 0-17) die with alpha; the accumulators that
// start later reuse their slots: the -lambda3 beta.dGt part of B_rhs (from
// Gt, slots 0-2), d_j d_k beta^k (3-5) and div beta (6), the last two
// cleared when beta^0 starts.  R 18-23, gt_rhs 24-29, At_rhs 30-35,
// Gt_rhs 36-38, chi_rhs 39, K_rhs 40.
constexpr int NSLOT_RHS = 41;
struct AccArr {  // the 12 small accumulators (PointT::A_*) in their slots
  double* p;     // slot 0 of the thread
  static __device__ __forceinline__ constexpr int slot(int i) {
    return i < 3 ? 36 + i : (i < 10 ? i - 3 : 29 + i);
  }
  __device__ __forceinline__ double& operator[](int i) { return p[slot(i) * NPT]; }
  __device__ __forceinline__ const double& operator[](int i) const { return p[slot(i) * NPT]; }
};
struct RhsStore {
  template <int N>
  using A = typename std::conditional<N == 12, AccArr, LdsArr<N>>::type;
};
using KPointR = dgr::PointT<RhsStore>;
static_assert(AccArr::slot(KPointR::A_GT) == 36 && AccArr::slot(KPointR::A_B) == 0 &&
                  AccArr::slot(KPointR::A_WB) == 3 && AccArr::slot(KPointR::A_DIVB) == 6 &&
                  AccArr::slot(KPointR::A_CHI) == 39 && AccArr::slot(KPointR::A_K) == 40,
              "accumulator slots");

// All operators along one axis through the point: first derivative (if
// D1), second derivative (if D2), upwinded derivative and Kreiss-Oliger
// term.  In interior blocks the 7-point line is read from LDS once into
// registers and every operator is evaluated from it with the reference's
// operation order (derivs.cpp); in boundary blocks the closure-aware forms
// above are used.
template <bool BND, bool D1, bool D2, int ORD>
__device__ __forceinline__ void axis_ops(const double* A, int o, int s, int q, const AxisC& x, double beta,
                                         double& d1, double& d2, double& ad, double& ko) {
#pragma clang fp contract(off)
  const int c = o + q * s;
  const double m3 = A[c - 3 * s], m2 = A[c - 2 * s], m1 = A[c - s], z0 = A[c], p1 = A[c + s], p2 = A[c + 2 * s],
               p3 = A[c + 3 * s];
  if constexpr (D1) d1 = d1_int<ORD>(m3, m2, m1, p1, p2, p3, x);
  if constexpr (D2) d2 = d2_int<ORD>(m3, m2, m1, z0, p1, p2, p3, x);
  if constexpr (ORD == 64) {
    // deriv64adv is the centred deriv64 (QuadGrav/src/derivs.cpp:2913-2928)
    ad = D1 ? d1 : d1_int<ORD>(m3, m2, m1, p1, p2, p3, x);
  } else if (beta > 0.0) {
    ad = (-3.0 * m1 - 10.0 * z0 + 18.0 * p1 - 6.0 * p2 + p3) * x.ih12;
  } else {
    ad = (-m3 + 6.0 * m2 - 18.0 * m1 + 10.0 * z0 + 3.0 * p1) * x.ih12;
  }
  ko = x.pre * (-m3 + 6.0 * m2 - 15.0 * m1 + 20.0 * z0 - 15.0 * p1 + 6.0 * p2 - p3);
  if constexpr (BND && ORD == 42) {
    // closures: the written-out deriv42 rows over the register window, for
    // each operator whose row differs from the interior one at this point
    if (near_face(q, x)) {
      using namespace dgr_forms;
      const double* sc = &x.ih2;
      if constexpr (D1) {
        const int f = form_d1(q, x.e, x.lf, x.hf);
        if (f != F_D1) d1 = closure42_d1(f, m2, m1, z0, p1, p2, sc);
      }
      if constexpr (D2) {
        const int f = form_d2(q, x.e, x.lf, x.hf);
        if (f != F_D2) d2 = closure42_d2(f, m3, m2, m1, z0, p1, p2, p3, sc);
      }
      const int fa = form_adv(q, x.e, x.lf, x.hf, beta > 0.0);
      if (fa != F_ADV_P && fa != F_ADV_N) ad = closure42_adv(fa, m3, m2, m1, z0, p1, p2, p3, sc);
      const int fk = form_ko(q, x.e, x.lf, x.hf);
      if (fk != F_KO) ko = closure42_ko(fk, m3, m2, m1, z0, p1, p2, p3, sc);
    }
  } else if constexpr (BND) {
    // 6th-order closures: one generic row evaluation per operator whose form
    // differs from the interior one at this point
    if (near_face(q, x)) {
      using namespace dgr_forms;
#pragma unroll 1
      for (int op = D1 ? 0 : (D2 ? 1 : 2); op < 4; ++op) {
        if (op == 1 && !D2) continue;
        const int f = op == 0 || op == 2 ? form64_d1(q, x.e, x.lf, x.hf)
                    : op == 1            ? form64_d2(q, x.e, x.lf, x.hf)
                                         : form64_ko(q, x.e, x.lf, x.hf);
        if (f == G_D1 || f == G_D2 || f == G_KO) continue;
        // the closure rows stay within the register window (+-4 for the
        // 6th-order rows); staged values are finite, so no masking
        // (q-4 is read only on the high side and q+4 only on the low side,
        // where they lie inside the block's interior)
        double w[dgr_forms::W];
        w[0] = (ORD == 64 && q >= x.e - 3) ? A[c - 4 * s] : 0.0;
        w[1] = m3; w[2] = m2; w[3] = m1; w[4] = z0; w[5] = p1; w[6] = p2; w[7] = p3;
        w[8] = (ORD == 64 && q <= 5) ? A[c + 4 * s] : 0.0;
        const Form& F = s_forms[f];
        double sum;
        {
#pragma clang fp contract(off)
          double up = 0.0, dn = 0.0;
#pragma unroll
          for (int m = 0; m < dgr_forms::W; ++m) up = up + F.c[m] * w[m];
#pragma unroll
          for (int m = dgr_forms::W - 1; m >= 0; --m) dn = dn + F.c[m] * w[m];
          sum = F.desc ? dn : up;
        }
        const double r = form_scale(sum, F.sel, &x.ih2);
        if (op == 0) d1 = r;
        else if (op == 1) d2 = r;
        else if (op == 2) ad = r;
        else ko = r;
      }
    }
  }
}

// Mixed-derivative inputs of one point, for the 6th-order path whose two
// planes are built one after the other in a shared buffer.
struct Mixed {
  double d1x, d1y, dxy, dxz, dyz;
};

template <int V, int FM, int ORD>
__device__ __forceinline__ void process_var(KPointR& s, Ctx& C, const double* su, const double* sgx,
                                            const double* sgy, const Mixed& M, const TileGeom& T,
                                            const IOPtrs* __restrict__ io, Outs<V>& out) {
  using namespace dgr;
  if (!C.act) return;
  // parked terms of the outputs this variable completes (gt_rhs and
  // chi_rhs at beta^2, Gt_rhs at K), all loaded here at once: emitted one
  // by one, each load would wait behind the previous emit's store, since
  // the compiler cannot tell the output arrays apart
  constexpr int NPK = V == V_BETA2 ? 7 : (V == V_K ? 3 : 0);
  double pk[NPK > 0 ? NPK : 1];
  if constexpr (V == V_BETA2) {
#pragma unroll
    for (int p = 0; p < 6; ++p) pk[p] = gmem(io->r.v[V_SG0 + p])[C.oo];
    pk[6] = gmem(io->r.v[V_CHI])[C.oo];
  }
  if constexpr (V == V_K) {
#pragma unroll
    for (int k = 0; k < 3; ++k) pk[k] = gmem(io->r.v[V_GT0 + k])[C.oo];
  }
  // B^k only enters b_rhs^k and B_rhs^k
  if constexpr (V >= V_BETA0 && V <= V_BETA2) s.B[V - V_BETA0] = gmem(io->u.v[V_B0 + (V - V_BETA0)])[C.o];
  if constexpr (V == V_BETA0) {
    // the shift accumulators start here, in slots the Christoffels used
#pragma unroll
    for (int j = 0; j < 3; ++j) s.acc[KPointR::A_WB + j] = 0.0;
    s.acc[KPointR::A_DIVB] = 0.0;
  }
  const AxisC* ax = C.ax;
  const int ti = C.ti, tj = C.tj, tk = C.tk, gi = C.gi, gj = C.gj, gk = C.gk;
  // su index of this point; lines through it along x, y, z are addressed
  // as su[o + q*stride] with q the block coordinate along the axis
  const int c = (ti + 3) + SX * ((tj + 3) + SY * (tk + 3));
  const int ox = c - gi, oy = c - gj * SX, oz = c - gk * (SX * SY);
  constexpr bool D2 = has_d2(V);
  // FM: flagged-face axes of the tile's block (bit 0 x, 1 y, 2 z); the
  // closure rows are compiled in along those axes only
  constexpr bool BND = FM != 0, CX = (FM & 1) != 0, CY = (FM & 2) != 0, CZ = (FM & 4) != 0;
  // first derivatives along x and y come from the LDS planes for D2 vars
  constexpr bool D1xy = !D2 && (has_d1(V) || BND);
  constexpr bool D1z = has_d1(V) || BND;
  double d1[3] = {0.0, 0.0, 0.0};
  double d2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double av[3], kv[3];
  axis_ops<CX, D1xy, D2, ORD>(su, ox, 1, gi, ax[0], s.beta[0], d1[0], d2[0], av[0], kv[0]);
  __builtin_amdgcn_sched_barrier(0);
  axis_ops<CY, D1xy, D2, ORD>(su, oy, SX, gj, ax[1], s.beta[1], d1[1], d2[3], av[1], kv[1]);
  __builtin_amdgcn_sched_barrier(0);
  axis_ops<CZ, D1z, D2, ORD>(su, oz, SX * SY, gk, ax[2], s.beta[2], d1[2], d2[5], av[2], kv[2]);
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (D2 && ORD == 64) {
    d1[0] = M.d1x;
    d1[1] = M.d1y;
    d2[1] = M.dxy;
    d2[2] = M.dxz;
    d2[4] = M.dyz;
  } else if constexpr (D2) {
    constexpr int GXY = Planes<ORD>::GXY;
    // x and y derivatives at the point are read back from the LDS planes
    d1[0] = sgx[ti + TX * ((tj + 2) + GXY * (tk + 2))];
    d1[1] = sgy[ti + TX * (tj + TY * (tk + 2))];
    // mixed derivatives by composition: d_y(d_x u), d_z(d_x u), d_z(d_y u).
    // sgx[a + TX*(b + GXY*cc)] holds d_x u at block coords (i0+a, j0-2+b, k0-2+cc);
    // sgy[a + TX*(b + TY*cc)] holds d_y u at (i0+a, j0+b, k0-2+cc).
    const int oxy = ti + TX * ((2 - T.j0) + GXY * (tk + 2));
    d2[1] = d1_op<CY, ORD>(sgx, oxy, TX, gj, ax[1]);
    const int oxz = ti + TX * ((tj + 2) + GXY * (2 - T.k0));
    d2[2] = d1_op<CZ, ORD>(sgx, oxz, TX * GXY, gk, ax[2]);
    const int oyz = ti + TX * (tj + TY * (2 - T.k0));
    d2[4] = d1_op<CZ, ORD>(sgy, oyz, TX * TY, gk, ax[2]);
    __builtin_amdgcn_sched_barrier(0);
  }
  // upwinded derivatives contracted with beta, summed in the order of the
  // reference's printed expression: beta0*a0 + beta1*a1 + beta2*a2
  const double adv = s.beta[0] * av[0] + s.beta[1] * av[1] + s.beta[2] * av[2];
  // Kreiss-Oliger sum, preceded on flagged faces by the Sommerfeld value
  double ex;
  {
#pragma clang fp contract(off)
    const double k0 = kv[0], k1 = kv[1], k2 = kv[2];
    const double ko = C.sigma * (k0 + k1 + k2);
    if (BND && C.on_face) {
      const double f = su[c];
      double x, y, z;
      const double inv_r = 1.0 / point_coords(C.geo, C.gi, C.gj, C.gk, x, y, z);
      const double bc = -inv_r * (x * d1[0] + y * d1[1] + z * d1[2] + c_falloff[V] * (f - c_asym[V]));
      ex = bc + ko;
    } else {
      ex = ko;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  // feed the staged equations; every right-hand side that is final (and the
  // dissipation terms parked in the output array) goes to `out`
  if constexpr (V >= V_SG0 && V <= V_SG5) {
    s.add_gt(V - V_SG0, d1, d2, adv);
    put_out<ORD>(io, C, out, 0, ex);
  }
  if constexpr (V == V_CHI) {
    s.add_chi(d1, d2, adv);
    put_out<ORD>(io, C, out, 0, ex);
  }
  if constexpr (V == V_ALPHA) put_out<ORD>(io, C, out, 0, rhs_value(C, s.add_alpha(C.P, d1, d2, adv), ex));
  if constexpr (V >= V_GT0 && V <= V_GT2) {
    s.add_Gt(C.P, V - V_GT0, d1, adv);
    put_out<ORD>(io, C, out, 0, ex);
  }
  if constexpr (V >= V_BETA0 && V <= V_BETA2) {
    double gto[6], chio = 0.0;
    put_out<ORD>(io, C, out, 0, rhs_value(C, s.add_beta(C.P, V - V_BETA0, d1, d2, adv, gto, &chio), ex));
    if constexpr (V == V_BETA2) {
#pragma unroll
      for (int p = 0; p < 6; ++p) put_out<ORD>(io, C, out, 1 + p, rhs_value(C, gto[p], pk[p]));
      put_out<ORD>(io, C, out, 7, rhs_value(C, chio, pk[6]));
    }
  }
  if constexpr (V == V_K) {
    put_out<ORD>(io, C, out, 0, rhs_value(C, s.add_K(d1, adv), ex));
#pragma unroll
    for (int k = 0; k < 3; ++k) put_out<ORD>(io, C, out, 1 + k, rhs_value(C, s.Gt_out(k), pk[k]));
  }
  if constexpr (V >= V_B0 && V <= V_B2) {
    s.B[V - V_B0] = su[c];
    put_out<ORD>(io, C, out, 0, rhs_value(C, s.add_B(C.P, V - V_B0, adv, C.eta), ex));
  }
  if constexpr (V >= V_SA0 && V <= V_SA5) put_out<ORD>(io, C, out, 0, rhs_value(C, s.add_At(V - V_SA0, adv), ex));
}

// Widened first-derivative planes for the mixed derivatives, walked in
// their fixed LDS shapes (compile-time divisors) and masked to the tile.
// Widened first-derivative planes for the mixed derivatives.  Thread t
// computes plane entries t, t + NT, ... of each plane; in interior tiles
// their staging offsets and axis coordinates, which depend only on the
// tile, are computed once per tile (PlaneSlots), so each variable's planes
// cost the stencils alone.
// sgx[a + TX*(b + GXY*cc)]: d_x u at block coords (i0+a, j0-GW+b, k0-GW+cc),
// su coords (a+3, b+3-GW, cc+3-GW).
// sgy[a + TX*(b + TY*cc)]: d_y u at (i0+a, j0+b, k0-GW+cc), su coords
// (a+3, b+3, cc+3-GW).
template <int ORD>
struct PlaneSlots {
  static constexpr int NPL =
      ((Planes<ORD>::NGX > Planes<ORD>::NGY ? Planes<ORD>::NGX : Planes<ORD>::NGY) + NT - 1) / NT;
  int ox[NPL], qx[NPL], oy[NPL], qy[NPL];
  uint32_t valid;  // bit k: x entry k, bit NPL + k: y entry k
};

template <int ORD>
__device__ __forceinline__ PlaneSlots<ORD> make_plane_slots(const TileGeom& T, int t) {
  using G = Planes<ORD>;
  constexpr int NPL = PlaneSlots<ORD>::NPL;
  PlaneSlots<ORD> P;
  P.valid = 0;
#pragma unroll
  for (int k = 0; k < NPL; ++k) {
    int e = t + k * NT;
    int a = e % TX, b = (e / TX) % G::GXY, cc = e / (TX * G::GXY);
    P.ox[k] = (a + 3) + SX * ((b + 3 - G::GW) + SY * (cc + 3 - G::GW)) - (T.i0 + a);
    P.qx[k] = T.i0 + a;
    if (e < G::NGX && a < T.nx && b < T.ny + 2 * G::GW && cc < T.nz + 2 * G::GW) P.valid |= 1u << k;
    b = (e / TX) % TY;
    cc = e / (TX * TY);
    P.oy[k] = (a + 3) + SX * ((b + 3) + SY * (cc + 3 - G::GW)) - (T.j0 + b) * SX;
    P.qy[k] = T.j0 + b;
    if (e < G::NGY && a < T.nx && b < T.ny && cc < T.nz + 2 * G::GW) P.valid |= 1u << (NPL + k);
  }
  return P;
}

// Boundary tiles (LOOP; their closure code leaves fewer registers)
// recompute the entry coordinates in the loop instead of holding them.  CL:
// closure rows along the plane's derivative axis.
template <bool LOOP, bool CL, int ORD>
__device__ __forceinline__ void plane_x(const double* su, double* sgx, const PlaneSlots<ORD>& P, const TileGeom& T,
                                        const AxisC* ax, int t) {
  using G = Planes<ORD>;
  if constexpr (LOOP) {
    for (int e = t; e < G::NGX; e += NT) {
      const int a = e % TX;
      const int b = (e / TX) % G::GXY;
      const int cc = e / (TX * G::GXY);
      if (a < T.nx && b < T.ny + 2 * G::GW && cc < T.nz + 2 * G::GW) {
        const int o = (a + 3) + SX * ((b + 3 - G::GW) + SY * (cc + 3 - G::GW)) - (T.i0 + a);
        sgx[e] = d1_op<CL, ORD>(su, o, 1, T.i0 + a, ax[0]);
      }
    }
  } else {
    constexpr int NPL = PlaneSlots<ORD>::NPL;
#pragma unroll
    for (int k = 0; k < NPL; ++k)
      if ((P.valid >> k) & 1u) sgx[t + k * NT] = d1_op<CL, ORD>(su, P.ox[k], 1, P.qx[k], ax[0]);
  }
}
template <bool LOOP, bool CL, int ORD>
__device__ __forceinline__ void plane_y(const double* su, double* sgy, const PlaneSlots<ORD>& P, const TileGeom& T,
                                        const AxisC* ax, int t) {
  using G = Planes<ORD>;
  if constexpr (LOOP) {
    for (int e = t; e < G::NGY; e += NT) {
      const int a = e % TX;
      const int b = (e / TX) % TY;
      const int cc = e / (TX * TY);
      if (a < T.nx && b < T.ny && cc < T.nz + 2 * G::GW) {
        const int o = (a + 3) + SX * ((b + 3) + SY * (cc + 3 - G::GW)) - (T.j0 + b) * SX;
        sgy[e] = d1_op<CL, ORD>(su, o, SX, T.j0 + b, ax[1]);
      }
    }
  } else {
    constexpr int NPL = PlaneSlots<ORD>::NPL;
#pragma unroll
    for (int k = 0; k < NPL; ++k)
      if ((P.valid >> (NPL + k)) & 1u) sgy[t + k * NT] = d1_op<CL, ORD>(su, P.oy[k], SX, P.qy[k], ax[1]);
  }
}

// Variables are staged through two LDS buffers: variable I is read from
// one while every thread writes variable I+1 (loaded into registers during
// variable I) into the other at the end of its own work on I, so one
// barrier per variable (plus one after the derivative planes) suffices.
template <int I, int FM, int ORD, bool GATHER>
__device__ __forceinline__ void run_vars(KPointR& s, Ctx& C, double (&pf)[NPF], double* su0, double* su1,
                                         double* sgx, double* sgy, const TileGeom& T,
                                         const IOPtrs* __restrict__ io, int t, const Slots& sl,
                                         const PlaneSlots<ORD>& ps, uint32_t zmask) {
  if constexpr (I < 24) {
    constexpr int V = h_order[I];
    using G = Planes<ORD>;
    constexpr bool BND = FM != 0, CX = (FM & 1) != 0, CY = (FM & 2) != 0, CZ = (FM & 4) != 0;
    double* const su = (I & 1) ? su1 : su0;  // variable I
    double* const sn = (I & 1) ? su0 : su1;  // variable I+1, written below
    // variable I is staged by every thread, nobody reads variable I-1 any
    // more (its buffer takes I+1), the planes of the last variable are read
    __syncthreads();
    Mixed M;
    if constexpr (has_d2(V) && G::SHARED) {
      // d_x plane -> d_x u, d_y d_x u, d_z d_x u; then the d_y plane in the
      // same buffer -> d_y u, d_z d_y u
      plane_x<BND, CX, ORD>(su, sgx, ps, T, C.ax, t);
      __syncthreads();
      if (C.act) {
        const int ti = C.ti, tj = C.tj, tk = C.tk;
        M.d1x = sgx[ti + TX * ((tj + G::GW) + G::GXY * (tk + G::GW))];
        M.dxy = d1_op<CY, ORD>(sgx, ti + TX * ((G::GW - T.j0) + G::GXY * (tk + G::GW)), TX, C.gj, C.ax[1]);
        M.dxz = d1_op<CZ, ORD>(sgx, ti + TX * ((tj + G::GW) + G::GXY * (G::GW - T.k0)), TX * G::GXY, C.gk, C.ax[2]);
      }
      __syncthreads();
      plane_y<BND, CY, ORD>(su, sgx, ps, T, C.ax, t);
      __syncthreads();
      if (C.act) {
        const int ti = C.ti, tj = C.tj, tk = C.tk;
        M.d1y = sgx[ti + TX * (tj + TY * (tk + G::GW))];
        M.dyz = d1_op<CZ, ORD>(sgx, ti + TX * (tj + TY * (G::GW - T.k0)), TX * TY, C.gk, C.ax[2]);
      }
    } else if constexpr (has_d2(V)) {
      plane_x<BND, CX, ORD>(su, sgx, ps, T, C.ax, t);
      plane_y<BND, CY, ORD>(su, sgy, ps, T, C.ax, t);
      __syncthreads();
    }
    if constexpr (V == dgr::V_CHI) {
      if (C.act) s.finish_metric();
    }
    Outs<V> out;
    process_var<V, FM, ORD>(s, C, su, sgx, sgy, M, T, io, out);
    // deriv42: the next variable into the other buffer, then this
    // variable's stores, then the loads of the variable after next.  The
    // 6th-order bodies store each output where it is computed (put_out):
    // their larger register windows spill more when the outputs stay live
    // across the staging (measured).
    if constexpr (I + 1 < 24) stage(pf, sn, T, t, zmask);
    if constexpr (ORD == 42) store_outs<V>(io, C, out);
    if constexpr (I + 2 < 24) prefetch<GATHER>(pf, io->u.v[h_order[I + 2]], T, sl);
    run_vars<I + 1, FM, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, io, t, sl, ps, zmask);
  }
}

// FM: 0 = tiles of blocks without flagged faces; FM_FACES = boundary tiles,
// each run through the body compiled for its block's flagged-face axes: bit
// 0/1/2 = closure tests along x/y/z, so 1, 2 or 4 for one axis, 7 for the
// edge and corner blocks (the bodies are run_vars<0, 1|2|4|7, ...>).
constexpr int FM_FACES = 8;

// A tile descriptor read as four 32-bit words (one scalar load of uniform
// data) and unpacked.
__device__ __forceinline__ DevTile load_tile(const DevTile* p) {
  static_assert(sizeof(DevTile) == 16, "tile descriptor is four words");
  const uint32_t* q = reinterpret_cast<const uint32_t*>(p);
  const uint32_t w0 = q[0], w1 = q[1], w2 = q[2], w3 = q[3];
  DevTile t;
  t.block = w0;
  t.i0 = (uint16_t)(w1 & 0xffffu);
  t.j0 = (uint16_t)(w1 >> 16);
  t.k0 = (uint16_t)(w2 & 0xffffu);
  t.nx = (uint8_t)((w2 >> 16) & 0xffu);
  t.ny = (uint8_t)(w2 >> 24);
  t.nz = (uint8_t)(w3 & 0xffu);
  t.pad = (uint8_t)((w3 >> 8) & 0xffu);
  return t;
}

template <int FM, int ORD, bool GATHER>
__global__ void __launch_bounds__(NT) k_bssn_rhs(const DevBlock* __restrict__ blocks,
                                                 const DevTile* __restrict__ tiles, int tile_begin, int ntiles,
                                                 const IOPtrs* __restrict__ io, DevParams P,
                                                 const long long* __restrict__ delta,
                                                 const uint32_t* __restrict__ slot, int stage_in) {
  __shared__ double sstage[2][NSU];  // two staging buffers (variables alternate)
  double* const su0 = sstage[0];
  double* const su1 = sstage[1];
  __shared__ double splanes[Planes<ORD>::N];
  double* const sgx = splanes;
  double* const sgy = Planes<ORD>::SHARED ? splanes : splanes + Planes<ORD>::NGX;
  __shared__ double sacc[NSLOT_RHS * NPT];
  __shared__ IOPtrs s_io;           // array pointers of this launch
  // pmin[3], h[3] and the per-axis stencil constants of the current block,
  // copied from the block table as one run of 8-byte words
  constexpr int NBW = (int)((offsetof(DevBlock, ax) + sizeof(AxisC) * 3 - offsetof(DevBlock, pmin)) / 8);
  static_assert(offsetof(DevBlock, h) == offsetof(DevBlock, pmin) + 24 &&
                    offsetof(DevBlock, ax) == offsetof(DevBlock, h) + 24 && sizeof(AxisC) % 8 == 0,
                "block geometry and axis constants are contiguous");
  __shared__ unsigned long long s_blk[NBW];
  double* const s_geo = reinterpret_cast<double*>(s_blk);
  AxisC* const s_ax = reinterpret_cast<AxisC*>(s_blk + 6);
  __shared__ long long s_delta[GATHER ? NNBR : 1];  // gather mode: source distance per ghost direction

  constexpr bool BND = FM != 0;
  const int t = threadIdx.x;
  // the array pointers are the same for every tile of the launch
  if (t < 24) s_io.u.v[t] = io->u.v[t];
  else if (t < 48) s_io.r.v[t - 24] = io->r.v[t - 24];
  if constexpr (BND) {
    // closure rows for the boundary stencils (visible after the first
    // barrier of the tile loop)
    constexpr int nf = ORD == 64 ? dgr_forms::G_NFORMS : dgr_forms::F_NFORMS;
    constexpr int nd = nf * (int)(sizeof(dgr_forms::Form) / sizeof(double));
    static_assert(sizeof(dgr_forms::Form) % sizeof(double) == 0, "Form is a whole number of doubles");
    const double* src = reinterpret_cast<const double*>(ORD == 64 ? c_forms64 : c_forms42);
    for (int i = t; i < nd; i += NT) reinterpret_cast<double*>(s_forms)[i] = src[i];
  }
  Ctx C;
  // point of this thread in the tile, in the bank-conflict-free order of
  // tile_map.h; threads beyond the tile's 343 points stay inactive
  if (t < NPT) {
    const int m = c_tile_map[t];
    C.ti = m & 7;
    C.tj = (m >> 3) & 7;
    C.tk = m >> 6;
  } else {
    C.ti = C.tj = C.tk = TX;
  }
  for (int i = 0; i < 4; ++i) C.P.lambda[i] = P.lambda[i];
  C.P.lambda_f[0] = P.lambda_f[0];
  C.P.lambda_f[1] = P.lambda_f[1];
  C.sigma = P.ko_sigma;

  __syncthreads();  // s_io (and the closure rows) visible
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const DevTile tl = load_tile(tiles + tile_begin + w);
    const DevBlock& B = blocks[tl.block];
    TileGeom T;
    T.off = B.offset;
    T.sz0 = B.sz[0]; T.sz1 = B.sz[1]; T.sz2 = B.sz[2];
    T.i0 = tl.i0; T.j0 = tl.j0; T.k0 = tl.k0;
    T.nx = tl.nx; T.ny = tl.ny; T.nz = tl.nz;
    T.rx = T.nx + 6; T.ry = T.ny + 6; T.rz = T.nz + 6;
    T.nreg = T.rx * T.ry * T.rz;

    if constexpr (FM == FM_FACES) {
      // a tile touching flagged faces (tl.pad != 0) takes the thread map of
      // those faces: its closure points fill the first waves
      if (t < NPT) {
        const int m = tl.pad ? c_face_map[tl.pad - 1][t] : c_tile_map[t];
        C.ti = m & 7;
        C.tj = (m >> 3) & 7;
        C.tk = m >> 6;
      }
    }
    C.act = t < NPT && C.ti < T.nx && C.tj < T.ny && C.tk < T.nz;
    C.gi = T.i0 + C.ti; C.gj = T.j0 + C.tj; C.gk = T.k0 + C.tk;
    C.bflag = B.bflag;
    // one load per thread, all in flight together
    if (t < NBW) s_blk[t] = reinterpret_cast<const unsigned long long*>(&B.pmin[0])[t];
    C.ax = s_ax;
    C.geo = s_geo;

    if constexpr (GATHER) {
      // source distance of each ghost direction: the same point of the
      // same-level neighbour's interior (corners sz-7 apart), or 0
      // (plan table, computed on the host from the neighbour links)
      if (t < NNBR) s_delta[t] = delta[(size_t)tl.block * NNBR + t];
      __syncthreads();
    }
    const bool from_stage = GATHER && stage_in;
    const Slots sl = make_slots<GATHER>(T, t, s_delta, from_stage);
    double pf[NPF];
    prefetch<GATHER>(pf, io->u.v[h_order[0]], T, sl);
    KPointR s;
    // the point's own values: padded index, or its compact slot in a stage buffer
    C.o = from_stage ? T.off + (uint64_t)((long long)(C.gi - 3) +
                                   (long long)(T.sz0 - 6) * ((C.gj - 3) + (long long)(T.sz1 - 6) * (C.gk - 3)) +
                                   s_delta[13])
              : T.off + (uint64_t)C.gi + (uint64_t)T.sz0 * (C.gj + (uint64_t)T.sz1 * C.gk);
    if (C.act) {
      // the state at the point, in flight with the first variable's loads
      const uint64_t o = C.o;
      s.alpha = gmem(s_io.u.v[dgr::V_ALPHA])[o];
      s.chi = gmem(s_io.u.v[dgr::V_CHI])[o];
      s.K = gmem(s_io.u.v[dgr::V_K])[o];
#pragma unroll
      for (int i = 0; i < 3; ++i) s.beta[i] = gmem(s_io.u.v[dgr::V_BETA0 + i])[o];
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        s.gt[p] = gmem(s_io.u.v[dgr::V_SG0 + p])[o];
        s.At[p] = gmem(s_io.u.v[dgr::V_SA0 + p])[o];
      }
    }
    __syncthreads();

    // point coordinates, eta (rhs.cpp:214-223) and the Sommerfeld face test
    // inactive threads (t >= NPT) never touch their slots
    s.C2.p = sacc + t;
    s.R.p = sacc + 18 * NPT + t;
    s.gt_rhs.p = sacc + 24 * NPT + t;
    s.At_rhs.p = sacc + 30 * NPT + t;
    s.acc.p = sacc + t;
    // RK4 stages (gather) write k in the compact tile-major layout of the
    // stage kernels: slot of the tile in block order x 343 + local index
    C.oo = GATHER ? (uint64_t)slot[tile_begin + w] * NPT + (uint64_t)(C.ti + TX * (C.tj + TY * C.tk)) : C.o;
    C.on_face = false;
    C.eta = P.eta_const;
    if (C.act) {
#pragma clang fp contract(off)
      double x, y, z;
      const double geo[6] = {B.pmin[0], B.pmin[1], B.pmin[2], B.h[0], B.h[1], B.h[2]};
      const double rr = point_coords(geo, C.gi, C.gj, C.gk, x, y, z);
      if (rr >= P.eta_r0) C.eta *= eta_damping(P.eta_r0 / rr, P.eta_damping_exp);
      const unsigned bf = C.bflag;
      C.on_face = ((bf & 1u) && C.gi == 3) || ((bf & 2u) && C.gi == T.sz0 - 4) || ((bf & 4u) && C.gj == 3) ||
                  ((bf & 8u) && C.gj == T.sz1 - 4) || ((bf & 16u) && C.gk == 3) ||
                  ((bf & 32u) && C.gk == T.sz2 - 4);
      s.begin();
    }
    const uint32_t zmask = BND ? beyond_face_mask(T, B.bflag, t) : 0u;
    PlaneSlots<ORD> ps;
    if constexpr (!BND) ps = make_plane_slots<ORD>(T, t);
    // the first variable into buffer 0, the second on its way
    stage(pf, su0, T, t, zmask);
    prefetch<GATHER>(pf, io->u.v[h_order[1]], T, sl);
    if constexpr (FM == FM_FACES) {
      // most boundary blocks have flagged faces along one axis only: their
      // tiles skip the closure tests along the other two
      const unsigned bf = C.bflag;
      const int fm = __builtin_amdgcn_readfirstlane(((bf & 3u) ? 1 : 0) | ((bf & 12u) ? 2 : 0) | ((bf & 48u) ? 4 : 0));
      if (fm == 1) run_vars<0, 1, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, &s_io, t, sl, ps, zmask);
      else if (fm == 2) run_vars<0, 2, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, &s_io, t, sl, ps, zmask);
      else if (fm == 4) run_vars<0, 4, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, &s_io, t, sl, ps, zmask);
      else run_vars<0, 7, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, &s_io, t, sl, ps, zmask);
    } else {
      run_vars<0, FM, ORD, GATHER>(s, C, pf, su0, su1, sgx, sgy, T, &s_io, t, sl, ps, zmask);
    }
    // the next tile rewrites the shared block constants and buffer 0
    __syncthreads();
  }
}


// ---------------------------------------------------------------------------
// Constraint monitor: physical_constraints (BSSN_GR/src/physcon.cpp:12-96)
// ---------------------------------------------------------------------------
// Hamiltonian and momentum constraints and the Weyl scalar psi4 on the
// interior points of every block.  Derivatives are deriv42 whatever the
// evolution's stencil order (BSSN_GR/include/physcon.h:8-14): first
// derivatives of 17 variables and second derivatives of chi and the
// conformal metric, mixed ones by composition (BSSN_GR/scripts/
// constraint_derivs.h), with the face closures on flagged faces.  The
// equations are those of GR/rhs_scripts/bssn/constraints.py.  The Ricci
// tensor and the Christoffels come from the staged point machinery of the
// right-hand side (bssn_staged.h); the momentum and psi4 terms that are
// linear in the derivatives of At are summed as each At component is
// processed, the rest is evaluated at the end.

// Processing order: metric (Christoffels), chi (closes them), Gt (Ricci), K, At.
constexpr int con_order[17] = {dgr::V_SG0, dgr::V_SG1, dgr::V_SG2, dgr::V_SG3, dgr::V_SG4, dgr::V_SG5,
                               dgr::V_CHI, dgr::V_GT0, dgr::V_GT1, dgr::V_GT2, dgr::V_K,
                               dgr::V_SA0, dgr::V_SA1, dgr::V_SA2, dgr::V_SA3, dgr::V_SA4, dgr::V_SA5};
__host__ __device__ constexpr bool con_d2(int v) {
  return v == dgr::V_CHI || (v >= dgr::V_SG0 && v <= dgr::V_SG5);
}

// Per-point constraint sums kept in registers.
struct ConAcc {
  double r[3], mr[3], mi[3];  // tetrad: unit radial vector, m = (mr + i mi) (constraints.py)
  double mom[3];              // sum_jk g^jk d_k At_ij
  double p3r, p3i;            // derivative parts of psi4_3 (before the 1/chi)
  double Gt[3], dK[3];
  bool axis;                  // within 1e-7 of the z axis: psi4 = 0 (physconeqs.cpp:1039-1044)
};

// Tetrad by Gram-Schmidt in the physical metric gt/chi (constraints.py).
__device__ __forceinline__ void con_tetrad(const KPoint& s, double x, double y, double z, ConAcc& A) {
  const double ichi = 1.0 / s.chi;
  auto dot = [&](const double* a, const double* b) {
    double v = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) v += s.gt[dgr::sp(i, j)] * ichi * a[i] * b[j];
    return v;
  };
  double r[3] = {x, y, z};
  double th[3] = {x * z, y * z, -(x * x + y * y)};
  double ph[3] = {-y, x, 0.0};
  const double nr = sqrt(dot(r, r));
  for (int i = 0; i < 3; ++i) r[i] /= nr;
  {
    const double p1 = dot(th, th), p2 = dot(r, th);
    for (int i = 0; i < 3; ++i) th[i] -= p2 * r[i];
    const double n = sqrt(p1 - p2 * p2);
    for (int i = 0; i < 3; ++i) th[i] /= n;
  }
  {
    const double p1 = dot(ph, ph), p2 = dot(r, ph), p3 = dot(th, ph);
    for (int i = 0; i < 3; ++i) ph[i] -= p2 * r[i] + p3 * th[i];
    const double n = sqrt(p1 - p2 * p2 - p3 * p3);
    for (int i = 0; i < 3; ++i) ph[i] /= n;
  }
  const double is2 = 0.7071067811865475244;
  for (int i = 0; i < 3; ++i) {
    A.r[i] = r[i];
    A.mr[i] = th[i] * is2;
    A.mi[i] = ph[i] * is2;
  }
}

template <int V, bool BND>
__device__ __forceinline__ void process_con(KPoint& s, ConAcc& A, const Ctx& C, const double* su,
                                            const double* sgx, const double* sgy, const TileGeom& T) {
  using namespace dgr;
  if (!C.act) return;
  const AxisC* ax = C.ax;
  const int ti = C.ti, tj = C.tj, tk = C.tk, gi = C.gi, gj = C.gj, gk = C.gk;
  const int c = (ti + 3) + SX * ((tj + 3) + SY * (tk + 3));
  const int ox = c - gi, oy = c - gj * SX, oz = c - gk * (SX * SY);
  constexpr bool D2 = con_d2(V);
  double d1[3] = {0.0, 0.0, 0.0};
  double d2[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double av, kv;  // unused operators of axis_ops
  axis_ops<BND, !D2, D2, 42>(su, ox, 1, gi, ax[0], 0.0, d1[0], d2[0], av, kv);
  axis_ops<BND, !D2, D2, 42>(su, oy, SX, gj, ax[1], 0.0, d1[1], d2[3], av, kv);
  axis_ops<BND, true, D2, 42>(su, oz, SX * SY, gk, ax[2], 0.0, d1[2], d2[5], av, kv);
  if constexpr (D2) {
    constexpr int GXY = Planes<42>::GXY;
    d1[0] = sgx[ti + TX * ((tj + 2) + GXY * (tk + 2))];
    d1[1] = sgy[ti + TX * (tj + TY * (tk + 2))];
    d2[1] = d1_op<BND, 42>(sgx, ti + TX * ((2 - T.j0) + GXY * (tk + 2)), TX, gj, ax[1]);
    d2[2] = d1_op<BND, 42>(sgx, ti + TX * ((tj + 2) + GXY * (2 - T.k0)), TX * GXY, gk, ax[2]);
    d2[4] = d1_op<BND, 42>(sgy, ti + TX * (tj + TY * (2 - T.k0)), TX * TY, gk, ax[2]);
  }
  if constexpr (V >= V_SG0 && V <= V_SG5) s.add_gt(V - V_SG0, d1, d2, 0.0);
  if constexpr (V == V_CHI) s.add_chi(d1, d2, 0.0);
  if constexpr (V >= V_GT0 && V <= V_GT2) {
    s.add_Gt(C.P, V - V_GT0, d1, 0.0);
    A.Gt[V - V_GT0] = su[c];
  }
  if constexpr (V == V_K) {
    for (int i = 0; i < 3; ++i) A.dK[i] = d1[i];
  }
  if constexpr (V >= V_SA0 && V <= V_SA5) {
    constexpr int p = V - V_SA0, a = spi(p), b = spj(p);
    // sum_jk g^jk d_k At_ij for (i, j) = (a, b) and (b, a)
    A.mom[a] += s.igt[sp(b, 0)] * d1[0] + s.igt[sp(b, 1)] * d1[1] + s.igt[sp(b, 2)] * d1[2];
    if constexpr (a != b) A.mom[b] += s.igt[sp(a, 0)] * d1[0] + s.igt[sp(a, 1)] * d1[1] + s.igt[sp(a, 2)] * d1[2];
    // psi4_3: U_ij, V_ij hold m^k d_j At_ki, for (k, i) = (a, b) and (b, a)
#pragma unroll
    for (int h = 0; h < (a != b ? 2 : 1); ++h) {
      const int k = h ? b : a, i = h ? a : b;
      for (int j = 0; j < 3; ++j) {
        const double MRij = A.mr[i] * A.r[j] - A.mr[j] * A.r[i];
        const double NRij = A.mi[i] * A.r[j] - A.mi[j] * A.r[i];
        A.p3r += (MRij * A.mr[k] - NRij * A.mi[k]) * d1[j];
        A.p3i += (NRij * A.mr[k] + MRij * A.mi[k]) * d1[j];
      }
    }
  }
}

template <int I, bool BND>
__device__ __forceinline__ void con_vars(KPoint& s, ConAcc& A, Ctx& C, double (&pf)[NPF], double* su, double* sgx,
                                         double* sgy, const TileGeom& T, const IOPtrs* __restrict__ io, int t,
                                         const Slots& sl, const PlaneSlots<42>& ps, uint32_t zmask) {
  if constexpr (I < 17) {
    constexpr int V = con_order[I];
    stage(pf, su, T, t, zmask);
    __syncthreads();
    if constexpr (con_d2(V)) {
      plane_x<BND, BND, 42>(su, sgx, ps, T, C.ax, t);
      plane_y<BND, BND, 42>(su, sgy, ps, T, C.ax, t);
    }
    if constexpr (I + 1 < 17) prefetch(pf, io->u.v[con_order[I + 1]], T, sl);
    if constexpr (con_d2(V)) __syncthreads();
    if constexpr (V == dgr::V_CHI) {
      if (C.act) s.finish_metric();
    }
    process_con<V, BND>(s, A, C, su, sgx, sgy, T);
    __syncthreads();
    con_vars<I + 1, BND>(s, A, C, pf, su, sgx, sgy, T, io, t, sl, ps, zmask);
  }
}

// Ham, mom, psi4 from the completed sums (constraints.py).
__device__ __forceinline__ void con_finish(const KPoint& s, const ConAcc& A, double out[6]) {
  using namespace dgr;
  const double chi = s.chi, ichi = 1.0 / chi, K = s.K;
  double Rm[6], trR = 0.0, sqrA = 0.0;
  for (int p = 0; p < 6; ++p) {
    Rm[p] = s.R[p];
    trR += spw(p) * s.igt[p] * Rm[p];
    sqrA += spw(p) * s.At[p] * s.AUU[p];
  }
  out[0] = chi * trR - sqrA + (2.0 / 3.0) * K * K;
  // Gamma^m_ki from the slots: C2[m*6 + sp(k, i)]
  double G[3][6];
  for (int m = 0; m < 3; ++m)
    for (int p = 0; p < 6; ++p) G[m][p] = s.C2[m * 6 + p];
  for (int i = 0; i < 3; ++i) {
    double cc = 0.0, gA = 0.0, dc = 0.0;
    for (int j = 0; j < 3; ++j) {
      for (int k = 0; k < 3; ++k) {
        double g = 0.0;
        for (int m = 0; m < 3; ++m) g += G[m][sp(k, i)] * s.At[sp(j, m)];
        cc += s.igt[sp(j, k)] * g;
        dc += s.igt[sp(j, k)] * s.At[sp(k, i)] * s.dchi[j];
      }
      gA += A.Gt[j] * s.At[sp(i, j)];
    }
    out[1 + i] = A.mom[i] - cc - gA - 1.5 * dc * ichi - (2.0 / 3.0) * A.dK[i];
  }
  if (A.axis) {
    out[4] = 0.0;
    out[5] = 0.0;
    return;
  }
  const double* r = A.r;
  const double* mr = A.mr;
  const double* mi = A.mi;
  double Av[3];
  for (int i = 0; i < 3; ++i) Av[i] = s.At[sp(i, 0)] * r[0] + s.At[sp(i, 1)] * r[1] + s.At[sp(i, 2)] * r[2];
  double p1r = 0.0, p1i = 0.0, p2r = 0.0, p2i = 0.0, p3r = A.p3r, p3i = A.p3i;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      const double MM = mr[i] * mr[j] - mi[i] * mi[j];
      const double NN = mr[i] * mi[j] + mr[j] * mi[i];
      p1r += Rm[sp(i, j)] * MM;
      p1i += Rm[sp(i, j)] * NN;
      p2r += s.At[sp(i, j)] * MM;
      p2i += s.At[sp(i, j)] * NN;
      // non-derivative part of U_ij, V_ij: m^k Gamma^m_ki At_mj
      double cu = 0.0, cv = 0.0;
      for (int k = 0; k < 3; ++k) {
        double g = 0.0;
        for (int m = 0; m < 3; ++m) g += G[m][sp(k, i)] * s.At[sp(m, j)];
        cu += mr[k] * g;
        cv += mi[k] * g;
      }
      const double MRij = mr[i] * r[j] - mr[j] * r[i];
      const double NRij = mi[i] * r[j] - mi[j] * r[i];
      p3r += MRij * cu - NRij * cv;
      p3i += NRij * cu + MRij * cv;
    }
  double rdchi = 0.0, rA = 0.0, mrdchi = 0.0, midchi = 0.0, mrA = 0.0, miA = 0.0;
  for (int i = 0; i < 3; ++i) {
    rdchi += r[i] * s.dchi[i];
    rA += Av[i] * r[i];
    mrdchi += mr[i] * s.dchi[i];
    midchi += mi[i] * s.dchi[i];
    mrA += mr[i] * Av[i];
    miA += mi[i] * Av[i];
  }
  const double Atmp = ichi * ichi * (rA + K * chi / 3.0 + 0.5 * rdchi);
  p2r *= Atmp;
  p2i *= Atmp;
  p3r *= ichi;
  p3i *= ichi;
  const double p4r = ichi * ichi * (mrA * (mrA + 0.5 * mrdchi) - miA * (miA + 0.5 * midchi));
  const double p4i = ichi * ichi * (mrA * (miA - 0.5 * midchi) + miA * (mrA - 0.5 * mrdchi));
  out[4] = p1r + p2r - p3r - p4r;
  out[5] = -(p1i + p2i - p3i - p4i);
}

template <bool BND>
__global__ void __launch_bounds__(NT) k_physcon(const DevBlock* __restrict__ blocks,
                                                const DevTile* __restrict__ tiles, int tile_begin, int ntiles,
                                                const IOPtrs* __restrict__ io, DevParams P) {
  __shared__ double su[NSU];
  __shared__ double splanes[Planes<42>::N];
  double* const sgx = splanes;
  double* const sgy = splanes + Planes<42>::NGX;
  __shared__ double sacc[NSLOT * NPT];
  __shared__ IOPtrs s_io;      // u: state, r[0..5]: ham, mom0..2, psi4_real, psi4_img
  __shared__ AxisC s_ax[3];
  __shared__ double s_geo[6];

  const int t = threadIdx.x;
  Ctx C;
  if (t < NPT) {
    const int m = c_tile_map[t];
    C.ti = m & 7;
    C.tj = (m >> 3) & 7;
    C.tk = m >> 6;
  } else {
    C.ti = C.tj = C.tk = TX;
  }
  for (int i = 0; i < 4; ++i) C.P.lambda[i] = P.lambda[i];
  C.P.lambda_f[0] = P.lambda_f[0];
  C.P.lambda_f[1] = P.lambda_f[1];
  C.sigma = 0.0;
  C.eta = 0.0;
  C.on_face = false;

  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const DevTile tl = tiles[tile_begin + w];
    const DevBlock& B = blocks[tl.block];
    TileGeom T;
    T.off = B.offset;
    T.sz0 = B.sz[0]; T.sz1 = B.sz[1]; T.sz2 = B.sz[2];
    T.i0 = tl.i0; T.j0 = tl.j0; T.k0 = tl.k0;
    T.nx = tl.nx; T.ny = tl.ny; T.nz = tl.nz;
    T.rx = T.nx + 6; T.ry = T.ny + 6; T.rz = T.nz + 6;
    T.nreg = T.rx * T.ry * T.rz;
    C.act = t < NPT && C.ti < T.nx && C.tj < T.ny && C.tk < T.nz;
    C.gi = T.i0 + C.ti; C.gj = T.j0 + C.tj; C.gk = T.k0 + C.tk;
    C.bflag = B.bflag;
    if (t < 24) s_io.u.v[t] = io->u.v[t];
    else if (t < 30) s_io.r.v[t - 24] = io->r.v[t - 24];
    else if (t < 33) s_ax[t - 30] = B.ax[t - 30];
    else if (t < 36) s_geo[t - 33] = B.pmin[t - 33];
    else if (t < 39) s_geo[t - 33] = B.h[t - 36];
    C.ax = s_ax;
    C.geo = s_geo;
    const Slots sl = make_slots<false>(T, t, nullptr, false);
    double pf[NPF];
    prefetch(pf, io->u.v[con_order[0]], T, sl);
    __syncthreads();

    KPoint s;
    s.C2.p = sacc + t;
    s.R.p = sacc + 18 * NPT + t;
    s.gt_rhs.p = sacc + 24 * NPT + t;
    s.At_rhs.p = sacc + 30 * NPT + t;
    s.acc.p = sacc + 36 * NPT + t;
    C.o = T.off + (uint64_t)C.gi + (uint64_t)T.sz0 * (C.gj + (uint64_t)T.sz1 * C.gk);
    ConAcc A;
    if (C.act) {
      const uint64_t o = C.o;
      s.alpha = 1.0;
      s.chi = gmem(s_io.u.v[dgr::V_CHI])[o];
      s.K = gmem(s_io.u.v[dgr::V_K])[o];
#pragma unroll
      for (int i = 0; i < 3; ++i) s.beta[i] = 0.0;
#pragma unroll
      for (int p = 0; p < 6; ++p) {
        s.gt[p] = gmem(s_io.u.v[dgr::V_SG0 + p])[o];
        s.At[p] = gmem(s_io.u.v[dgr::V_SA0 + p])[o];
      }
      s.begin();
      double x, y, z;
      point_coords(C.geo, C.gi, C.gj, C.gk, x, y, z);
      A.axis = fabs(x) <= 1e-7 && fabs(y) <= 1e-7;
      if (!A.axis) con_tetrad(s, x, y, z, A);
      for (int i = 0; i < 3; ++i) A.mom[i] = 0.0;
      A.p3r = A.p3i = 0.0;
    }
    const uint32_t zmask = BND ? beyond_face_mask(T, B.bflag, t) : 0u;
    PlaneSlots<42> ps;
    if constexpr (!BND) ps = make_plane_slots<42>(T, t);
    con_vars<0, BND>(s, A, C, pf, su, sgx, sgy, T, &s_io, t, sl, ps, zmask);
    if (C.act) {
      double out[6];
      con_finish(s, A, out);
#pragma unroll
      for (int q = 0; q < 6; ++q) gmem(s_io.r.v[q])[C.o] = out[q];
    }
    __syncthreads();  // s_io, s_ax, s_geo are rewritten by the next tile
  }
}

// Masked constraint norms (computeConstraintL2Norm, BSSN_GR/include/
// grUtils.tcc:40-88): over the block-interior points where the mask is at
// least the threshold, each node once (a point on a block's high face is
// left to a same-level neighbour on that side, which holds it on its low
// face).  One workgroup per tile writes 13 partials: sum of squares and max
// |.| of the 6 outputs, and the point count.
struct Ptr6 {
  const double* v[6];
};
constexpr int NORM_T = 256;

// Each unique node once: only points that own their node (the zip's owner
// flags) enter the sums, as computeConstraintL2Norm sums over the zipped
// vector (BSSN_GR/include/grUtils.tcc:40-88).
__global__ void __launch_bounds__(NORM_T) k_con_norm_partial(const DevBlock* __restrict__ blocks,
                                                             const DevTile* __restrict__ tiles, int ntiles,
                                                             const uint8_t* __restrict__ own, Ptr6 con,
                                                             const double* __restrict__ mask, double thr,
                                                             double* __restrict__ part) {
  __shared__ double red[NORM_T];
  const int t = threadIdx.x;
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const DevTile tl = tiles[w];
    const DevBlock& B = blocks[tl.block];
    double sq[6] = {0, 0, 0, 0, 0, 0}, mx[6] = {0, 0, 0, 0, 0, 0}, cnt = 0.0;
    const int np = tl.nx * tl.ny * tl.nz;
    for (int e = t; e < np; e += NORM_T) {
      const int i = tl.i0 + e % tl.nx, j = tl.j0 + (e / tl.nx) % tl.ny, k = tl.k0 + e / (tl.nx * tl.ny);
      const uint64_t o = B.offset + (uint64_t)i + (uint64_t)B.sz[0] * (j + (uint64_t)B.sz[1] * k);
      if (!own[o] || mask[o] < thr) continue;
      cnt += 1.0;
      for (int q = 0; q < 6; ++q) {
        const double v = con.v[q][o];
        sq[q] += v * v;
        mx[q] = fmax(mx[q], fabs(v));
      }
    }
    // fixed-order tree reductions (deterministic)
    for (int q = 0; q < 13; ++q) {
      red[t] = q < 6 ? sq[q] : (q < 12 ? mx[q - 6] : cnt);
      __syncthreads();
      for (int h = NORM_T / 2; h > 0; h >>= 1) {
        if (t < h) red[t] = q >= 6 && q < 12 ? fmax(red[t], red[t + h]) : red[t] + red[t + h];
        __syncthreads();
      }
      if (t == 0) part[(size_t)w * 13 + q] = red[0];
      __syncthreads();
    }
  }
}

// Mesh::zip (include/mesh.tcc:3775-3822): node k of the zipped vector takes
// the value at padded index idx[k] (its owner's copy).
__global__ void __launch_bounds__(256) k_zip(const uint64_t* __restrict__ idx, uint64_t n,
                                             const double* __restrict__ in, double* __restrict__ out) {
  for (uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x; k < n; k += (uint64_t)gridDim.x * 256)
    out[k] = in[idx[k]];
}

// Sum of the per-tile partials; out[13] = L2[6], max[6], count.  One
// workgroup per output value: thread t takes tiles t, t + NORM_T, ... in
// order, then a fixed tree combines the threads (deterministic).
__global__ void __launch_bounds__(NORM_T) k_con_norm_final(const double* __restrict__ part, int ntiles,
                                                           double* __restrict__ out) {
  __shared__ double red[NORM_T];
  const int q = blockIdx.x, t = threadIdx.x;
  const bool mx = q >= 6 && q < 12;
  double a = 0.0;
  for (int w = t; w < ntiles; w += NORM_T) {
    const double v = part[(size_t)w * 13 + q];
    a = mx ? fmax(a, v) : a + v;
  }
  red[t] = a;
  __syncthreads();
  for (int h = NORM_T / 2; h > 0; h >>= 1) {
    if (t < h) red[t] = mx ? fmax(red[t], red[t + h]) : red[t] + red[t + h];
    __syncthreads();
  }
  if (t == 0) out[q] = red[0];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {
thread_local char g_errbuf[256];
}  // namespace

int dgr_internal::hip_fail(hipError_t e, const char* what) {
  snprintf(g_errbuf, sizeof(g_errbuf), "%s: %s", what, hipGetErrorString(e));
  return DGR_E_HIP;
}

namespace {

void free_plan(dgr_plan_s* plan) {
  (void)hipFree(plan->d_blocks);
  (void)hipFree(plan->d_tiles);
  (void)hipFree(plan->d_tiles_blk);
  (void)hipFree(plan->d_tile_slot);
  (void)hipFree(plan->d_io);
  (void)hipFree(plan->d_nbr);
  (void)hipFree(plan->d_delta);
  (void)hipFree(plan->d_cdelta);
  (void)hipFree(plan->d_fill_items);
  (void)hipFree(plan->d_fill_groups);
  (void)hipFree(plan->d_fill_copies);
  (void)hipFree(plan->d_io_con);
  (void)hipFree(plan->d_norm);
  (void)hipFree(plan->d_tiles_sub);
  (void)hipFree(plan->d_io_blk);
  (void)hipFree(plan->d_flag);
  (void)hipFree(plan->d_own);
  (void)hipFree(plan->d_zidx);
  if (plan->ev_fork) (void)hipEventDestroy(plan->ev_fork);
  if (plan->ev_join) (void)hipEventDestroy(plan->ev_join);
  if (plan->aux) (void)hipStreamDestroy(plan->aux);
  delete plan;
}

// Thread map of a boundary tile (DevTile::pad): 0 = the default map, else
// 1 + index of the face map of the flagged faces that have closure points
// (3 or fewer points from the face) in the tile, sx + 3 sy + 9 sz - 1 with
// s = 0 (none), 1 (low face), 2 (high face) per axis.  A face map applies
// when, along every axis with such a face, the tile spans 7 points and
// begins (low face) or ends (high face) at the block's interior edge, and no
// axis has closure points of both faces.
uint8_t face_map_id(const dgr_block& bl, const DevTile& t) {
  const int q0[3] = {t.i0, t.j0, t.k0}, n[3] = {t.nx, t.ny, t.nz};
  int idx = 0, mul = 1;
  for (int a = 0; a < 3; ++a, mul *= 3) {
    const int sz = (int)bl.sz[a];
    const bool lo = ((bl.bflag >> (2 * a)) & 1u) && q0[a] <= 5;
    const bool hi = ((bl.bflag >> (2 * a + 1)) & 1u) && q0[a] + n[a] - 1 >= sz - 6;
    if (lo && hi) return 0;
    if (lo) {
      if (n[a] != 7 || q0[a] != 3) return 0;
      idx += mul;
    } else if (hi) {
      if (n[a] != 7 || q0[a] + 6 != sz - 4) return 0;
      idx += 2 * mul;
    }
  }
  return (uint8_t)idx;
}

// Stencil constants of one axis, with the operations of derivs.cpp
// (idx = 1/dx; idx_by_2 = 0.5*idx; idx_by_12 = idx/12; idx_sqrd = 1/(dx*dx);
// pre_factor_6_dx = -1/64/dx; smr3 = 59/48*64*dx, ...).
AxisC host_axis(double h, int sz, unsigned bflag, int a) {
  AxisC c;
  const double ih = 1.0 / h;
  c.ih2 = 0.5 * ih;
  c.ih12 = ih / 12.0;
  c.isq = 1.0 / (h * h);
  c.isq12 = c.isq / 12.0;
  c.pre = -1.0 / 64.0 / h;
  c.smr3 = 59.0 / 48.0 * 64 * h;
  c.smr2 = 43.0 / 48.0 * 64 * h;
  c.smr1 = 49.0 / 48.0 * 64 * h;
  c.ih60 = ih / 60.0;        // QuadGrav/src/derivs.cpp:2147
  c.isq180 = c.isq / 180.0;  // QuadGrav/src/derivs.cpp:2574
  c.e = sz - 3;
  c.lf = (bflag >> (2 * a)) & 1u;
  c.hf = (bflag >> (2 * a + 1)) & 1u;
  c.pad = 0;
  return c;
}

// Device descriptor of a block: offset, sizes, spacing and per-axis stencil
// constants.
void make_devblock(const dgr_block& bl, DevBlock& db) {
  db.offset = bl.offset;
  for (int d = 0; d < 3; ++d) {
    db.sz[d] = bl.sz[d];
    db.pmin[d] = bl.pmin[d];
    db.h[d] = (bl.pmax[d] - bl.pmin[d]) / (bl.sz[d] - 1);  // rhs.cpp:148-150
  }
  for (int d = 0; d < 3; ++d) db.ax[d] = host_axis(db.h[d], (int)bl.sz[d], bl.bflag, d);
  db.bflag = bl.bflag;
  db.group = bl.group;
  db.cbase = 0;
}

// Same-level neighbour table: blocks of equal size and spacing whose
// corners are sz-7 points apart along an axis are neighbours.  Spacings are
// compared by a rounded log2 (equal up to rounding); corners on the integer
// lattice (pmin - org)/h of their class (size, spacing, group), org being
// the corner of the class's first block.  A block whose corner lies off that
// lattice by more than 1e-6 of a spacing gets no link and is no one's
// neighbour: such blocks overlap their lattice neighbours, and a ghost copied
// from one would come from a different place.
void build_neighbours(const dgr_block* blocks, const std::vector<DevBlock>& hb, uint32_t nblocks,
                      std::vector<int32_t>& nbr) {
  using Cls = std::tuple<uint32_t, uint32_t, uint32_t, long long, long long, long long, uint32_t>;
  using Key = std::tuple<Cls, long long, long long, long long>;
  auto lev = [](double h) { return llround(std::log2(h) * 64.0); };
  auto cls = [&](uint32_t b) {
    const DevBlock& B = hb[b];
    return Cls(B.sz[0], B.sz[1], B.sz[2], lev(B.h[0]), lev(B.h[1]), lev(B.h[2]), B.group);
  };
  std::map<Cls, uint32_t> first;
  for (uint32_t b = 0; b < nblocks; ++b) first.emplace(cls(b), b);
  // lattice coordinates of block b's corner; false when off the lattice
  auto lattice = [&](uint32_t b, long long K[3]) {
    const DevBlock& B = hb[b];
    const DevBlock& O = hb[first[cls(b)]];
    for (int a = 0; a < 3; ++a) {
      const double s = (B.pmin[a] - O.pmin[a]) / B.h[a];
      K[a] = llround(s);
      if (!(std::fabs(s - (double)K[a]) < 1e-6)) return false;
    }
    return true;
  };
  std::map<Key, int32_t> pos;
  std::vector<char> on(nblocks, 0);
  for (uint32_t b = 0; b < nblocks; ++b) {
    long long K[3];
    if (!lattice(b, K)) continue;
    on[b] = 1;
    pos.emplace(Key(cls(b), K[0], K[1], K[2]), (int32_t)b);
  }
  nbr.assign((size_t)nblocks * NNBR, -1);
  for (uint32_t b = 0; b < nblocks; ++b) {
    if (!on[b]) continue;
    const DevBlock& B = hb[b];
    long long K[3];
    lattice(b, K);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          if (!dx && !dy && !dz) continue;
          const Key k(cls(b), K[0] + (long long)dx * (B.sz[0] - 7), K[1] + (long long)dy * (B.sz[1] - 7),
                      K[2] + (long long)dz * (B.sz[2] - 7));
          auto it = pos.find(k);
          if (it != pos.end()) nbr[(size_t)b * NNBR + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] = it->second;
        }
  }
  (void)blocks;
}

// Split n interior points into chunks of at most 7, as even as possible.
// Cut n interior points into the fewest near-equal runs of at most tmax.
void split(int n, int tmax, std::vector<std::pair<int, int>>& out) {
  out.clear();
  const int parts = (n + tmax - 1) / tmax;
  int start = 3;
  for (int p = 0; p < parts; ++p) {
    const int len = n / parts + (p < n % parts ? 1 : 0);
    out.emplace_back(start, len);
    start += len;
  }
}

}  // namespace

extern "C" const char* dgr_error_string(int code) {
  switch (code) {
    case DGR_OK: return "ok";
    case DGR_E_INVALID: return "invalid argument";
    case DGR_E_HIP: return g_errbuf[0] ? g_errbuf : "HIP error";
    case DGR_E_NOMEM: return "device allocation failed";
    case DGR_E_UNSUPPORTED: return "unsupported configuration";
    case DGR_E_CONSTRAINT: return "constraint projection failed";
    default: return "unknown error";
  }
}

namespace {
int noalloc_init(int dev);
}  // namespace

extern "C" int dgr_init(int device) {
  int n = 0;
  DGR_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return DGR_E_INVALID;
  DGR_HIP(hipSetDevice(device));
  // set-up of the no-allocation entry point on this device
  return noalloc_init(device);
}

extern "C" void dgr_schedule_default(dgr_schedule* s) {
  if (!s) return;
  for (int c = 0; c < DGR_SIZE_CLASSES; ++c) s->tile_max[c] = TX;
  s->order = DGR_ORDER_BLOCK;
  s->grid_mult = 1;
}

namespace {
int size_class(const dgr_block& b) {
  const uint32_t m = std::max(b.sz[0], std::max(b.sz[1], b.sz[2]));
  return m <= 13 ? 0 : m <= 19 ? 1 : m <= 25 ? 2 : 3;
}

// Scheduled order of the tiles [begin, end) of `blk` (block order, a block's
// tiles contiguous).  XCD-aware: workgroup w of a launch runs on XCD w % 8
// and the persistent loop hands it tiles w, w + grid, ...; so a tile at
// position t runs on XCD t % 8 whenever the grid is a multiple of 8.  Blocks
// are dealt to 8 queues (fewest tiles first) and the queues interleaved,
// which keeps each block's tiles, and the halo they share, on one L2.
void order_tiles(std::vector<DevTile>& t, size_t begin, size_t end, uint32_t order, const dgr_block* blocks) {
  if (order == DGR_ORDER_LARGEST) {
    std::stable_sort(t.begin() + begin, t.begin() + end, [&](const DevTile& a, const DevTile& b) {
      const dgr_block &A = blocks[a.block], &B = blocks[b.block];
      return (uint64_t)A.sz[0] * A.sz[1] * A.sz[2] > (uint64_t)B.sz[0] * B.sz[1] * B.sz[2];
    });
  } else if (order == DGR_ORDER_XCD || order == DGR_ORDER_SFC) {
    constexpr int NXCD = 8;
    std::vector<std::vector<DevTile>> q(NXCD);
    // runs of one block's tiles
    std::vector<std::pair<size_t, size_t>> runs;
    for (size_t i = begin; i < end;) {
      size_t j = i;
      while (j < end && t[j].block == t[i].block) ++j;
      runs.emplace_back(i, j);
      i = j;
    }
    if (order == DGR_ORDER_SFC) {
      // Morton order of the block corners on the finest spacing present,
      // then 8 contiguous runs of about equal tile counts
      double hmin = 1e300, lo[3] = {1e300, 1e300, 1e300};
      for (auto& r : runs) {
        const dgr_block& B = blocks[t[r.first].block];
        for (int d = 0; d < 3; ++d) {
          hmin = std::min(hmin, (B.pmax[d] - B.pmin[d]) / (B.sz[d] - 1));
          lo[d] = std::min(lo[d], B.pmin[d]);
        }
      }
      auto key = [&](const std::pair<size_t, size_t>& r) {
        const dgr_block& B = blocks[t[r.first].block];
        uint64_t c[3], k = 0;
        for (int d = 0; d < 3; ++d) c[d] = (uint64_t)std::llround((B.pmin[d] - lo[d]) / hmin) & 0x1fffff;
        for (int bit = 20; bit >= 0; --bit)
          for (int d = 2; d >= 0; --d) k = (k << 1) | ((c[d] >> bit) & 1u);
        return k;
      };
      std::vector<std::pair<uint64_t, size_t>> ord(runs.size());
      for (size_t r = 0; r < runs.size(); ++r) ord[r] = {key(runs[r]), r};
      std::stable_sort(ord.begin(), ord.end());
      const size_t total = end - begin;
      size_t acc = 0;
      for (auto& kr : ord) {
        const auto& r = runs[kr.second];
        const int x = (int)std::min<size_t>(NXCD - 1, acc * NXCD / std::max<size_t>(total, 1));
        q[x].insert(q[x].end(), t.begin() + r.first, t.begin() + r.second);
        acc += r.second - r.first;
      }
    } else {
      for (auto& r : runs) {
        int best = 0;
        for (int x = 1; x < NXCD; ++x)
          if (q[x].size() < q[best].size()) best = x;
        q[best].insert(q[best].end(), t.begin() + r.first, t.begin() + r.second);
      }
    }
    std::vector<size_t> pos(NXCD, 0);
    size_t o = begin;
    while (o < end)
      for (int x = 0; x < NXCD && o < end; ++x)
        if (pos[x] < q[x].size()) t[o++] = q[x][pos[x]++];
  }
}
}  // namespace

extern "C" int dgr_plan_create(dgr_plan_t* plan_out, const dgr_block* blocks, uint32_t nblocks,
                               const dgr_bssn_params* p) {
  return dgr_plan_create_ex(plan_out, blocks, nblocks, p, nullptr);
}

extern "C" int dgr_plan_create_ex(dgr_plan_t* plan_out, const dgr_block* blocks, uint32_t nblocks,
                                  const dgr_bssn_params* p, const dgr_schedule* sched) {
  if (!plan_out || !blocks || !p || nblocks == 0) return DGR_E_INVALID;
  if (p->deriv_order != 42 && p->deriv_order != 64) return DGR_E_UNSUPPORTED;
  dgr_schedule sc;
  dgr_schedule_default(&sc);
  if (sched) {
    sc = *sched;
    for (int c = 0; c < DGR_SIZE_CLASSES; ++c)
      if (sc.tile_max[c] < 3 || sc.tile_max[c] > (uint32_t)TX) return DGR_E_INVALID;
    if (sc.order > DGR_ORDER_SFC || sc.grid_mult > 64) return DGR_E_INVALID;
  }
  std::vector<DevBlock> hb(nblocks);
  std::vector<DevTile> ht;
  dgr_plan_s* plan = new dgr_plan_s();
  plan->grid_mult = sc.grid_mult;
  plan->order = (int)p->deriv_order;
  plan->block_tile_begin.resize(nblocks);
  plan->block_tile_count.resize(nblocks);
  std::vector<std::pair<int, int>> sx, sy, sz;
  // interior blocks' tiles first, then boundary blocks' tiles: the two
  // ranges run through the interior- and boundary-specialised kernels
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t b = 0; b < nblocks; ++b) {
      const dgr_block& bl = blocks[b];
      if ((pass == 0) != (bl.bflag == 0)) continue;
      // the reference's bssnrhs takes any block with its 3-point padding;
      // one-sided closures need at least 6 interior points along an axis
      // with flagged faces (sz >= 12), as its own bssnrhs does
      const uint32_t szmin = bl.bflag ? 12u : 7u;
      for (int d = 0; d < 3; ++d) {
        if (bl.sz[d] < szmin || bl.sz[d] > 65535) { delete plan; return DGR_E_INVALID; }
      }
      DevBlock& db = hb[b];
      make_devblock(bl, db);
      plan->ngroups = std::max<uint32_t>(plan->ngroups, bl.group + 1);
      plan->block_tile_begin[b] = (uint32_t)ht.size();
      const int tmax = (int)sc.tile_max[size_class(bl)];
      split(bl.sz[0] - 6, tmax, sx);
      split(bl.sz[1] - 6, tmax, sy);
      split(bl.sz[2] - 6, tmax, sz);
      for (auto& cz : sz)
        for (auto& cy : sy)
          for (auto& cx : sx) {
            DevTile t;
            t.block = b;
            t.i0 = (uint16_t)cx.first; t.j0 = (uint16_t)cy.first; t.k0 = (uint16_t)cz.first;
            t.nx = (uint8_t)cx.second; t.ny = (uint8_t)cy.second; t.nz = (uint8_t)cz.second;
            t.pad = face_map_id(bl, t);
            ht.push_back(t);
          }
      plan->block_tile_count[b] = (uint32_t)ht.size() - plan->block_tile_begin[b];
      plan->points += (uint64_t)(bl.sz[0] - 6) * (bl.sz[1] - 6) * (bl.sz[2] - 6);
    }
    if (pass == 0) plan->ntiles_interior = (uint32_t)ht.size();
  }
  plan->nblocks = nblocks;
  plan->ntiles = (uint32_t)ht.size();
  // batched launches walk the tiles in the scheduled order; per-block
  // launches keep the block-order table
  std::vector<DevTile> hs = ht;
  order_tiles(hs, 0, plan->ntiles_interior, sc.order, blocks);
  order_tiles(hs, plan->ntiles_interior, hs.size(), sc.order, blocks);
  plan->h_tiles_blk = ht;
  // block-order index of every scheduled tile (compact k of the RK4 stages)
  std::vector<uint32_t> hslot(hs.size());
  {
    std::map<std::tuple<uint32_t, uint16_t, uint16_t, uint16_t>, uint32_t> pos;
    for (uint32_t q = 0; q < (uint32_t)ht.size(); ++q) pos[std::make_tuple(ht[q].block, ht[q].i0, ht[q].j0, ht[q].k0)] = q;
    for (size_t q = 0; q < hs.size(); ++q) hslot[q] = pos[std::make_tuple(hs[q].block, hs[q].i0, hs[q].j0, hs[q].k0)];
  }
  for (uint32_t b = 0; b < nblocks; ++b)
    plan->dof_end = std::max<uint64_t>(plan->dof_end, blocks[b].offset + (uint64_t)blocks[b].sz[0] * blocks[b].sz[1] *
                                                                           blocks[b].sz[2]);
  // compact region of the RK4 stage buffers: every block's interior points,
  // block after block
  for (uint32_t b = 0; b < nblocks; ++b) {
    if (plan->ncomp > 0xffffffffull) {
      delete plan;
      return DGR_E_INVALID;
    }
    hb[b].cbase = (uint32_t)plan->ncomp;
    plan->ncomp += (uint64_t)(blocks[b].sz[0] - 6) * (blocks[b].sz[1] - 6) * (blocks[b].sz[2] - 6);
  }
  plan->chi_floor = p->chi_floor;
  build_neighbours(blocks, hb, nblocks, plan->nbr);
  for (int i = 0; i < 4; ++i) plan->params.lambda[i] = p->lambda[i];
  plan->params.lambda_f[0] = p->lambda_f[0];
  plan->params.lambda_f[1] = p->lambda_f[1];
  plan->params.eta_const = p->eta_const;
  plan->params.eta_r0 = p->eta_r0;
  plan->params.eta_damping_exp = p->eta_damping_exp;
  plan->params.ko_sigma = p->ko_sigma;
  hipError_t e = hipGetDevice(&plan->device);
  if (e != hipSuccess) { delete plan; return hip_fail(e, "hipGetDevice"); }
  int ncu = 0;
  e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, plan->device);
  if (e != hipSuccess) { delete plan; return hip_fail(e, "hipDeviceGetAttribute"); }
  // the kernel holds ~140 KB of LDS, so one workgroup per CU is resident
  plan->grid = ncu;
  if (hipMalloc(&plan->d_blocks, sizeof(DevBlock) * nblocks) != hipSuccess ||
      hipMalloc(&plan->d_tiles, sizeof(DevTile) * ht.size()) != hipSuccess ||
      hipMalloc(&plan->d_tiles_blk, sizeof(DevTile) * ht.size()) != hipSuccess ||
      hipMalloc(&plan->d_tile_slot, sizeof(uint32_t) * std::max<size_t>(1, hslot.size())) != hipSuccess ||
      hipMalloc(&plan->d_io, sizeof(IOPtrs)) != hipSuccess ||
      hipMalloc(&plan->d_nbr, sizeof(int32_t) * plan->nbr.size()) != hipSuccess ||
      hipMalloc(&plan->d_delta, sizeof(long long) * plan->nbr.size()) != hipSuccess ||
      hipMalloc(&plan->d_cdelta, sizeof(long long) * plan->nbr.size()) != hipSuccess ||
      hipMalloc(&plan->d_flag, 256) != hipSuccess) {
    free_plan(plan);
    return DGR_E_NOMEM;
  }
  // the uploads run on the plan's own non-blocking stream, which is then
  // synchronised alone: a copy on the legacy default stream could wait for
  // unrelated work that shares its hardware queue
  e = hipStreamCreateWithFlags(&plan->aux, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&plan->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&plan->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) {
    free_plan(plan);
    return hip_fail(e, "hipStreamCreate(plan)");
  }
  const hipStream_t up = plan->aux;
  e = hipMemcpyAsync(plan->d_blocks, hb.data(), sizeof(DevBlock) * nblocks, hipMemcpyHostToDevice, up);
  if (e == hipSuccess) e = hipMemcpyAsync(plan->d_tiles, hs.data(), sizeof(DevTile) * hs.size(), hipMemcpyHostToDevice, up);
  if (e == hipSuccess)
    e = hipMemcpyAsync(plan->d_tiles_blk, ht.data(), sizeof(DevTile) * ht.size(), hipMemcpyHostToDevice, up);
  if (e == hipSuccess)
    e = hipMemcpyAsync(plan->d_nbr, plan->nbr.data(), sizeof(int32_t) * plan->nbr.size(), hipMemcpyHostToDevice, up);
  // gathering RHS: index distance from a ghost point to the same point of
  // the neighbour's interior, per block and ghost direction (0 where no
  // neighbour: the block's own padding); alive until the uploads are done
  std::vector<long long> hd(plan->nbr.size(), 0), hcd(plan->nbr.size(), 0);
  if (e == hipSuccess) {
    const long long dofe = (long long)plan->dof_end;
    for (uint32_t b = 0; b < nblocks; ++b) {
      const long long s0 = hb[b].sz[0], s1 = hb[b].sz[1], s2 = hb[b].sz[2];
      const long long n0 = s0 - 6, n1 = s1 - 6;
      // the block's own interior in a stage buffer's compact region (> 0:
      // dof_end exceeds the block's offset by its size)
      hcd[(size_t)b * NNBR + 13] = dofe + (long long)hb[b].cbase - (long long)hb[b].offset;
      for (int d = 0; d < NNBR; ++d) {
        const int n = plan->nbr[(size_t)b * NNBR + d];
        if (n < 0 || d == 13) continue;
        const long long dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
        hd[(size_t)b * NNBR + d] = (long long)hb[n].offset - (long long)hb[b].offset -
                                   (dx * (s0 - 7) + s0 * (dy * (s1 - 7) + s1 * dz * (s2 - 7)));
        // same point of the neighbour's interior, compact region; > 0, since
        // the shift is below the block's interior size
        hcd[(size_t)b * NNBR + d] = dofe + (long long)hb[n].cbase - (long long)hb[b].offset -
                                    (dx * (s0 - 7) + n0 * (dy * (s1 - 7) + n1 * dz * (s2 - 7)));
      }
    }
    e = hipMemcpyAsync(plan->d_delta, hd.data(), sizeof(long long) * hd.size(), hipMemcpyHostToDevice, up);
    if (e == hipSuccess)
      e = hipMemcpyAsync(plan->d_cdelta, hcd.data(), sizeof(long long) * hcd.size(), hipMemcpyHostToDevice, up);
  }
  if (e == hipSuccess && !hslot.empty())
    e = hipMemcpyAsync(plan->d_tile_slot, hslot.data(), sizeof(uint32_t) * hslot.size(), hipMemcpyHostToDevice, up);
  if (e == hipSuccess) e = hipStreamSynchronize(up);
  if (e != hipSuccess) {
    free_plan(plan);
    return hip_fail(e, "hipMemcpy(plan)");
  }
  const int rc = build_level_fill(plan, blocks, hb);
  if (rc != DGR_OK) {
    free_plan(plan);
    return rc;
  }
  plan->h_blocks = std::move(hb);
  *plan_out = plan;
  return DGR_OK;
}

extern "C" int dgr_block_neighbours(const dgr_block* blocks, uint32_t nblocks, int32_t* out) {
  if (!blocks || !out) return DGR_E_INVALID;
  std::vector<DevBlock> hb(nblocks);
  for (uint32_t b = 0; b < nblocks; ++b) {
    const dgr_block& bl = blocks[b];
    DevBlock& db = hb[b];
    for (int d = 0; d < 3; ++d) {
      if (bl.sz[d] < 2) return DGR_E_INVALID;
      db.sz[d] = bl.sz[d];
      db.pmin[d] = bl.pmin[d];
      db.h[d] = (bl.pmax[d] - bl.pmin[d]) / (bl.sz[d] - 1);
    }
    db.group = bl.group;
  }
  std::vector<int32_t> nbr;
  build_neighbours(blocks, hb, nblocks, nbr);
  std::memcpy(out, nbr.data(), sizeof(int32_t) * nbr.size());
  return DGR_OK;
}

extern "C" int dgr_plan_destroy(dgr_plan_t plan) {
  if (!plan) return DGR_E_INVALID;
  free_plan(plan);
  return DGR_OK;
}

extern "C" uint64_t dgr_plan_points(dgr_plan_t plan) { return plan ? plan->points : 0; }
extern "C" uint32_t dgr_plan_tiles(dgr_plan_t plan) { return plan ? plan->ntiles : 0; }

namespace {
// Upload the 48 array pointers to the plan's device table when they differ
// from the last call (stream-ordered, so calls on one stream see their own
// pointers).
int upload_ptrs(dgr_plan_t plan, double* const* rhs, const double* const* uin, hipStream_t st) {
  if (!rhs || !uin) return DGR_E_INVALID;
  IOPtrs io;
  for (int v = 0; v < 24; ++v) {
    if (!rhs[v] || !uin[v]) return DGR_E_INVALID;
    io.u.v[v] = uin[v];
    io.r.v[v] = rhs[v];
  }
  if (plan->io_valid && std::memcmp(&io, &plan->h_io, sizeof(IOPtrs)) == 0) return DGR_OK;
  plan->h_io = io;
  DGR_HIP(hipMemcpyAsync(plan->d_io, &plan->h_io, sizeof(IOPtrs), hipMemcpyHostToDevice, st));
  plan->io_valid = true;
  return DGR_OK;
}
}  // namespace

namespace {
using RhsKernel = decltype(&k_bssn_rhs<0, 42, false>);
// Kernel of one tile range: interior or boundary tiles, stencil order,
// gather mode.  Boundary tiles run the per-face-axis bodies.
RhsKernel rhs_kernel(bool bnd, int order, bool gather) {
  if (order == 64) {
    if (gather) return bnd ? k_bssn_rhs<FM_FACES, 64, true> : k_bssn_rhs<0, 64, true>;
    return bnd ? k_bssn_rhs<FM_FACES, 64, false> : k_bssn_rhs<0, 64, false>;
  }
  if (gather) return bnd ? k_bssn_rhs<FM_FACES, 42, true> : k_bssn_rhs<0, 42, true>;
  return bnd ? k_bssn_rhs<FM_FACES, 42, false> : k_bssn_rhs<0, 42, false>;
}
}  // namespace

int dgr_internal::launch_rhs(dgr_plan_t plan, const IOPtrs* d_io, hipStream_t st, bool gather, bool stage) {
  if (stage && !gather) return DGR_E_INVALID;
  const int ni = (int)plan->ntiles_interior;
  const int nbd = (int)plan->ntiles - ni;
  auto grid = [&](int n) { return plan->grid_mult == 0 ? n : std::min(plan->grid * (int)plan->grid_mult, n); };
  const auto kin = rhs_kernel(false, plan->order, gather);
  const auto kbd = rhs_kernel(true, plan->order, gather);
  const long long* nb = gather ? (stage ? plan->d_cdelta : plan->d_delta) : nullptr;
  const uint32_t* sl = gather ? plan->d_tile_slot : nullptr;
  // both ranges present: the boundary kernel (the longer tiles, launched
  // first) runs on the plan's second stream beside the interior kernel
  const bool fork = ni > 0 && nbd > 0;
  hipStream_t sb = st;
  if (fork) {
    DGR_HIP(hipEventRecord(plan->ev_fork, st));
    DGR_HIP(hipStreamWaitEvent(plan->aux, plan->ev_fork, 0));
    sb = plan->aux;
  }
  if (nbd > 0)
    hipLaunchKernelGGL(kbd, dim3(grid(nbd)), dim3(NT), 0, sb, plan->d_blocks, plan->d_tiles, ni, nbd, d_io,
                       plan->params, nb, sl, (int)stage);
  if (ni > 0)
    hipLaunchKernelGGL(kin, dim3(grid(ni)), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles, 0, ni, d_io,
                       plan->params, nb, sl, (int)stage);
  DGR_HIP(hipGetLastError());
  if (fork) {
    DGR_HIP(hipEventRecord(plan->ev_join, plan->aux));
    DGR_HIP(hipStreamWaitEvent(st, plan->ev_join, 0));
  }
  return DGR_OK;
}

extern "C" int dgr_bssn_rhs(dgr_plan_t plan, double* const* rhs, const double* const* uin, void* stream) {
  if (!plan) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  int rc = upload_ptrs(plan, rhs, uin, st);
  if (rc) return rc;
  return launch_rhs(plan, plan->d_io, st, false);
}

extern "C" int dgr_bssn_rhs_per_block(dgr_plan_t plan, double* const* rhs, const double* const* uin,
                                      void* stream) {
  if (!plan) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  int rc = upload_ptrs(plan, rhs, uin, st);
  if (rc) return rc;
  for (uint32_t b = 0; b < plan->nblocks; ++b) {
    const int tb = (int)plan->block_tile_begin[b];
    const int nt = (int)plan->block_tile_count[b];
    const auto k = rhs_kernel(tb >= (int)plan->ntiles_interior, plan->order, false);
    hipLaunchKernelGGL(k, dim3(nt), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles_blk, tb, nt, plan->d_io,
                       plan->params, (const long long*)nullptr, (const uint32_t*)nullptr, 0);
  }
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_bssn_rhs_blkwise(dgr_plan_t plan, double* const* rhs, const double* const* uin,
                                    const uint32_t* blk_ids, uint32_t n, void* stream) {
  if (!plan || (n > 0 && !blk_ids)) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  std::vector<DevTile> sub;
  uint32_t nin = 0;
  // interior blocks' tiles, then boundary blocks' tiles (two kernels)
  for (int pass = 0; pass < 2; ++pass)
    for (uint32_t q = 0; q < n; ++q) {
      const uint32_t b = blk_ids[q];
      if (b >= plan->nblocks) return DGR_E_INVALID;
      const uint32_t tb = plan->block_tile_begin[b], tc = plan->block_tile_count[b];
      if ((pass == 0) != (tb < plan->ntiles_interior)) continue;
      sub.insert(sub.end(), plan->h_tiles_blk.begin() + tb, plan->h_tiles_blk.begin() + tb + tc);
      if (pass == 0) nin += tc;
    }
  if (sub.empty()) return DGR_OK;
  int rc = upload_ptrs(plan, rhs, uin, st);
  if (rc) return rc;
  if (!plan->d_tiles_sub) DGR_HIP(hipMalloc(&plan->d_tiles_sub, sizeof(DevTile) * plan->ntiles));
  // the subset list is reused by the next call: wait for earlier launches
  DGR_HIP(hipStreamSynchronize(st));
  DGR_HIP(hipMemcpy(plan->d_tiles_sub, sub.data(), sizeof(DevTile) * sub.size(), hipMemcpyHostToDevice));
  const auto kin = rhs_kernel(false, plan->order, false);
  const auto kbd = rhs_kernel(true, plan->order, false);
  const int nbd = (int)sub.size() - (int)nin;
  auto grid = [&](int m) { return std::min(plan->grid * 2, m); };
  if (nin > 0)
    hipLaunchKernelGGL(kin, dim3(grid((int)nin)), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles_sub, 0, (int)nin,
                       plan->d_io, plan->params, (const long long*)nullptr, (const uint32_t*)nullptr, 0);
  if (nbd > 0)
    hipLaunchKernelGGL(kbd, dim3(grid(nbd)), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles_sub, (int)nin, nbd,
                       plan->d_io, plan->params, (const long long*)nullptr, (const uint32_t*)nullptr, 0);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_bssn_rhs_blk(dgr_plan_t plan, double* out, const double* in, uint32_t dof, uint32_t blk_id,
                                void* stream) {
  if (!plan || !out || !in || dof != 24 || blk_id >= plan->nblocks) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  const uint32_t tb = plan->block_tile_begin[blk_id], tc = plan->block_tile_count[blk_id];
  // block-local arrays: the kernels address u[v] + offset + local index, so
  // each pointer is shifted back by the block's offset
  uint64_t off = 0, nn = 0;
  {
    DevBlock B;
    DGR_HIP(hipMemcpy(&B, plan->d_blocks + blk_id, sizeof(DevBlock), hipMemcpyDeviceToHost));
    off = B.offset;
    nn = (uint64_t)B.sz[0] * B.sz[1] * B.sz[2];
  }
  IOPtrs io;
  for (int v = 0; v < 24; ++v) {
    io.u.v[v] = reinterpret_cast<const double*>(reinterpret_cast<uintptr_t>(in + v * nn) - off * sizeof(double));
    io.r.v[v] = reinterpret_cast<double*>(reinterpret_cast<uintptr_t>(out + v * nn) - off * sizeof(double));
  }
  if (!plan->d_io_blk) DGR_HIP(hipMalloc(&plan->d_io_blk, sizeof(IOPtrs)));
  DGR_HIP(hipStreamSynchronize(st));
  DGR_HIP(hipMemcpy(plan->d_io_blk, &io, sizeof(IOPtrs), hipMemcpyHostToDevice));
  const auto k = rhs_kernel(tb >= plan->ntiles_interior, plan->order, false);
  hipLaunchKernelGGL(k, dim3(tc), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles_blk, (int)tb, (int)tc,
                     plan->d_io_blk, plan->params, (const long long*)nullptr, (const uint32_t*)nullptr, 0);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_physcon(dgr_plan_t plan, double* const* con, const double* const* uin, void* stream) {
  if (!plan || !con || !uin) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  IOPtrs io{};
  for (int v = 0; v < 24; ++v) {
    if (!uin[v]) return DGR_E_INVALID;
    io.u.v[v] = uin[v];
  }
  for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q) {
    if (!con[q]) return DGR_E_INVALID;
    io.r.v[q] = con[q];
  }
  if (!plan->d_io_con) DGR_HIP(hipMalloc(&plan->d_io_con, sizeof(IOPtrs)));
  if (!plan->io_con_valid || std::memcmp(&io, &plan->h_io_con, sizeof(IOPtrs)) != 0) {
    plan->h_io_con = io;
    DGR_HIP(hipMemcpyAsync(plan->d_io_con, &plan->h_io_con, sizeof(IOPtrs), hipMemcpyHostToDevice, st));
    plan->io_con_valid = true;
  }
  const int ni = (int)plan->ntiles_interior;
  const int nbd = (int)plan->ntiles - ni;
  auto grid = [&](int n) { return plan->grid_mult == 0 ? n : std::min(plan->grid * (int)plan->grid_mult, n); };
  if (ni > 0)
    hipLaunchKernelGGL(k_physcon<false>, dim3(grid(ni)), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles, 0, ni,
                       plan->d_io_con, plan->params);
  if (nbd > 0)
    hipLaunchKernelGGL(k_physcon<true>, dim3(grid(nbd)), dim3(NT), 0, st, plan->d_blocks, plan->d_tiles, ni, nbd,
                       plan->d_io_con, plan->params);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

namespace {
// Owner of every unique node (the zipped vector's entries), from the block
// geometry the plan already holds.  A node on the interior boundary of a
// block can be shared:
//   - with same-level neighbours (face, edge or corner contact): of the
//     blocks around the node the one lying furthest in +z, then +y, then +x
//     owns it, so a block skips the node when a same-level neighbour that
//     also holds it lies in a direction whose last nonzero component (z
//     before y before x) is positive (the copies are equal);
//   - across a level jump: a fine block's node on a side whose neighbour is
//     coarser either lies between the coarse nodes (a hanging node, which has
//     no zipped entry: Mesh::zip copies only nodes whose owning element is
//     the block's own, include/mesh.tcc:3805-3812) or coincides with a coarse
//     node, which the coarse block owns; the fine block owns neither.
// Zipped order: blocks in order, each block's owned points x fastest.
int build_zip(dgr_plan_t plan) {
  if (plan->zip_built) return DGR_OK;
  std::vector<uint8_t> own(plan->dof_end, 0);
  std::vector<uint64_t> idx;
  idx.reserve(plan->points);
  for (uint32_t b = 0; b < plan->nblocks; ++b) {
    const DevBlock& B = plan->h_blocks[b];
    const int32_t* nb = &plan->nbr[(size_t)b * NNBR];
    const uint32_t cd = plan->coarse_dirs.empty() ? 0u : plan->coarse_dirs[b];
    const int s0 = (int)B.sz[0], s1 = (int)B.sz[1], s2 = (int)B.sz[2];
    for (int k = 3; k < s2 - 3; ++k)
      for (int j = 3; j < s1 - 3; ++j)
        for (int i = 3; i < s0 - 3; ++i) {
          // interior points hold no shared node
          if (i > 3 && i < s0 - 4 && j > 3 && j < s1 - 4 && k > 3 && k < s2 - 4) {
            const uint64_t o = B.offset + (uint64_t)i + (uint64_t)s0 * (j + (uint64_t)s1 * k);
            own[o] = 1;
            idx.push_back(o);
            continue;
          }
          bool skip = false;
          for (int d = 0; d < NNBR && !skip; ++d) {
            if (d == 13) continue;
            const int dx = d % 3 - 1, dy = (d / 3) % 3 - 1, dz = d / 9 - 1;
            const bool on = (dx == 0 || i == (dx < 0 ? 3 : s0 - 4)) && (dy == 0 || j == (dy < 0 ? 3 : s1 - 4)) &&
                            (dz == 0 || k == (dz < 0 ? 3 : s2 - 4));
            if (!on) continue;
            const bool up = dz > 0 || (dz == 0 && (dy > 0 || (dy == 0 && dx > 0)));
            skip = ((cd >> d) & 1u) || (up && nb[d] >= 0);
          }
          if (skip) continue;
          const uint64_t o = B.offset + (uint64_t)i + (uint64_t)s0 * (j + (uint64_t)s1 * k);
          own[o] = 1;
          idx.push_back(o);
        }
  }
  DGR_HIP(hipMalloc(&plan->d_own, std::max<size_t>(1, own.size())));
  DGR_HIP(hipMalloc(&plan->d_zidx, sizeof(uint64_t) * std::max<size_t>(1, idx.size())));
  DGR_HIP(hipMemcpy(plan->d_own, own.data(), own.size(), hipMemcpyHostToDevice));
  if (!idx.empty()) DGR_HIP(hipMemcpy(plan->d_zidx, idx.data(), sizeof(uint64_t) * idx.size(), hipMemcpyHostToDevice));
  plan->nzip = idx.size();
  plan->zip_built = true;
  return DGR_OK;
}
}  // namespace

extern "C" int dgr_zip_size(dgr_plan_t plan, uint64_t* n) {
  if (!plan || !n) return DGR_E_INVALID;
  const int rc = build_zip(plan);
  if (rc != DGR_OK) return rc;
  *n = plan->nzip;
  return DGR_OK;
}

extern "C" int dgr_zip(dgr_plan_t plan, double* zipped, const double* unzipped, void* stream) {
  if (!plan || !zipped || !unzipped) return DGR_E_INVALID;
  const int rc = build_zip(plan);
  if (rc != DGR_OK) return rc;
  if (plan->nzip == 0) return DGR_OK;
  const uint64_t nwg = std::min<uint64_t>((plan->nzip + 255) / 256, 65536);
  hipLaunchKernelGGL(k_zip, dim3((unsigned)nwg), dim3(256), 0, (hipStream_t)stream, plan->d_zidx, plan->nzip,
                     unzipped, zipped);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_physcon_norms(dgr_plan_t plan, const double* const* con, const double* mask, double threshold,
                                 double* out, void* stream) {
  if (!plan || !con || !mask || !out) return DGR_E_INVALID;
  {
    const int rc = build_zip(plan);
    if (rc != DGR_OK) return rc;
  }
  hipStream_t st = (hipStream_t)stream;
  Ptr6 c6;
  for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q) {
    if (!con[q]) return DGR_E_INVALID;
    c6.v[q] = con[q];
  }
  const int nt = (int)plan->ntiles;
  if (nt == 0) {
    for (int q = 0; q < 13; ++q) out[q] = 0.0;
    return DGR_OK;
  }
  if (!plan->d_norm) DGR_HIP(hipMalloc(&plan->d_norm, sizeof(double) * 13 * ((size_t)nt + 1)));
  double* res = plan->d_norm + (size_t)13 * nt;
  hipLaunchKernelGGL(k_con_norm_partial, dim3(std::min(nt, 8192)), dim3(NORM_T), 0, st, plan->d_blocks,
                     plan->d_tiles_blk, nt, plan->d_own, c6, mask, threshold, plan->d_norm);
  hipLaunchKernelGGL(k_con_norm_final, dim3(13), dim3(NORM_T), 0, st, plan->d_norm, nt, res);
  DGR_HIP(hipGetLastError());
  double h[13];
  DGR_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
  DGR_HIP(hipStreamSynchronize(st));
  const double n = h[12];
  for (int q = 0; q < 6; ++q) {
    out[q] = n > 0 ? std::sqrt(h[q] / n) : 0.0;
    out[6 + q] = h[6 + q];
  }
  out[12] = n;
  return DGR_OK;
}

extern "C" int dgr_bssn_rhs_blocks(double* const* rhs, const double* const* u, const dgr_block* blocks,
                                   uint32_t nblocks, const dgr_bssn_params* p, void* stream) {
  dgr_plan_t plan = nullptr;
  int rc = dgr_plan_create(&plan, blocks, nblocks, p);
  if (rc) return rc;
  rc = dgr_bssn_rhs(plan, rhs, u, stream);
  if (rc == DGR_OK) {
    hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamSynchronize");
  }
  dgr_plan_destroy(plan);
  return rc;
}

// ---------------------------------------------------------------------------
// One-shot bssnRHS without device allocation (no counterpart in the
// reference).  The block table, the tile table and the 48 array pointers go
// to `scratch`, a file-scope device array; each GPU holds its own copy of it
// until the process ends.  Host code never passes the array's name to a
// kernel or a copy: it uses the device address hipGetSymbolAddress returns,
// looked up once per device with that device current.
// ---------------------------------------------------------------------------
namespace {
constexpr size_t NA_OFF_TILES = sizeof(DevBlock) * DGR_NOALLOC_MAX_BLOCKS;
constexpr size_t NA_OFF_IO = NA_OFF_TILES + sizeof(DevTile) * DGR_NOALLOC_MAX_TILES;
constexpr size_t NA_BYTES = (NA_OFF_IO + sizeof(IOPtrs) + 255) / 256 * 256;
static_assert(NA_OFF_TILES % 256 == 0 && NA_OFF_IO % 256 == 0, "scratch sections 256-byte aligned");
static_assert(NA_BYTES == DGR_NOALLOC_SCRATCH_BYTES, "DGR_NOALLOC_SCRATCH_BYTES matches the scratch layout");

__device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
struct NoAlloc {
  std::mutex m;                       // host tables and the per-device addresses
  unsigned char* ptr[NA_MAX_DEVICES] = {};  // device address of `scratch` per device
  int ncu[NA_MAX_DEVICES] = {};
  DevBlock hb[DGR_NOALLOC_MAX_BLOCKS];
  DevTile ht[DGR_NOALLOC_MAX_TILES];
};
NoAlloc g_na;  // static storage: the entry point allocates nothing on the host either

// Device address of `scratch` on the current device (looked up on the first
// call on that device).  Caller holds g_na.m.
int noalloc_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= NA_MAX_DEVICES) return DGR_E_UNSUPPORTED;
  if (!g_na.ptr[dev]) {
    void* p = nullptr;
    DGR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    int ncu = 0;
    DGR_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    g_na.ncu[dev] = ncu;
    g_na.ptr[dev] = static_cast<unsigned char*>(p);
  }
  *out = g_na.ptr[dev];
  return DGR_OK;
}

int noalloc_init(int dev) {
  std::lock_guard<std::mutex> lock(g_na.m);
  unsigned char* p = nullptr;
  return noalloc_ptr(dev, &p);
}
}  // namespace

extern "C" int dgr_bssn_rhs_blocks_noalloc(double* const* rhs, const double* const* u, const dgr_block* blocks,
                                           uint32_t nblocks, const dgr_bssn_params* p, void* stream) {
  if (!rhs || !u || !blocks || !p || nblocks == 0) return DGR_E_INVALID;
  if (nblocks > DGR_NOALLOC_MAX_BLOCKS) return DGR_E_INVALID;
  if (p->deriv_order != 42 && p->deriv_order != 64) return DGR_E_UNSUPPORTED;
  IOPtrs io;
  for (int v = 0; v < 24; ++v) {
    if (!rhs[v] || !u[v]) return DGR_E_INVALID;
    io.u.v[v] = u[v];
    io.r.v[v] = rhs[v];
  }
  // tiles of at most 7 points per axis (the default schedule, block order)
  auto parts = [](uint32_t n) { return (n + TX - 1) / TX; };
  uint64_t ntiles = 0;
  for (uint32_t b = 0; b < nblocks; ++b) {
    const dgr_block& bl = blocks[b];
    const uint32_t szmin = bl.bflag ? 12u : 7u;
    for (int d = 0; d < 3; ++d)
      if (bl.sz[d] < szmin || bl.sz[d] > 65535) return DGR_E_INVALID;
    ntiles += (uint64_t)parts(bl.sz[0] - 6) * parts(bl.sz[1] - 6) * parts(bl.sz[2] - 6);
  }
  if (ntiles > DGR_NOALLOC_MAX_TILES) return DGR_E_INVALID;
  const hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lock(g_na.m);
  int dev = 0;
  DGR_HIP(hipGetDevice(&dev));
  unsigned char* ptr = nullptr;
  const int rc = noalloc_ptr(dev, &ptr);
  if (rc != DGR_OK) return rc;
  // interior blocks' tiles first, then the boundary blocks' (the two
  // kernels), each block's tiles as dgr_plan_create cuts them
  uint32_t nt = 0, ni = 0;
  for (int pass = 0; pass < 2; ++pass) {
    for (uint32_t b = 0; b < nblocks; ++b) {
      const dgr_block& bl = blocks[b];
      if ((pass == 0) != (bl.bflag == 0)) continue;
      make_devblock(bl, g_na.hb[b]);
      uint32_t np[3], n[3];
      for (int d = 0; d < 3; ++d) {
        n[d] = bl.sz[d] - 6;
        np[d] = parts(n[d]);
      }
      uint32_t c0[3] = {3, 3, 3};
      for (uint32_t pz = 0; pz < np[2]; ++pz) {
        const uint32_t lz = n[2] / np[2] + (pz < n[2] % np[2] ? 1 : 0);
        c0[1] = 3;
        for (uint32_t py = 0; py < np[1]; ++py) {
          const uint32_t ly = n[1] / np[1] + (py < n[1] % np[1] ? 1 : 0);
          c0[0] = 3;
          for (uint32_t px = 0; px < np[0]; ++px) {
            const uint32_t lx = n[0] / np[0] + (px < n[0] % np[0] ? 1 : 0);
            DevTile& t = g_na.ht[nt++];
            t.block = b;
            t.i0 = (uint16_t)c0[0]; t.j0 = (uint16_t)c0[1]; t.k0 = (uint16_t)c0[2];
            t.nx = (uint8_t)lx; t.ny = (uint8_t)ly; t.nz = (uint8_t)lz;
            t.pad = face_map_id(bl, t);
            c0[0] += lx;
          }
          c0[1] += ly;
        }
        c0[2] += lz;
      }
    }
    if (pass == 0) ni = nt;
  }
  DevBlock* d_blocks = reinterpret_cast<DevBlock*>(ptr);
  DevTile* d_tiles = reinterpret_cast<DevTile*>(ptr + NA_OFF_TILES);
  IOPtrs* d_io = reinterpret_cast<IOPtrs*>(ptr + NA_OFF_IO);
  DGR_HIP(hipMemcpyAsync(d_blocks, g_na.hb, sizeof(DevBlock) * nblocks, hipMemcpyHostToDevice, st));
  DGR_HIP(hipMemcpyAsync(d_tiles, g_na.ht, sizeof(DevTile) * nt, hipMemcpyHostToDevice, st));
  DGR_HIP(hipMemcpyAsync(d_io, &io, sizeof(IOPtrs), hipMemcpyHostToDevice, st));
  DevParams P;
  for (int i = 0; i < 4; ++i) P.lambda[i] = p->lambda[i];
  P.lambda_f[0] = p->lambda_f[0];
  P.lambda_f[1] = p->lambda_f[1];
  P.eta_const = p->eta_const;
  P.eta_r0 = p->eta_r0;
  P.eta_damping_exp = p->eta_damping_exp;
  P.ko_sigma = p->ko_sigma;
  // persistent grid of one workgroup per CU, both tile ranges on the
  // caller's stream one after the other
  const int order = (int)p->deriv_order;
  const int nbd = (int)(nt - ni);
  auto grid = [&](int n) { return std::min(std::max(g_na.ncu[dev], 1), n); };
  if (ni > 0)
    hipLaunchKernelGGL(rhs_kernel(false, order, false), dim3(grid((int)ni)), dim3(NT), 0, st, d_blocks, d_tiles, 0,
                       (int)ni, d_io, P, nullptr, nullptr, 0);
  if (nbd > 0)
    hipLaunchKernelGGL(rhs_kernel(true, order, false), dim3(grid(nbd)), dim3(NT), 0, st, d_blocks, d_tiles,
                       (int)ni, nbd, d_io, P, nullptr, nullptr, 0);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

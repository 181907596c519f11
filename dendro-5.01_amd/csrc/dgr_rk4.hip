/*
 * RK4 time step on device-resident padded blocks.
 *
 * Reference behaviour (BSSN_GR/src/rkBSSN.cpp:803-1011, coefficients
 * BSSN_GR/include/rkBSSN.h:55-58): per stage s = 0..2
 *     k_s  = RHS(u_s)                      (u_0 = u_prev)
 *     u_{s+1} = u_prev + U[s+1] dt k_s,   then enforce_bssn_constraints
 *     ghost exchange + unzip of u_{s+1}
 * and finally u = u_prev + sum_s C[s] dt k_s (s = 0..3 in order), then
 * enforce.  The reference does the combine on zipped nodal vectors; here
 * the state stays unzipped: the combine runs on every block's interior
 * points (shared face nodes are duplicated in both blocks and receive
 * identical values), and the "unzip" is a same-level ghost fill: the RHS
 * of each stage reads a ghost point that has a same-level neighbour from
 * that neighbour's interior (gather mode of the RHS kernel), which is what a
 * refill before it would have copied there.
 *
 * During one evolution call the state lives in a state buffer of the work
 * memory laid out like the stage buffer (padded arrays for the ghosts, then
 * the interior points block after block): packed from the caller's arrays
 * when the call starts, unpacked into them when it ends.  Every read and
 * write of the state inside a step then covers whole 64-byte lines instead
 * of the 7-point row fragments of the padded layout.
 *
 * Kernels:
 *   k_stage          fused stage combine + running final sum + constraint
 *                    projection (BSSN_GR/src/bssn_constraints.cpp:12-172)
 *   k_hybrid         caller's state <-> the state buffer (pack at the start
 *                    of an evolution call; at its end unpack with the
 *                    same-level ghost refill fused in)
 *   k_enforce        the projection alone
 *   k_ghost_refill   padding <- neighbour interiors (same-level ghosts;
 *                    dgr_ghost_refill)
 * One step (four RHS + stage launches) is captured once into a hipGraph and
 * replayed.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/dgr_bssn.h"
#include "dgr_internal.h"

using namespace dgr_internal;

namespace {

constexpr int RK_NT = 256;
enum { V_ALPHA = 0, V_CHI = 1, V_SG0 = 12, V_SA0 = 18 };

// The 24 variable arrays of one state.  Kernels take these tables by value:
// pointers in the kernel-argument segment are known to be global and cannot
// be clobbered by the kernel's own stores, so they stay in scalar registers
// instead of being re-read after every store.
struct Ptrs24 {
  double* v[24];
};

// RK4 work memory: one allocation holding the state buffer (h0) and the
// stage state (stg), each 24 hybrid arrays (dof padded points, then the
// compact interior region), then in a compact tile-major layout
// (343 slots per tile in block order, the layout the RHS kernels write in
// their gather mode) k_1, the running final sum, k_0 (later k_3) and k_2;
// all addressed arithmetically so that only the caller's state needs a
// pointer table.  With k_0 and k_2 kept, stages 0 and 2 store no running
// sum: stage 1 forms it from k_0 and the last stage adds k_2 before k_3,
// the same operations in the same order.
struct Work {
  double* base;   // stage state: per variable dof padded points, then the compact interior region
  uint64_t dof;
  uint64_t sstride;  // doubles per variable of the stage state (dof + compact region)
  double* kc;     // k of this stage, compact: 24 x nacc
  double* acc;    // running sum, compact: 24 x nacc
  uint64_t nacc;
  double* kx;     // stage 1: k_0; last stage: k_2 (compact)
};

// Per-thread cursor over the 24 variables of one point in the work memory.
// The cursor is advanced one variable at a time and kept opaque to the
// optimiser, so the 72 per-variable base addresses are not hoisted into
// scalar registers (which would spill).
typedef __attribute__((address_space(1))) double gdouble;  // global memory
struct WorkCursor {
  double* q;        // stage state of variable v at padded point o
  __device__ __forceinline__ WorkCursor(const Work& W, uint64_t o) : q(W.base + o) {}
  __device__ __forceinline__ gdouble& stg() { return ((gdouble*)q)[0]; }
  __device__ __forceinline__ void next(uint64_t n) {
    q += n;
    asm volatile("" : "+v"(q));
  }
};
// The running sum is never seen outside k_stage, so it lives in a compact
// tile-major layout (slot w * 343 + thread of tile w): whole-line reads and
// writes instead of the 7-point row fragments of the padded layout.
struct AccCursor {
  double* q;
  __device__ __forceinline__ AccCursor(double* base, uint64_t co) : q(base + co) {}
  __device__ __forceinline__ gdouble& acc() { return ((gdouble*)q)[0]; }
  __device__ __forceinline__ void next(uint64_t n) {
    q += n;
    asm volatile("" : "+v"(q));
  }
};

enum StageKind { ST_FIRST = 0, ST_SECOND = 1, ST_THIRD = 2, ST_LAST = 3 };

// enforce_bssn_constraints on one point (bssn_constraints.cpp:12-172):
// rescale gt to unit determinant, remove the trace of At, floor chi and
// alpha.  Returns a nonzero flag where the reference would exit.
__device__ __forceinline__ int enforce_point(double (&u)[24], double chi_floor) {
#pragma clang fp contract(off)
  double g00 = u[V_SG0], g01 = u[V_SG0 + 1], g02 = u[V_SG0 + 2], g11 = u[V_SG0 + 3], g12 = u[V_SG0 + 4],
         g22 = u[V_SG0 + 5];
  double det = g00 * (g11 * g22 - g12 * g12) - g01 * g01 * g22 + 2.0 * g01 * g02 * g12 - g02 * g02 * g11;
  if (det < 0.0) return 1;
  const double one_third = 1.0 / 3.0;
  const double s = 1.0 / pow(det, one_third);
  g00 *= s; g01 *= s; g02 *= s; g11 *= s; g12 *= s; g22 *= s;
  det = g00 * (g11 * g22 - g12 * g12) - g01 * g01 * g22 + 2.0 * g01 * g02 * g12 - g02 * g02 * g11;
  int bad = fabs(det - 1.0) > 1.0e-6 ? 2 : 0;
  const double idet = 1.0 / det;
  const double gu00 = idet * (g11 * g22 - g12 * g12);
  const double gu01 = idet * (-g01 * g22 + g02 * g12);
  const double gu02 = idet * (g01 * g12 - g02 * g11);
  const double gu11 = idet * (g00 * g22 - g02 * g02);
  const double gu12 = idet * (-g00 * g12 + g01 * g02);
  const double gu22 = idet * (g00 * g11 - g01 * g01);
  double a00 = u[V_SA0], a01 = u[V_SA0 + 1], a02 = u[V_SA0 + 2], a11 = u[V_SA0 + 3], a12 = u[V_SA0 + 4],
         a22 = u[V_SA0 + 5];
  const double t3 = one_third * (a00 * gu00 + a11 * gu11 + a22 * gu22 + 2.0 * (a01 * gu01 + a02 * gu02 + a12 * gu12));
  a00 -= t3 * g00; a01 -= t3 * g01; a02 -= t3 * g02; a11 -= t3 * g11; a12 -= t3 * g12; a22 -= t3 * g22;
  const double trA = a00 * gu00 + a11 * gu11 + a22 * gu22 + 2.0 * (a01 * gu01 + a02 * gu02 + a12 * gu12);
  if (fabs(trA) > 1.0e-6) bad |= 4;
  u[V_SA0] = a00; u[V_SA0 + 1] = a01; u[V_SA0 + 2] = a02; u[V_SA0 + 3] = a11; u[V_SA0 + 4] = a12; u[V_SA0 + 5] = a22;
  u[V_SG0] = g00; u[V_SG0 + 1] = g01; u[V_SG0 + 2] = g02; u[V_SG0 + 3] = g11; u[V_SG0 + 4] = g12; u[V_SG0 + 5] = g22;
  if (u[V_CHI] < chi_floor) u[V_CHI] = chi_floor;
  u[V_ALPHA] = fmax(u[V_ALPHA], chi_floor);
  return bad;
}

// Interior point of a tile handled by thread t (plain row-major order):
// its padded index o and, if asked, its slot in the compact interior region
// of a stage buffer (block after block, x fastest).
__device__ __forceinline__ bool tile_point(const DevTile& tl, const DevBlock& B, int t, uint64_t& o,
                                           uint64_t* oc = nullptr) {
  const int ti = t % 7, tj = (t / 7) % 7, tk = t / 49;
  if (t >= 343 || ti >= tl.nx || tj >= tl.ny || tk >= tl.nz) return false;
  const uint64_t i = tl.i0 + ti, j = tl.j0 + tj, k = tl.k0 + tk;
  o = B.offset + i + (uint64_t)B.sz[0] * (j + (uint64_t)B.sz[1] * k);
  if (oc) *oc = B.cbase + (i - 3) + (uint64_t)(B.sz[0] - 6) * ((j - 3) + (uint64_t)(B.sz[1] - 6) * (k - 3));
  return true;
}

// RK4_U[s+1] and RK4_C[s] of BSSN_GR/include/rkBSSN.h:55-58
__constant__ double c_rkU[3] = {1.0 / 2.0, 1.0 / 2.0, 1.0};
__constant__ double c_rkC[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

// Stage s of the RK4 step, one workgroup per tile.  `state` is the
// evolver's state buffer: during a run the state lives in its compact
// interior region (whole lines), unpacked into the caller's arrays when the
// run ends.  gdt[g] is the time step
// of block group g; the coefficients are formed as the reference forms them
// (RK4_U[s+1] * dt, RK4_C[s] * dt) and then multiply k.  u_prev is always
// the state at the start of the step; KIND fixes whether the running sum
// starts from it (stage 0) and where the new values go (the stage buffer,
// or the state after the last stage).  Every operand of a point is loaded
// before the first store.  The stores are non-temporal: no kernel reads
// them back before the whole stage is written, and they would only evict
// the inputs still to be read.  The stage state goes to the compact region
// of the stage buffer (whole lines; the padded layout's 7-point row
// fragments cost a third of the stage's time), which the next RHS and the
// level-jump fills read.
template <int KIND>
__global__ void __launch_bounds__(384) k_stage(const DevBlock* __restrict__ blocks, const DevTile* __restrict__ tiles,
                                              int ntiles, Ptrs24 state, Work W, const double* __restrict__ gdt,
                                              double chi_floor, int stage, unsigned* __restrict__ flag) {
  {
    const int w = blockIdx.x;
    if (w >= ntiles) return;
    const DevTile tl = tiles[w];
    const DevBlock& B = blocks[tl.block];
    uint64_t o, oc;
    if (!tile_point(tl, B, threadIdx.x, o, &oc)) return;
    const double dt = gdt[B.group];
    const uint64_t co = (uint64_t)w * 343 + threadIdx.x;
    double un[24], kk[24];
    if (KIND == ST_FIRST || KIND == ST_THIRD) {
      // no running sum here: stage 1 forms u_prev + (RK4_C[0] dt) k_0, the
      // last stage adds (RK4_C[2] dt) k_2
      const double cu = c_rkU[stage] * dt;
      AccCursor c(W.kc, co);
#pragma unroll
      for (int v = 0; v < 24; ++v) {
        un[v] = state.v[v][W.dof + oc];  // u_prev: the state at the start of the step
        kk[v] = c.acc();
        c.next(W.nacc);
      }
#pragma unroll
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        un[v] = un[v] + cu * kk[v];
      }
    } else if (KIND == ST_SECOND) {
      const double cu = c_rkU[stage] * dt, cc = c_rkC[stage] * dt;
      double acc[24];
      AccCursor c(W.kc, co);
      AccCursor a(W.kx, co);
#pragma unroll
      for (int v = 0; v < 24; ++v) {
        un[v] = state.v[v][W.dof + oc];  // u_prev: the state at the start of the step
        kk[v] = c.acc();
        acc[v] = a.acc();       // k_0
        c.next(W.nacc);
        a.next(W.nacc);
      }
      const double c0 = c_rkC[0] * dt;
#pragma unroll
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        acc[v] = un[v] + c0 * acc[v];
      }
#pragma unroll
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        acc[v] = acc[v] + cc * kk[v];
        un[v] = un[v] + cu * kk[v];
      }
      AccCursor d(W.acc, co);
#pragma unroll
      for (int v = 0; v < 24; ++v) {
        __builtin_nontemporal_store(acc[v], &d.acc());
        d.next(W.nacc);
      }
    } else {
      const double c2 = c_rkC[2] * dt, cc = c_rkC[3] * dt;
      double k2[24];
      AccCursor c(W.kc, co);
      AccCursor a(W.acc, co);
      AccCursor x(W.kx, co);
#pragma unroll
      for (int v = 0; v < 24; ++v) {
        un[v] = a.acc();
        kk[v] = c.acc();
        k2[v] = x.acc();
        c.next(W.nacc);
        a.next(W.nacc);
        x.next(W.nacc);
      }
#pragma unroll
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        un[v] = un[v] + c2 * k2[v];
        un[v] = un[v] + cc * kk[v];
      }
    }
    const int bad = enforce_point(un, chi_floor);
    if (bad) atomicOr(flag, (unsigned)bad);
    if (KIND == ST_LAST) {
#pragma unroll
      for (int v = 0; v < 24; ++v) __builtin_nontemporal_store(un[v], (gdouble*)&state.v[v][W.dof + oc]);
    } else {
      WorkCursor d(W, W.dof + oc);
#pragma unroll
      for (int v = 0; v < 24; ++v) {
        __builtin_nontemporal_store(un[v], &d.stg());
        d.next(W.sstride);
      }
    }
  }
}

// Between the caller's padded arrays u and the state buffer h, one block
// per workgroup iteration.  cdelta is the plan's stage-buffer direction
// table: per block and ghost direction the index distance to the same point
// of a neighbour's (or for direction 13 the block's own) interior in the
// compact region, 0 where no same-level neighbour exists.
// PACK: interior points to the compact region, and the ghosts without a
// same-level neighbour that the stages read from the padding (level jumps,
// concave corners) to the padded part.  Ghosts beyond a flagged face are
// staged as zeros by the RHS whatever the padding holds, and are skipped.
// UNPACK: the interior points, and every ghost that has a same-level
// neighbour from that neighbour's interior (the same-level refill); the
// other ghosts are left to the level fills or keep their values.
// A block none of whose ghosts needs this pass (PACK: every ghost direction
// has a same-level neighbour or lies beyond a flagged face; UNPACK: no
// direction has one) is walked over its interior points only.
// XCD-aware: workgroup w runs on XCD w % 8, so the grid (a multiple of 8)
// is renumbered to give each XCD a contiguous run of blocks, whose
// neighbours' interiors the unpack reads from the same L2.
template <bool PACK>
__global__ void __launch_bounds__(RK_NT) k_hybrid(const DevBlock* __restrict__ blocks,
                                                  const long long* __restrict__ cdelta, int nblocks, Ptrs24 u,
                                                  Ptrs24 h) {
  __shared__ int s_ghosts;
  const int first = (int)(blockIdx.x % 8) * (int)(gridDim.x / 8) + (int)(blockIdx.x / 8);
  for (int b = first; b < nblocks; b += gridDim.x) {
    const DevBlock& B = blocks[b];
    const int s0 = B.sz[0], s1 = B.sz[1], s2 = B.sz[2];
    __syncthreads();  // the previous block's s_ghosts has been read
    if (threadIdx.x == 0) s_ghosts = 0;
    __syncthreads();
    if (threadIdx.x < NNBR && threadIdx.x != 13) {
      const int dir = (int)threadIdx.x, dx = dir % 3, dy = (dir / 3) % 3, dz = dir / 9;
      const unsigned f = B.bflag;
      const bool beyond = (dx == 0 && (f & 1u)) || (dx == 2 && (f & 2u)) || (dy == 0 && (f & 4u)) ||
                          (dy == 2 && (f & 8u)) || (dz == 0 && (f & 16u)) || (dz == 2 && (f & 32u));
      const long long d = cdelta[(size_t)b * NNBR + dir];
      if (PACK ? (d == 0 && !beyond) : (d != 0)) s_ghosts = 1;
    }
    __syncthreads();
    const bool all = s_ghosts != 0;
    const int m0 = s0 - 6, m1 = s1 - 6;
    const int n = all ? s0 * s1 * s2 : m0 * m1 * (s2 - 6);
    for (int q = threadIdx.x; q < n; q += RK_NT) {
      const int i = all ? q % s0 : 3 + q % m0;
      const int j = all ? (q / s0) % s1 : 3 + (q / m0) % m1;
      const int k = all ? q / (s0 * s1) : 3 + q / (m0 * m1);
      const int e = i + s0 * (j + s1 * k);
      const int dx = i < 3 ? 0 : (i >= s0 - 3 ? 2 : 1);
      const int dy = j < 3 ? 0 : (j >= s1 - 3 ? 2 : 1);
      const int dz = k < 3 ? 0 : (k >= s2 - 3 ? 2 : 1);
      const int dir = dx + 3 * dy + 9 * dz;
      const long long d = cdelta[(size_t)b * NNBR + dir];
      const uint64_t o = B.offset + e;
      // the point's slot in the compact region (through a neighbour's
      // interior for a ghost), valid where d != 0
      const uint64_t c = (uint64_t)((long long)B.offset + (i - 3) +
                                    (long long)(s0 - 6) * ((j - 3) + (long long)(s1 - 6) * (k - 3)) + d);
      double t[24];
      if (PACK) {
        const unsigned f = B.bflag;
        const bool beyond = ((f & 1u) && i < 3) || ((f & 2u) && i >= s0 - 3) || ((f & 4u) && j < 3) ||
                            ((f & 8u) && j >= s1 - 3) || ((f & 16u) && k < 3) || ((f & 32u) && k >= s2 - 3);
        if (dir != 13 && (d != 0 || beyond)) continue;
        const uint64_t to = dir == 13 ? c : o;
#pragma unroll
        for (int v = 0; v < 24; ++v) t[v] = u.v[v][o];
#pragma unroll
        for (int v = 0; v < 24; ++v) h.v[v][to] = t[v];
      } else {
        if (d == 0) continue;
#pragma unroll
        for (int v = 0; v < 24; ++v) t[v] = h.v[v][c];
#pragma unroll
        for (int v = 0; v < 24; ++v) u.v[v][o] = t[v];
      }
    }
  }
}

__global__ void __launch_bounds__(384) k_enforce(const DevBlock* __restrict__ blocks, const DevTile* __restrict__ tiles,
                                                int ntiles, Ptrs24 tab, double chi_floor,
                                                unsigned* __restrict__ flag) {
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const DevTile tl = tiles[w];
    const DevBlock& B = blocks[tl.block];
    uint64_t o;
    if (!tile_point(tl, B, threadIdx.x, o)) continue;
    double u[24];
#pragma unroll
    for (int v = 0; v < 24; ++v) u[v] = tab.v[v][o];
    const int bad = enforce_point(u, chi_floor);
    if (bad) atomicOr(flag, (unsigned)bad);
    if (bad & 1) continue;  // the reference exits before touching this node
#pragma unroll
    for (int v = 0; v < 24; ++v) tab.v[v][o] = u[v];
  }
}

// Ghost refill: every padded point outside [3, sz-3)^3 takes the value of
// the same physical point in the interior of the same-level neighbour.
__global__ void __launch_bounds__(RK_NT) k_ghost_refill(const DevBlock* __restrict__ blocks,
                                                       const int32_t* __restrict__ nbr, int nblocks,
                                                       Ptrs24 tab) {
  for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const DevBlock& B = blocks[b];
    const int s0 = B.sz[0], s1 = B.sz[1], s2 = B.sz[2];
    const int n = s0 * s1 * s2;
    for (int e = threadIdx.x; e < n; e += RK_NT) {
      const int i = e % s0, j = (e / s0) % s1, k = e / (s0 * s1);
      const int di = i < 3 ? -1 : (i >= s0 - 3 ? 1 : 0);
      const int dj = j < 3 ? -1 : (j >= s1 - 3 ? 1 : 0);
      const int dk = k < 3 ? -1 : (k >= s2 - 3 ? 1 : 0);
      if ((di | dj | dk) == 0) continue;
      const int nb = nbr[(size_t)b * NNBR + (di + 1) + 3 * (dj + 1) + 9 * (dk + 1)];
      if (nb < 0) continue;
      const DevBlock& N = blocks[nb];
      // neighbour corners are sz-7 points apart along each axis
      const int si = i - di * (s0 - 7), sj = j - dj * (s1 - 7), sk = k - dk * (s2 - 7);
      const uint64_t src = N.offset + (uint64_t)si + (uint64_t)s0 * (sj + (uint64_t)s1 * sk);
      const uint64_t dst = B.offset + (uint64_t)e;
      double t[24];  // all loads before the stores (see k_stage)
#pragma unroll
      for (int v = 0; v < 24; ++v) t[v] = tab.v[v][src];
#pragma unroll
      for (int v = 0; v < 24; ++v) tab.v[v][dst] = t[v];
    }
  }
}

// One workgroup per tile or block, capped well above what fills the chip.
int grid_for(dgr_plan_t plan, int n) {
  const int g = plan->grid * 64;
  return n < g ? (n > 0 ? n : 1) : g;
}
// The same, rounded up to a multiple of the 8 XCDs (k_hybrid).
int grid_xcd(dgr_plan_t plan, int n) { return (grid_for(plan, n) + 7) / 8 * 8; }

}  // namespace

struct dgr_rk4_s {
  dgr_plan_t plan = nullptr;
  uint64_t dof = 0;
  double* buf = nullptr;        // state buffer, stage state (24 x sstride each), k, running sum, k_0 (24 x 343 per tile each)
  uint64_t sstride = 0;         // stage state per variable: dof padded points, then the compact interior region
  Ptrs24 state{}, h0{}, stg{}, k{};  // the caller's state; the state buffer, stage buffer and k arrays in buf
  uint64_t nacc = 0;           // compact running-sum slots per variable (343 per tile)
  IOPtrs* d_io = nullptr;       // RHS pointer table of each stage (input -> its k array)
  double* d_gdt = nullptr;      // time step per block group
  std::vector<double> h_gdt;
  bool gdt_set = false;
  unsigned* d_flag = nullptr;
  bool own_buf = true;
  hipStream_t st = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool use_graph = true;
};

namespace {

int enqueue_step(dgr_rk4_s* r, hipStream_t st) {
  dgr_plan_t p = r->plan;
  const int nt = (int)p->ntiles;
  for (int s = 0; s < 4; ++s) {
    // k_s = RHS(u_s): stage 0 reads the state buffer, later stages the
    // stage buffer (both hybrid: whole-line compact interiors); k_s goes to
    // the array of table s (k_1, k_0 / k_3, k_2)
    int rc = launch_rhs(p, r->d_io + s, st, true, true);
    if (rc) return rc;
    // stage combine + projection; the next RHS reads ghosts from the
    // neighbours' interiors (gather), so no refill runs between stages
    double* const c0 = r->buf + 48 * r->sstride;
    double* const k1 = c0, *const acc = c0 + 24 * r->nacc, *const k03 = c0 + 48 * r->nacc, *const k2 = c0 + 72 * r->nacc;
    double* const kc = s == 0 || s == 3 ? k03 : (s == 1 ? k1 : k2);
    const Work W = {r->stg.v[0], r->dof, r->sstride, kc, acc, r->nacc, s == 1 ? k03 : k2};
    const dim3 g(std::max(nt, 1)), b(384);
    if (s == 0)
      hipLaunchKernelGGL(k_stage<ST_FIRST>, g, b, 0, st, p->d_blocks, p->d_tiles_blk, nt, r->h0, W, r->d_gdt,
                         p->chi_floor, s, r->d_flag);
    else if (s == 1)
      hipLaunchKernelGGL(k_stage<ST_SECOND>, g, b, 0, st, p->d_blocks, p->d_tiles_blk, nt, r->h0, W, r->d_gdt,
                         p->chi_floor, s, r->d_flag);
    else if (s == 2)
      hipLaunchKernelGGL(k_stage<ST_THIRD>, g, b, 0, st, p->d_blocks, p->d_tiles_blk, nt, r->h0, W, r->d_gdt,
                         p->chi_floor, s, r->d_flag);
    else
      hipLaunchKernelGGL(k_stage<ST_LAST>, g, b, 0, st, p->d_blocks, p->d_tiles_blk, nt, r->h0, W, r->d_gdt,
                         p->chi_floor, s, r->d_flag);
    // ghosts across level jumps (adaptive layouts) of the new stage state
    OutPtrs o;
    for (int v = 0; v < 24; ++v) o.v[v] = s < 3 ? r->stg.v[v] : r->h0.v[v];
    rc = launch_level_fill(p, o, st, true);
    if (rc) return rc;
  }
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

int set_group_dt(dgr_rk4_s* r, const double* dt, uint32_t n) {
  if (n < r->plan->ngroups) return DGR_E_INVALID;
  bool same = r->gdt_set;
  for (uint32_t g = 0; g < r->plan->ngroups; ++g) {
    if (!(dt[g] > 0.0)) return DGR_E_INVALID;
    same = same && r->h_gdt[g] == dt[g];
  }
  if (same) return DGR_OK;
  for (uint32_t g = 0; g < r->plan->ngroups; ++g) r->h_gdt[g] = dt[g];
  // stream-ordered before the next step; the host copy stays unchanged
  // until the next call, which first waits for the evolver's stream
  DGR_HIP(hipStreamSynchronize(r->st));
  DGR_HIP(hipMemcpyAsync(r->d_gdt, r->h_gdt.data(), sizeof(double) * r->plan->ngroups, hipMemcpyHostToDevice,
                         r->st));
  r->gdt_set = true;
  return DGR_OK;
}

int check_ptrs(double* const* u) {
  if (!u) return DGR_E_INVALID;
  for (int v = 0; v < 24; ++v)
    if (!u[v]) return DGR_E_INVALID;
  return DGR_OK;
}

void destroy(dgr_rk4_s* r) {
  if (r->exec) (void)hipGraphExecDestroy(r->exec);
  if (r->graph) (void)hipGraphDestroy(r->graph);
  if (r->st) (void)hipStreamDestroy(r->st);
  if (r->own_buf) (void)hipFree(r->buf);
  (void)hipFree(r->d_io);
  (void)hipFree(r->d_gdt);
  (void)hipFree(r->d_flag);
  delete r;
}

}  // namespace

extern "C" int dgr_ghost_refill(dgr_plan_t plan, double* const* u, void* stream) {
  if (!plan || check_ptrs(u)) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  Ptrs24 h;
  OutPtrs o;
  for (int v = 0; v < 24; ++v) h.v[v] = o.v[v] = u[v];
  hipLaunchKernelGGL(k_ghost_refill, dim3(grid_for(plan, (int)plan->nblocks)), dim3(RK_NT), 0, st, plan->d_blocks,
                     plan->d_nbr, (int)plan->nblocks, h);
  DGR_HIP(hipGetLastError());
  int rc = launch_level_fill(plan, o, st);
  if (rc) return rc;
  DGR_HIP(hipStreamSynchronize(st));
  return DGR_OK;
}

extern "C" int dgr_enforce_constraints(dgr_plan_t plan, double* const* u, void* stream) {
  if (!plan || check_ptrs(u)) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  Ptrs24 h;
  for (int v = 0; v < 24; ++v) h.v[v] = u[v];
  // the plan's flag word: no allocation per call (calls on one plan are
  // ordered by their synchronisation below)
  unsigned* d_flag = plan->d_flag;
  DGR_HIP(hipMemsetAsync(d_flag, 0, sizeof(unsigned), st));
  const int nt = (int)plan->ntiles;
  hipLaunchKernelGGL(k_enforce, dim3(grid_for(plan, nt)), dim3(384), 0, st, plan->d_blocks, plan->d_tiles_blk, nt, h,
                     plan->chi_floor, d_flag);
  DGR_HIP(hipGetLastError());
  unsigned flag = 0;
  DGR_HIP(hipMemcpyAsync(&flag, d_flag, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  DGR_HIP(hipStreamSynchronize(st));
  return flag ? DGR_E_CONSTRAINT : DGR_OK;
}

// state buffer and stage buffer (each per variable the padded layout, whose
// ghosts the level fills, the pack and the reset write, then the interior
// points in the compact region the stage kernels write); k_1, the running
// sum, k_0 / k_3 and k_2 compact
extern "C" uint64_t dgr_rk4_work_doubles(dgr_plan_t plan) {
  return plan ? 48ull * (plan->dof_end + plan->ncomp) + 96ull * 343ull * plan->ntiles : 0;
}

extern "C" int dgr_rk4_create_with_work(dgr_rk4_t* out, dgr_plan_t plan, double* const* u, double* work,
                                        int use_graph) {
  if (!out || !plan || check_ptrs(u)) return DGR_E_INVALID;
  dgr_rk4_s* r = new dgr_rk4_s();
  r->plan = plan;
  r->dof = plan->dof_end;
  r->sstride = plan->dof_end + plan->ncomp;
  r->use_graph = use_graph != 0;
  r->h_gdt.assign(plan->ngroups, 0.0);
  for (int v = 0; v < 24; ++v) r->state.v[v] = u[v];
  const size_t nb = sizeof(double) * r->dof;
  r->own_buf = work == nullptr;
  r->buf = work;
  r->nacc = 343ull * plan->ntiles;
  if ((r->own_buf && hipMalloc(&r->buf, sizeof(double) * dgr_rk4_work_doubles(plan)) != hipSuccess) ||
      hipMalloc(&r->d_io, sizeof(IOPtrs) * 4) != hipSuccess ||
      hipMalloc(&r->d_gdt, sizeof(double) * plan->ngroups) != hipSuccess || hipMalloc(&r->d_flag, 256) != hipSuccess) {
    destroy(r);
    return DGR_E_NOMEM;
  }
  Ptrs24& stg = r->stg;
  Ptrs24& k = r->k;
  for (int v = 0; v < 24; ++v) {
    r->h0.v[v] = r->buf + (size_t)v * r->sstride;
    stg.v[v] = r->buf + (size_t)(24 + v) * r->sstride;
    k.v[v] = r->buf + (size_t)48 * r->sstride + (size_t)v * r->nacc;
  }
  // RHS of stage s: input (the state buffer, then the stage buffer) -> k_1,
  // k_0 / k_3 or k_2 (compact arrays after the stage buffer)
  IOPtrs io[4];
  double* const kb = r->buf + (size_t)48 * r->sstride;
  for (int v = 0; v < 24; ++v) {
    io[0].u.v[v] = r->h0.v[v];
    io[0].r.v[v] = kb + (size_t)(48 + v) * r->nacc;  // k_0
    io[1].u.v[v] = stg.v[v];
    io[1].r.v[v] = k.v[v];                           // k_1
    io[2].u.v[v] = stg.v[v];
    io[2].r.v[v] = kb + (size_t)(72 + v) * r->nacc;  // k_2
    io[3].u.v[v] = stg.v[v];
    io[3].r.v[v] = kb + (size_t)(48 + v) * r->nacc;  // k_3 over k_0
  }
  // set-up copies on the evolver's own non-blocking stream, synchronised
  // alone (a copy on the legacy default stream could wait for unrelated work
  // sharing its hardware queue)
  hipError_t e = hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMemcpyAsync(r->d_io, io, sizeof(io), hipMemcpyHostToDevice, r->st);
  if (e == hipSuccess) e = hipMemsetAsync(r->d_flag, 0, 256, r->st);
  // ghosts the refill never reaches (no same-level neighbour) keep the
  // state's values in the stage buffer
  for (int v = 0; v < 24 && e == hipSuccess; ++v)
    e = hipMemcpyAsync(stg.v[v], r->state.v[v], nb, hipMemcpyDeviceToDevice, r->st);
  if (e == hipSuccess) e = hipStreamSynchronize(r->st);
  if (e != hipSuccess) {
    destroy(r);
    return hip_fail(e, "dgr_rk4_create");
  }
  *out = r;
  return DGR_OK;
}

extern "C" int dgr_rk4_create(dgr_rk4_t* out, dgr_plan_t plan, double* const* u, int use_graph) {
  return dgr_rk4_create_with_work(out, plan, u, nullptr, use_graph);
}

extern "C" int dgr_rk4_reset_stage(dgr_rk4_t r, void* stream) {
  if (!r) return DGR_E_INVALID;
  const hipStream_t st = (hipStream_t)stream;
  const size_t nb = sizeof(double) * r->dof;
  for (int v = 0; v < 24; ++v) DGR_HIP(hipMemcpyAsync(r->stg.v[v], r->state.v[v], nb, hipMemcpyDeviceToDevice, st));
  return DGR_OK;
}

extern "C" int dgr_rk4_set_group_dt(dgr_rk4_t r, const double* dt, uint32_t ngroups) {
  if (!r || !dt) return DGR_E_INVALID;
  return set_group_dt(r, dt, ngroups);
}

extern "C" int dgr_rk4_step(dgr_rk4_t r, double dt, void* stream) { return dgr_rk4_steps(r, dt, 1, stream); }

extern "C" int dgr_rk4_steps(dgr_rk4_t r, double dt, int nsteps, void* stream) {
  if (!r || !(dt > 0.0)) return DGR_E_INVALID;
  std::vector<double> all(r->plan->ngroups, dt);
  int rc = set_group_dt(r, all.data(), (uint32_t)all.size());
  if (rc) return rc;
  return dgr_rk4_run(r, nsteps, stream);
}

extern "C" int dgr_rk4_run(dgr_rk4_t r, int nsteps, void* stream) {
  if (!r || nsteps < 0 || !r->gdt_set) return DGR_E_INVALID;
  hipStream_t caller = (hipStream_t)stream;
  // order after the caller's work, run on the evolver's stream (graphs
  // cannot be captured on the legacy default stream), order the caller after
  hipEvent_t ev;
  DGR_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  hipError_t e = hipEventRecord(ev, caller);
  if (e == hipSuccess) e = hipStreamWaitEvent(r->st, ev, 0);
  if (e != hipSuccess) {
    (void)hipEventDestroy(ev);
    return hip_fail(e, "dgr_rk4_run");
  }
  int rc = DGR_OK;
  if (r->use_graph && !r->exec) {
    // capture one step once; later steps replay it
    e = hipStreamBeginCapture(r->st, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) rc = hip_fail(e, "hipStreamBeginCapture");
    if (rc == DGR_OK) {
      rc = enqueue_step(r, r->st);
      e = hipStreamEndCapture(r->st, &r->graph);
      if (rc == DGR_OK && e != hipSuccess) rc = hip_fail(e, "hipStreamEndCapture");
    }
    if (rc == DGR_OK) {
      e = hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphInstantiate");
    }
  }
  const int nb = (int)r->plan->nblocks;
  if (rc == DGR_OK && nsteps > 0) {
    // the caller's state into the state buffer
    hipLaunchKernelGGL(k_hybrid<true>, dim3(grid_xcd(r->plan, nb)), dim3(RK_NT), 0, r->st, r->plan->d_blocks,
                       r->plan->d_cdelta, nb, r->state, r->h0);
    e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "dgr_rk4_run: pack");
  }
  for (int n = 0; rc == DGR_OK && n < nsteps; ++n) {
    if (r->use_graph) {
      e = hipGraphLaunch(r->exec, r->st);
      if (e != hipSuccess) rc = hip_fail(e, "hipGraphLaunch");
    } else {
      rc = enqueue_step(r, r->st);
    }
  }
  if (rc == DGR_OK && nsteps > 0) {
    // the new state back into the caller's arrays, interiors and same-level
    // ghosts in one pass, then the ghosts at level jumps
    hipLaunchKernelGGL(k_hybrid<false>, dim3(grid_xcd(r->plan, nb)), dim3(RK_NT), 0, r->st, r->plan->d_blocks,
                       r->plan->d_cdelta, nb, r->state, r->h0);
    e = hipGetLastError();
    if (e != hipSuccess) rc = hip_fail(e, "dgr_rk4_run: unpack");
    if (rc == DGR_OK) {
      OutPtrs o;
      for (int v = 0; v < 24; ++v) o.v[v] = r->state.v[v];
      rc = launch_level_fill(r->plan, o, r->st, false);
    }
  }
  if (rc == DGR_OK) {
    e = hipEventRecord(ev, r->st);
    if (e == hipSuccess) e = hipStreamWaitEvent(caller, ev, 0);
    if (e != hipSuccess) rc = hip_fail(e, "dgr_rk4_run");
  }
  (void)hipEventDestroy(ev);
  return rc;
}

extern "C" int dgr_rk4_status(dgr_rk4_t r) {
  if (!r) return DGR_E_INVALID;
  unsigned flag = 0;
  DGR_HIP(hipStreamSynchronize(r->st));
  DGR_HIP(hipMemcpy(&flag, r->d_flag, sizeof(unsigned), hipMemcpyDeviceToHost));
  DGR_HIP(hipMemset(r->d_flag, 0, sizeof(unsigned)));
  return flag ? DGR_E_CONSTRAINT : DGR_OK;
}

extern "C" int dgr_rk4_destroy(dgr_rk4_t r) {
  if (!r) return DGR_E_INVALID;
  (void)hipStreamSynchronize(r->st);
  destroy(r);
  return DGR_OK;
}

// ---------------------------------------------------------------------------
// Explicit Runge-Kutta stepper with a Butcher tableau: the ETS time stepper
// (ODE/include/ets.h:337-445 tableaux, :469-545 evolve) driving the BSSN
// context (BSSN_GR/src/bssnCtx.cpp:151-181 rhs: constraint projection of the
// stage input, unzip, bssnRHS).  Per stage s
//     tmp = u + sum_{p<s} (a_sp dt) k_p     (p in order, on interior points)
//     enforce_bssn_constraints(tmp); ghost fill of tmp; k_s = RHS(tmp)
// then u += (b_k dt) k_k for k in order (VecFMA), and u's ghosts are
// refilled.  u itself is not projected (the reference projects only the
// stage inputs).
// ---------------------------------------------------------------------------
namespace {

constexpr int ETS_MAX_STAGES = 8;
struct EtsTerms {
  double c[ETS_MAX_STAGES];
  int idx[ETS_MAX_STAGES];
  int n;
};

// out = base + sum_q c[q] k_{idx[q]} on the interior points of every tile;
// k_p variable v in the compact layout of the RHS in gather mode, at
// kb + (p*24 + v) * nacc + (tile * 343 + local index); base is read at the
// padded points, out written to the compact interior region of a stage
// buffer (cdof doubles past each variable's padded points).  PROJECT: the
// stage input, projected (enforce_bssn_constraints) before it is stored; a
// point the projection rejects (det gt < 0) keeps the combined values, as
// the projection leaves such a node untouched.
template <bool PROJECT>
__global__ void __launch_bounds__(384) k_ets_combine(const DevBlock* __restrict__ blocks,
                                                    const DevTile* __restrict__ tiles, int ntiles, Ptrs24 out,
                                                    Ptrs24 base, const double* __restrict__ kb, uint64_t nacc,
                                                    EtsTerms T, double chi_floor, unsigned* __restrict__ flag,
                                                    uint64_t cdof) {
  for (int w = blockIdx.x; w < ntiles; w += gridDim.x) {
    const DevTile tl = tiles[w];
    const DevBlock& B = blocks[tl.block];
    uint64_t o, oc;
    if (!tile_point(tl, B, threadIdx.x, o, &oc)) continue;
    const uint64_t co = (uint64_t)w * 343 + threadIdx.x;
    if constexpr (PROJECT) {
      double u[24];
#pragma unroll
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        double x = base.v[v][o];
        for (int q = 0; q < T.n; ++q) x = x + T.c[q] * kb[((uint64_t)T.idx[q] * 24 + v) * nacc + co];
        u[v] = x;
      }
      const int bad = enforce_point(u, chi_floor);
      if (bad) atomicOr(flag, (unsigned)bad);
      // a stage input: the compact interior region of the stage buffer
#pragma unroll
      for (int v = 0; v < 24; ++v) out.v[v][cdof + oc] = u[v];
    } else {
      for (int v = 0; v < 24; ++v) {
#pragma clang fp contract(off)
        double x = base.v[v][o];
        for (int q = 0; q < T.n; ++q) x = x + T.c[q] * kb[((uint64_t)T.idx[q] * 24 + v) * nacc + co];
        out.v[v][cdof + oc] = x;
      }
    }
  }
}

}  // namespace

struct dgr_ets_s {
  dgr_plan_t plan = nullptr;
  uint64_t dof = 0;
  int ns = 0;
  std::vector<double> a, b, c;  // a[s*ns + p], b[k], c[s]
  double time = 0.0;
  uint64_t nacc = 0;      // compact k slots per variable (343 per tile)
  uint64_t sstride = 0;   // tmp per variable: dof padded points, then the compact interior region
  double* buf = nullptr;  // tmp (24 x sstride) then k_0 .. k_{ns-1} (24 x nacc each, compact)
  Ptrs24 state{}, tmp{};
  IOPtrs* d_io = nullptr;  // RHS tables: stage s reads tmp, writes k_s
  unsigned* d_flag = nullptr;
};

namespace {
void ets_destroy(dgr_ets_s* e) {
  (void)hipFree(e->buf);
  (void)hipFree(e->d_io);
  (void)hipFree(e->d_flag);
  delete e;
}
}  // namespace

extern "C" int dgr_ets_create_tableau(dgr_ets_t* out, dgr_plan_t plan, double* const* u, const double* aij,
                                      const double* bi, const double* ci, uint32_t stages) {
  if (!out || !plan || check_ptrs(u) || !aij || !bi || !ci || stages < 1 || stages > ETS_MAX_STAGES)
    return DGR_E_INVALID;
  // explicit: a_sp = 0 for p >= s
  for (uint32_t s = 0; s < stages; ++s)
    for (uint32_t p = s; p < stages; ++p)
      if (aij[s * stages + p] != 0.0) return DGR_E_INVALID;
  dgr_ets_s* e = new dgr_ets_s();
  e->plan = plan;
  e->dof = plan->dof_end;
  e->ns = (int)stages;
  e->a.assign(aij, aij + stages * stages);
  e->b.assign(bi, bi + stages);
  e->c.assign(ci, ci + stages);
  for (int v = 0; v < 24; ++v) e->state.v[v] = u[v];
  const size_t nb = sizeof(double) * e->dof;
  e->nacc = 343ull * plan->ntiles;
  e->sstride = plan->dof_end + plan->ncomp;
  const size_t nbuf = sizeof(double) * 24 * (e->sstride + (size_t)stages * e->nacc);
  if (hipMalloc(&e->buf, nbuf) != hipSuccess ||
      hipMalloc(&e->d_io, sizeof(IOPtrs) * stages) != hipSuccess || hipMalloc(&e->d_flag, 256) != hipSuccess) {
    ets_destroy(e);
    return DGR_E_NOMEM;
  }
  std::vector<IOPtrs> io(stages);
  for (int v = 0; v < 24; ++v) e->tmp.v[v] = e->buf + (size_t)v * e->sstride;
  for (uint32_t s = 0; s < stages; ++s)
    for (int v = 0; v < 24; ++v) {
      io[s].u.v[v] = e->tmp.v[v];
      io[s].r.v[v] = e->buf + (size_t)24 * e->sstride + ((size_t)s * 24 + v) * e->nacc;
    }
  hipError_t er = hipMemcpy(e->d_io, io.data(), sizeof(IOPtrs) * stages, hipMemcpyHostToDevice);
  if (er == hipSuccess) er = hipMemset(e->d_flag, 0, 256);
  if (er == hipSuccess) er = hipMemset(e->buf, 0, nbuf);
  // ghosts no fill reaches keep the state's values in the stage input
  for (int v = 0; v < 24 && er == hipSuccess; ++v)
    er = hipMemcpy(e->tmp.v[v], e->state.v[v], nb, hipMemcpyDeviceToDevice);
  if (er != hipSuccess) {
    ets_destroy(e);
    return hip_fail(er, "dgr_ets_create");
  }
  *out = e;
  return DGR_OK;
}

extern "C" int dgr_ets_create(dgr_ets_t* out, dgr_plan_t plan, double* const* u, int type) {
  // ETS::set_ets_coefficients(ETSType) (ODE/include/ets.h:388-436); RK5 is
  // disabled there and here
  if (type == DGR_ETS_RK3) {
    const double a[9] = {0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 1.0 / 4.0, 1.0 / 4.0, 0.0};
    const double b[3] = {1.0 / 6.0, 1.0 / 6.0, 2.0 / 3.0};
    const double c[3] = {0.0, 1.0, 1.0 / 2.0};
    return dgr_ets_create_tableau(out, plan, u, a, b, c, 3);
  }
  if (type == DGR_ETS_RK4) {
    const double a[16] = {0.0, 0.0, 0.0, 0.0, 1.0 / 2.0, 0.0, 0.0, 0.0, 0.0, 1.0 / 2.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0};
    const double b[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
    const double c[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
    return dgr_ets_create_tableau(out, plan, u, a, b, c, 4);
  }
  return DGR_E_UNSUPPORTED;
}

extern "C" int dgr_ets_evolve(dgr_ets_t e, double dt, void* stream) {
  if (!e || !(dt > 0.0)) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  dgr_plan_t p = e->plan;
  const int nt = (int)p->ntiles;
  const int g = grid_for(p, nt);
  const double* kb = e->buf + (size_t)24 * e->sstride;
  OutPtrs tmp_o;
  for (int v = 0; v < 24; ++v) tmp_o.v[v] = e->tmp.v[v];
  for (int s = 0; s < e->ns; ++s) {
    EtsTerms T{};
    for (int q = 0; q < s; ++q) {
      const double aq = e->a[(size_t)s * e->ns + q];
      if (aq == 0.0) continue;  // adding (0 dt) k leaves the sum unchanged
      T.c[T.n] = aq * dt;
      T.idx[T.n++] = q;
    }
    // the stage input goes to tmp's compact region (whole lines), which the
    // level-jump fills and the gathering RHS read
    hipLaunchKernelGGL(k_ets_combine<true>, dim3(g), dim3(384), 0, st, p->d_blocks, p->d_tiles_blk, nt, e->tmp,
                       e->state, kb, e->nacc, T, p->chi_floor, e->d_flag, e->dof);
    int rc = launch_level_fill(p, tmp_o, st, true);
    if (rc) return rc;
    rc = launch_rhs(p, e->d_io + s, st, true, true);
    if (rc) return rc;
  }
  EtsTerms T{};
  for (int k = 0; k < e->ns; ++k) {
    T.c[k] = e->b[k] * dt;
    T.idx[k] = k;
  }
  T.n = e->ns;
  // the new state into tmp's compact region (the last RHS has read it),
  // then into u's interiors and same-level ghosts in one pass
  hipLaunchKernelGGL(k_ets_combine<false>, dim3(g), dim3(384), 0, st, p->d_blocks, p->d_tiles_blk, nt, e->tmp,
                     e->state, kb, e->nacc, T, p->chi_floor, e->d_flag, e->dof);
  const int nb = (int)p->nblocks;
  hipLaunchKernelGGL(k_hybrid<false>, dim3(grid_xcd(p, nb)), dim3(RK_NT), 0, st, p->d_blocks, p->d_cdelta, nb,
                     e->state, e->tmp);
  DGR_HIP(hipGetLastError());
  OutPtrs so;
  for (int v = 0; v < 24; ++v) so.v[v] = e->state.v[v];
  int rc = launch_level_fill(p, so, st);
  if (rc) return rc;
  e->time += dt;
  return DGR_OK;
}

extern "C" double dgr_ets_time(dgr_ets_t e) { return e ? e->time : 0.0; }

extern "C" int dgr_ets_status(dgr_ets_t e, void* stream) {
  if (!e) return DGR_E_INVALID;
  hipStream_t st = (hipStream_t)stream;
  unsigned flag = 0;
  DGR_HIP(hipMemcpyAsync(&flag, e->d_flag, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  DGR_HIP(hipStreamSynchronize(st));
  DGR_HIP(hipMemsetAsync(e->d_flag, 0, sizeof(unsigned), st));
  DGR_HIP(hipStreamSynchronize(st));
  return flag ? DGR_E_CONSTRAINT : DGR_OK;
}

extern "C" int dgr_ets_destroy(dgr_ets_t e) {
  if (!e) return DGR_E_INVALID;
  (void)hipDeviceSynchronize();
  ets_destroy(e);
  return DGR_OK;
}

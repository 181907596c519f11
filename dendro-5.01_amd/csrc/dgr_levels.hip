/*
 * Ghost fill across level jumps of a 2:1-balanced block layout.
 *
 * Reference behaviour: Mesh::unzip (include/mesh.tcc:7436-8694) fills the
 * padding of a block whose neighbour is coarser by interpolating the coarse
 * element to the positions of the (virtual) fine element next to the block,
 * with the order-p reference element's parent-to-child operators applied
 * along x, then y, then z (FEM/include/refel.h:249-300), and the padding of a
 * block whose neighbour is finer from the fine element's nodes.
 *
 * Here, for blocks of one element of order 6 (sz = 13):
 *   fine side    a ghost box (3 layers deep, in one of the 26 directions)
 *                lies inside one element of the coarser neighbour; its
 *                points sit at half-integer multiples of the coarse node
 *                spacing, so each axis needs the 7 Lagrange weights of the
 *                coarse nodes at one of the 13 positions 0, 1/2, ..., 6.  The
 *                box is interpolated separably, x then y then z, in LDS
 *                (k_fill_interp, one wave per variable of a group of up to
 *                8 boxes interpolated from one coarse element);
 *   coarse side  every ghost point coincides with a node of a finer block
 *                (coarse spacing = 2 fine spacings) and copies it
 *                (k_fill_copy).
 * Ghosts with a same-level neighbour are not touched here (k_ghost_refill,
 * or the gathering RHS of the RK4 stages, handle them); ghosts outside the
 * domain keep their values.
 *
 * There is no buildable reference for this path (Mesh::unzip needs MPI and
 * the full mesh), so it is checked by what interpolation of order 6 must
 * do: polynomials of degree <= 6 are reproduced and copies are exact.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "../../include/dgr_bssn.h"
#include "dgr_internal.h"

using namespace dgr_internal;

namespace {

typedef __attribute__((address_space(1))) double gdouble;

constexpr int NF = 256;  // threads per interpolation workgroup

// Lagrange weights W[s][j] of node j (0..6) at position s/2, s = 0..12.
__device__ __forceinline__ void lagrange_table(double (*W)[7], int t) {
  for (int e = t; e < 13 * 7; e += NF) {
    const int s = e / 7, j = e % 7;
    const double x = 0.5 * s;
    double w = 1.0;
    for (int m = 0; m < 7; ++m)
      if (m != j) w *= (x - m) / (double)(j - m);
    W[s][j] = w;
  }
}

// Waves per interpolation workgroup: each wave interpolates one variable of
// one group on its own, so its sweeps synchronise only within the wave.
constexpr int FW = NF / 64;

// Lane-level order of a wave's LDS accesses: a wave's LDS operations run in
// issue order, so only the compiler has to be kept from moving them across
// this point (and the values other lanes wrote made visible to this one).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One task per wave: variable v of the group of boxes interpolated from one
// coarse element (items[groups[g]] .. items[groups[g+1]-1]), task g * 24 + v,
// so the waves of a workgroup take neighbouring variables of one element.
// Lane (b, c) < 49 of the wave holds the element's x line (b, c) of v in
// registers for the whole group; every box of the group is interpolated
// from them, x then y then z, through two LDS buffers of the wave.
__global__ void __launch_bounds__(NF) k_fill_interp(const FillItem* __restrict__ items,
                                                     const uint32_t* __restrict__ groups, uint32_t ngroups,
                                                     OutPtrs u, int stage) {
  __shared__ double W[13][7];
  __shared__ double sX[FW][343], sY[FW][343];
  const int lane = (int)threadIdx.x & 63, wv = (int)threadIdx.x >> 6;
  lagrange_table(W, (int)threadIdx.x);
  __syncthreads();
  double* const X = sX[wv];
  double* const Y = sY[wv];
  const uint64_t ntask = (uint64_t)ngroups * 24;
  for (uint64_t task = (uint64_t)blockIdx.x * FW + wv; task < ntask; task += (uint64_t)gridDim.x * FW) {
    const uint32_t g = (uint32_t)(task / 24);
    const int v = (int)(task % 24);
    const uint32_t i0 = groups[g], i1 = groups[g + 1];
    const FillItem E = items[i0];  // src and ssz are the group's
    // the element's nodes: padded layout, or a stage buffer's compact region
    const uint64_t es = stage ? E.csrc : E.src;
    const uint64_t e0 = stage ? E.ssz[0] - 6 : E.ssz[0], e1 = stage ? E.ssz[1] - 6 : E.ssz[1];
    const gdouble* src = (const gdouble*)u.v[v] + es;
    double line[7];  // C[a][b][c], a = 0..6, of lane (b, c)
    if (lane < 49) {
      const int b = lane % 7, c = lane / 7;
#pragma unroll
      for (int a = 0; a < 7; ++a) line[a] = src[a + e0 * (b + e1 * c)];
    }
    for (uint32_t it = i0; it < i1; ++it) {
      const FillItem I = items[it];
      const int nx = I.n[0], ny = I.n[1], nz = I.n[2];
      // One lane per line along the sweep's axis: the lane holds the
      // line's 7 values and forms every output of the line, with weights
      // that are the same for the whole wave (LDS broadcasts).
      // along x: X[p][b][c] = sum_a W[s0x + p][a] C[a][b][c]; lane = (b, c)
      if (lane < 49) {
        const int b = lane % 7, c = lane / 7;
        for (int p = 0; p < nx; ++p) {
          const double* wt = W[I.s0[0] + p];
          double r = 0.0;
#pragma unroll
          for (int a = 0; a < 7; ++a) r += wt[a] * line[a];
          X[p + nx * (b + 7 * c)] = r;
        }
      }
      wave_sync();
      // along y: Y[p][q][c] = sum_b W[s0y + q][b] X[p][b][c]; lane = (p, c)
      if (lane < nx * 7) {
        const int p = lane % nx, c = lane / nx;
        double col[7];
#pragma unroll
        for (int b = 0; b < 7; ++b) col[b] = X[p + nx * (b + 7 * c)];
        for (int q = 0; q < ny; ++q) {
          const double* wt = W[I.s0[1] + q];
          double r = 0.0;
#pragma unroll
          for (int b = 0; b < 7; ++b) r += wt[b] * col[b];
          Y[p + nx * (q + ny * c)] = r;
        }
      }
      wave_sync();
      // along z, into the fine block's padding; lane = (p, q)
      gdouble* dst = (gdouble*)u.v[v] + I.dst;
      if (lane < nx * ny) {
        const int p = lane % nx, q = lane / nx;
        double col[7];
#pragma unroll
        for (int c = 0; c < 7; ++c) col[c] = Y[p + nx * (q + ny * c)];
        for (int r3 = 0; r3 < nz; ++r3) {
          const double* wt = W[I.s0[2] + r3];
          double r = 0.0;
#pragma unroll
          for (int c = 0; c < 7; ++c) r += wt[c] * col[c];
          dst[p + (uint64_t)I.dsz[0] * (q + (uint64_t)I.dsz[1] * r3)] = r;
        }
      }
      wave_sync();  // the next box (or task) rewrites X and Y
    }
  }
}

__global__ void __launch_bounds__(256) k_fill_copy(const FillCopy* __restrict__ c, uint64_t n, OutPtrs u,
                                                   int stage) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) {
    const FillCopy C = c[i];
    const uint64_t src = stage ? C.csrc : C.src;  // a fine block's interior node
    double x[24];  // all loads before the stores
#pragma unroll
    for (int v = 0; v < 24; ++v) x[v] = ((const gdouble*)u.v[v])[src];
#pragma unroll
    for (int v = 0; v < 24; ++v) ((gdouble*)u.v[v])[C.dst] = x[v];
  }
}

// Element lattice: blocks of one spacing keyed by the integer position of
// their element corner relative to a common origin.  Spacings are keyed by
// a rounded log2, so spacings that differ only by rounding match and
// factors of two are distinct.
struct Lattice {
  using Key = std::tuple<long long, long long, long long, long long, uint32_t>;
  std::map<Key, uint32_t> at;
  double org[3];

  static long long level(double h) { return llround(std::log2(h) * 64.0); }
  int find(double h, const long long ix[3], uint32_t group) const {
    auto it = at.find(Key(level(h), ix[0], ix[1], ix[2], group));
    return it == at.end() ? -1 : (int)it->second;
  }
};

bool regular(const DevBlock& B) {
  return B.sz[0] == 13 && B.sz[1] == 13 && B.sz[2] == 13 && std::fabs(B.h[1] - B.h[0]) < 1e-9 * B.h[0] &&
         std::fabs(B.h[2] - B.h[0]) < 1e-9 * B.h[0];
}

}  // namespace

int dgr_internal::build_level_fill(dgr_plan_t plan, const dgr_block* blocks, const std::vector<DevBlock>& hb) {
  (void)blocks;
  const uint32_t nb = plan->nblocks;
  Lattice L;
  for (int a = 0; a < 3; ++a) L.org[a] = 1e300;
  for (uint32_t b = 0; b < nb; ++b)
    if (regular(hb[b]))
      for (int a = 0; a < 3; ++a) L.org[a] = std::min(L.org[a], hb[b].pmin[a] + 3 * hb[b].h[a]);
  auto elem_index = [&](double x, double E, int a) { return (long long)std::floor((x - L.org[a]) / E + 1e-9); };
  for (uint32_t b = 0; b < nb; ++b) {
    const DevBlock& B = hb[b];
    if (!regular(B)) continue;
    const double E = 6 * B.h[0];
    long long ix[3];
    for (int a = 0; a < 3; ++a) ix[a] = llround((B.pmin[a] + 3 * B.h[a] - L.org[a]) / E);
    L.at.emplace(Lattice::Key(Lattice::level(B.h[0]), ix[0], ix[1], ix[2], B.group), b);
  }

  std::vector<FillItem> items;
  std::vector<FillCopy> copies;
  plan->coarse_dirs.assign(nb, 0u);
  // one spacing (per group) only: no level jumps to look for
  bool jumps = false;
  long long lev0 = 0;
  bool any = false;
  for (uint32_t b = 0; b < nb && !jumps; ++b) {
    if (!regular(hb[b])) continue;
    const long long l = Lattice::level(hb[b].h[0]);
    if (any && l != lev0) jumps = true;
    lev0 = l;
    any = true;
  }
  for (uint32_t b = 0; b < nb && jumps; ++b) {
    const DevBlock& B = hb[b];
    if (!regular(B)) continue;
    const int n = 13;
    const double h = B.h[0], E = 6 * h;
    for (int d = 0; d < NNBR; ++d) {
      if (d == 13 || plan->nbr[(size_t)b * NNBR + d] >= 0) continue;
      const int dd[3] = {d % 3 - 1, (d / 3) % 3 - 1, d / 9 - 1};
      // beyond a flagged (domain-boundary) face there is nothing to fill
      bool outside = false;
      for (int a = 0; a < 3; ++a)
        if (dd[a] != 0 && ((B.bflag >> (2 * a + (dd[a] > 0 ? 1 : 0))) & 1u)) outside = true;
      if (outside) continue;
      int lo[3], cnt[3];
      for (int a = 0; a < 3; ++a) {
        lo[a] = dd[a] < 0 ? 0 : (dd[a] > 0 ? n - 3 : 3);
        cnt[a] = dd[a] == 0 ? n - 6 : 3;
      }
      // coarser neighbour: the element of spacing 2h holding the box's middle
      double mid[3];
      long long ic[3];
      for (int a = 0; a < 3; ++a) {
        mid[a] = B.pmin[a] + (lo[a] + 0.5 * (cnt[a] - 1)) * h;
        ic[a] = elem_index(mid[a], 2 * E, a);
      }
      const int c = L.find(2 * h, ic, B.group);
      if (c >= 0) {
        const DevBlock& C = hb[c];
        FillItem it;
        std::memset(&it, 0, sizeof(it));
        bool ok = true;
        for (int a = 0; a < 3; ++a) {
          // first box point in half coarse spacings (= fine spacings) from C's corner
          const double s = (B.pmin[a] + lo[a] * h - (C.pmin[a] + 3 * C.h[a])) / h;
          const long long si = llround(s);
          ok = ok && std::fabs(s - si) < 1e-6 && si >= 0 && si + cnt[a] - 1 <= 12;
          it.s0[a] = (uint8_t)std::max<long long>(0, si);
          it.n[a] = (uint8_t)cnt[a];
          it.dsz[a] = B.sz[a];
          it.ssz[a] = C.sz[a];
        }
        if (!ok) continue;
        it.dst = B.offset + lo[0] + (uint64_t)B.sz[0] * (lo[1] + (uint64_t)B.sz[1] * lo[2]);
        it.src = C.offset + 3 + (uint64_t)C.sz[0] * (3 + (uint64_t)C.sz[1] * 3);
        it.csrc = plan->dof_end + C.cbase;  // node (3,3,3): the block's first interior point
        items.push_back(it);
        plan->coarse_dirs[b] |= 1u << d;
        continue;
      }
      // finer neighbours: each ghost point is a node of a block of spacing h/2
      for (int k = lo[2]; k < lo[2] + cnt[2]; ++k)
        for (int j = lo[1]; j < lo[1] + cnt[1]; ++j)
          for (int i = lo[0]; i < lo[0] + cnt[0]; ++i) {
            const int q[3] = {i, j, k};
            double x[3];
            long long f[3];
            for (int a = 0; a < 3; ++a) {
              x[a] = B.pmin[a] + q[a] * h;
              f[a] = elem_index(x[a], E / 2, a);
            }
            // a point on an element face belongs to both elements: try both
            int F = -1;
            for (int m = 0; m < 8 && F < 0; ++m) {
              long long g[3];
              bool valid = true;
              for (int a = 0; a < 3; ++a) {
                g[a] = f[a] - ((m >> a) & 1);
                if ((m >> a) & 1) {
                  const double r = (x[a] - L.org[a]) / (E / 2) - f[a];
                  valid = valid && std::fabs(r) < 1e-6;
                }
              }
              if (valid) F = L.find(h / 2, g, B.group);
            }
            if (F < 0) continue;
            const DevBlock& FB = hb[F];
            uint64_t src = FB.offset, csrc = plan->dof_end + FB.cbase;
            uint64_t stride = 1, cstride = 1;
            bool ok = true;
            for (int a = 0; a < 3; ++a) {
              const double xi = (x[a] - (FB.pmin[a] + 3 * FB.h[a])) / FB.h[a];
              const long long xr = llround(xi);
              ok = ok && std::fabs(xi - xr) < 1e-6 && xr >= 0 && xr <= 6;
              src += (uint64_t)(3 + xr) * stride;
              csrc += (uint64_t)xr * cstride;
              stride *= FB.sz[a];
              cstride *= FB.sz[a] - 6;
            }
            if (!ok) continue;
            copies.push_back({B.offset + i + (uint64_t)n * (j + (uint64_t)n * k), src, csrc});
          }
    }
  }
  // boxes interpolated from one coarse element form groups of at most
  // kGroupMax (one workgroup stages the element once per pass for all boxes
  // of its group; the cap keeps enough workgroups to fill the chip: 8
  // measured 26% faster than whole elements, ~24 boxes each, and than 4)
  constexpr uint32_t kGroupMax = 8;
  std::stable_sort(items.begin(), items.end(),
                   [](const FillItem& a, const FillItem& b) { return a.src < b.src; });
  std::vector<uint32_t> groups;
  for (uint32_t i = 0; i < (uint32_t)items.size(); ++i)
    if (i == 0 || items[i].src != items[i - 1].src || i - groups.back() >= kGroupMax) groups.push_back(i);
  groups.push_back((uint32_t)items.size());
  plan->nfill_items = (uint32_t)items.size();
  plan->nfill_groups = (uint32_t)groups.size() - 1;
  plan->nfill_copies = copies.size();
  // uploads on the plan's own stream (see dgr_plan_create_ex)
  const hipStream_t up = plan->aux;
  if (!items.empty()) {
    DGR_HIP(hipMalloc(&plan->d_fill_items, sizeof(FillItem) * items.size()));
    DGR_HIP(hipMemcpyAsync(plan->d_fill_items, items.data(), sizeof(FillItem) * items.size(), hipMemcpyHostToDevice,
                           up));
    DGR_HIP(hipMalloc(&plan->d_fill_groups, sizeof(uint32_t) * groups.size()));
    DGR_HIP(hipMemcpyAsync(plan->d_fill_groups, groups.data(), sizeof(uint32_t) * groups.size(),
                           hipMemcpyHostToDevice, up));
  }
  if (!copies.empty()) {
    DGR_HIP(hipMalloc(&plan->d_fill_copies, sizeof(FillCopy) * copies.size()));
    DGR_HIP(hipMemcpyAsync(plan->d_fill_copies, copies.data(), sizeof(FillCopy) * copies.size(),
                           hipMemcpyHostToDevice, up));
  }
  DGR_HIP(hipStreamSynchronize(up));
  return DGR_OK;
}

int dgr_internal::launch_level_fill(dgr_plan_t plan, const OutPtrs& u, hipStream_t st, bool stage) {
  if (plan->nfill_copies) {
    const uint64_t g = std::min<uint64_t>((plan->nfill_copies + 255) / 256, 65536);
    hipLaunchKernelGGL(k_fill_copy, dim3((unsigned)g), dim3(256), 0, st, plan->d_fill_copies, plan->nfill_copies,
                       u, (int)stage);
  }
  if (plan->nfill_items) {
    // one wave per group and variable
    const uint64_t g = std::min<uint64_t>(((uint64_t)plan->nfill_groups * 24 + FW - 1) / FW, 65536);
    hipLaunchKernelGGL(k_fill_interp, dim3((unsigned)g), dim3(NF), 0, st, plan->d_fill_items, plan->d_fill_groups,
                       plan->nfill_groups, u, (int)stage);
  }
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_plan_level_fills(dgr_plan_t plan, uint32_t* boxes, uint64_t* copies) {
  if (!plan) return DGR_E_INVALID;
  if (boxes) *boxes = plan->nfill_items;
  if (copies) *copies = plan->nfill_copies;
  return DGR_OK;
}

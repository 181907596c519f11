/*
 * Initial data on device: the puncture data of punctureData
 * (BSSN_GR/src/grUtils.cpp:398-702) evaluated at every padded point of every
 * block, so evolutions start device-resident without staging padded state
 * from the host.
 *
 * Per point (x = pmin + i h, as the RHS kernels place points): the
 * Brill-Lindquist conformal factor psi = 1 + m1/(2 r1) + m2/(2 r2) plus the
 * first-order spin / boost / spin-boost corrections of the reference
 * (arXiv:0711.1165, grUtils.cpp:478-606), alpha = max(psi^-2, floor),
 * chi = max(psi^-4, floor), gt = delta, K = beta = Gt = B = 0, and the
 * Bowen-York extrinsic curvature At_ij = sum over holes of
 * 3/(psi^6 r^3) (spin term + r/2 momentum term) (grUtils.cpp:644-700).
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "bssn_staged.h"
#include "dgr_internal.h"

namespace {

using namespace dgr_internal;

struct Ptrs24Out {
  double* v[24];
};

struct Hole {
  double m, c[3], p[3], s[3], pabs, sabs;
};
struct Holes {
  Hole h[2];
  double floor;
};

// conformal-factor correction of one hole at distance r along unit n
__device__ double psi_correction(const Hole& H, double r, const double n[3]) {
  double corr = 0.0;
  const double R = 2.0 * r / H.m;
  const double l = 1.0 / (1.0 + R);
  const double l2 = l * l, l3 = l2 * l, l4 = l3 * l, l5 = l4 * l;
  if (fabs(H.sabs) > 1.e-6) {  // spinning puncture
    const double J = 4.0 * H.sabs / (H.m * H.m);
    const double u0 = (l + l2 + l3 - 4.0 * l4 + 2.0 * l5) / 40.0;
    const double u2 = -l5 / 20.0;
    const double mu = (n[0] * H.s[0] + n[1] * H.s[1] + n[2] * H.s[2]) / fabs(H.sabs);
    const double P2 = (3.0 * mu * mu - 1.0) / 2.0;
    corr += J * J * (u0 + u2 * R * R * P2);
  }
  if (H.pabs > 1.e-6) {  // boosted puncture
    const double Pa = 2.0 * H.pabs / H.m;
    const double u0 = ((l - 2.0 * l2 + 2.0 * l3) - l4 + 0.20 * l5) * (5.0 / 32.0);
    double u2 = 15.0 * l + 132.0 * l2 + 53.0 * l3;
    u2 = u2 + 96.0 * l4 + 82.0 * l5;
    u2 = u2 + (84.0 / R) * (l5 + log(l) / R);
    u2 = u2 / (80.0 * R);
    const double mu = (n[0] * H.p[0] + n[1] * H.p[1] + n[2] * H.p[2]) / H.pabs;
    const double P2 = (3.0 * mu * mu - 1.0) / 2.0;
    corr += Pa * Pa * (u0 + u2 * P2);
  }
  if (H.pabs > 1.e-6 && fabs(H.sabs) > 1.e-6) {  // spinning and boosted
    const double pxs = (H.p[1] * H.s[2] - H.p[2] * H.s[1]) * n[0] + (H.p[2] * H.s[0] - H.p[0] * H.s[2]) * n[1] +
                       (H.p[0] * H.s[1] - H.p[1] * H.s[0]) * n[2];
    const double v1 = pxs * (16.0 / (H.m * H.m * H.m * H.m)) * r;
    const double v2 = 1.0 + 5.0 * R + 10.0 * R * R;
    corr += v1 * v2 * l5 / 80.0;
  }
  return corr;
}

__global__ void __launch_bounds__(256) k_init_puncture(const DevBlock* __restrict__ blocks, uint32_t nblocks,
                                                       Ptrs24Out u, Holes H) {
  for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const DevBlock& B = blocks[b];
    const int s0 = B.sz[0], s1 = B.sz[1], n = s0 * s1 * (int)B.sz[2];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const int i = e % s0, j = (e / s0) % s1, k = e / (s0 * s1);
      double x[3];
      {
#pragma clang fp contract(off)
        x[0] = B.pmin[0] + i * B.h[0];
        x[1] = B.pmin[1] + j * B.h[1];
        x[2] = B.pmin[2] + k * B.h[2];
      }
      double r[2], nv[2][3];
      double psi = 1.0, corr = 0.0;
      for (int q = 0; q < 2; ++q) {
        const Hole& Hq = H.h[q];
        const double d0 = x[0] - Hq.c[0], d1 = x[1] - Hq.c[1], d2 = x[2] - Hq.c[2];
        r[q] = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
        nv[q][0] = d0 / r[q];
        nv[q][1] = d1 / r[q];
        nv[q][2] = d2 / r[q];
        psi += Hq.m / (2.0 * r[q]);
      }
      for (int q = 0; q < 2; ++q) corr += psi_correction(H.h[q], r[q], nv[q]);
      const double psu = psi + corr;
      double* const* v = u.v;
      const uint64_t o = B.offset + (uint64_t)e;
      v[dgr::V_ALPHA][o] = fmax(1.0 / (psu * psu), H.floor);
      v[dgr::V_CHI][o] = fmax(1.0 / (psu * psu * psu * psu), H.floor);
      v[dgr::V_K][o] = 0.0;
      for (int a = 0; a < 3; ++a) {
        v[dgr::V_GT0 + a][o] = 0.0;
        v[dgr::V_BETA0 + a][o] = 0.0;
        v[dgr::V_B0 + a][o] = 0.0;
      }
      const double psu6 = psu * psu * psu * psu * psu * psu;
      for (int p = 0; p < 6; ++p) {
        const int a = dgr::spi(p), c = dgr::spj(p);
        v[dgr::V_SG0 + p][o] = a == c ? 1.0 : 0.0;
        double A = 0.0;
        for (int q = 0; q < 2; ++q) {
          const Hole& Hq = H.h[q];
          const double* nn = nv[q];
          // spin part: eps_a kl S_k n_l n_c + (a <-> c)
          const double sxn[3] = {Hq.s[1] * nn[2] - Hq.s[2] * nn[1], Hq.s[2] * nn[0] - Hq.s[0] * nn[2],
                                 Hq.s[0] * nn[1] - Hq.s[1] * nn[0]};
          const double v2 = sxn[a] * nn[c] + sxn[c] * nn[a];
          const double pn = Hq.p[0] * nn[0] + Hq.p[1] * nn[1] + Hq.p[2] * nn[2];
          const double v3 = Hq.p[a] * nn[c] + Hq.p[c] * nn[a] + pn * (nn[a] * nn[c] - (a == c ? 1.0 : 0.0));
          A += 3.0 / (psu6 * r[q] * r[q] * r[q]) * (v2 + (r[q] / 2.0) * v3);
        }
        v[dgr::V_SA0 + p][o] = A;
      }
    }
  }
}

// Perturbed flat space: the flat values plus, on every variable v,
// amp sum_b exp(-|x - c_vb|^2 / width^2) over the bumps in order, with the
// operations in the order of the host generator (pydgr.data.flat_bumps).
constexpr int kMaxBumps = 16;
__global__ void __launch_bounds__(256) k_init_bumps(const DevBlock* __restrict__ blocks, uint32_t nblocks,
                                                    Ptrs24Out u, const double* __restrict__ centers, int nbumps,
                                                    double amp, double iw2) {
  __shared__ double sc[24 * kMaxBumps * 3];
  for (int e = threadIdx.x; e < 24 * nbumps * 3; e += blockDim.x) sc[e] = centers[e];
  __syncthreads();
  for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const DevBlock& B = blocks[b];
    const int s0 = B.sz[0], s1 = B.sz[1], n = s0 * s1 * (int)B.sz[2];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
#pragma clang fp contract(off)
      const int i = e % s0, j = (e / s0) % s1, k = e / (s0 * s1);
      const double x = B.pmin[0] + i * B.h[0], y = B.pmin[1] + j * B.h[1], z = B.pmin[2] + k * B.h[2];
      const uint64_t o = B.offset + (uint64_t)e;
      for (int v = 0; v < 24; ++v) {
        double acc = 0.0;
        for (int q = 0; q < nbumps; ++q) {
          const double* c = sc + 3 * (v * nbumps + q);
          const double dx = x - c[0], dy = y - c[1], dz = z - c[2];
          acc = acc + exp(-(dx * dx + dy * dy + dz * dz) * iw2);
        }
        const bool one = v == dgr::V_ALPHA || v == dgr::V_CHI || v == dgr::V_SG0 || v == dgr::V_SG0 + 3 ||
                         v == dgr::V_SG0 + 5;
        u.v[v][o] = (one ? 1.0 : 0.0) + amp * acc;
      }
    }
  }
}

}  // namespace

extern "C" int dgr_init_bumps(dgr_plan_t plan, double* const* u, const double* centers, int nbumps, double amp,
                              double width, void* stream) {
  if (!plan || !u || !centers || nbumps < 0 || nbumps > kMaxBumps || !(width > 0.0)) return DGR_E_INVALID;
  Ptrs24Out pu;
  for (int v = 0; v < 24; ++v) {
    if (!u[v]) return DGR_E_INVALID;
    pu.v[v] = u[v];
  }
  if (plan->nblocks == 0) return DGR_OK;
  const uint32_t g = std::min<uint32_t>(plan->nblocks, 16384);
  hipLaunchKernelGGL(k_init_bumps, dim3(g), dim3(256), 0, (hipStream_t)stream, plan->d_blocks, plan->nblocks, pu,
                     centers, nbumps, amp, 1.0 / (width * width));
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_init_puncture(dgr_plan_t plan, double* const* u, const dgr_puncture* holes, void* stream) {
  if (!plan || !u || !holes) return DGR_E_INVALID;
  Ptrs24Out pu;
  for (int v = 0; v < 24; ++v) {
    if (!u[v]) return DGR_E_INVALID;
    pu.v[v] = u[v];
  }
  Holes H;
  for (int q = 0; q < 2; ++q) {
    const dgr_puncture& P = holes[q];
    if (!(P.mass > 0.0)) return DGR_E_INVALID;
    Hole& h = H.h[q];
    h.m = P.mass;
    for (int d = 0; d < 3; ++d) {
      h.c[d] = P.pos[d];
      h.p[d] = P.p[d];
    }
    h.pabs = std::sqrt(P.p[0] * P.p[0] + P.p[1] * P.p[1] + P.p[2] * P.p[2]);
    h.sabs = P.spin;
    h.s[0] = P.spin * std::sin(P.spin_theta) * std::cos(P.spin_phi);
    h.s[1] = P.spin * std::sin(P.spin_theta) * std::sin(P.spin_phi);
    h.s[2] = P.spin * std::cos(P.spin_theta);
  }
  H.floor = plan->chi_floor;
  if (plan->nblocks == 0) return DGR_OK;
  const uint32_t g = std::min<uint32_t>(plan->nblocks, 16384);
  hipLaunchKernelGGL(k_init_puncture, dim3(g), dim3(256), 0, (hipStream_t)stream, plan->d_blocks, plan->nblocks,
                     pu, H);
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

/*
 * Staged BSSN point equations (host + device).
 *
 * The 24 right-hand sides of the constant-eta, standard-gauge BSSN system
 * that the reference evaluates per point in bssnrhs (BSSN_GR/src/rhs.cpp:
 * 210-262, equations stated in CodeGen/bssn.py:60-88 with the helpers of
 * GR/rhs_scripts/bssn/dendro.py), rewritten so that a GPU thread consumes
 * each variable's derivatives as soon as they are computed and emits each
 * right-hand side as soon as its last input has arrived.  Call order:
 *
 *   begin()                     state -> inverse metric, A^{ij}
 *   add_gt(p, d1, d2, adv)      p = 0..5: d gt kept, d2 gt into Ricci
 *   finish_metric()             Christoffels, Gamma^i, quadratic Ricci part
 *   add_chi(d1, d2, adv)        conformal-factor Ricci part
 *   add_alpha(d1, d2, adv)      D_i D_j alpha, K laplacian   -> a_rhs
 *   add_Gt(k, d1, adv)          k = 0..2: d Gt^k Ricci part (k = 2 closes
 *                               the Ricci tensor into the At source)
 *   add_beta(k, d1, d2, adv)    k = 0..2: Lie / shift terms  -> b_rhs^k;
 *                               after k = 2: gt_rhs (6) and chi_rhs ready
 *   add_K(d1, adv)              -> K_rhs, Gt_rhs (3) ready
 *   add_B(k, adv, eta)          -> B_rhs^k
 *   add_At(p, adv)              -> At_rhs_p
 *
 * adv is beta^k times the upwinded derivative d_k of the variable.
 * Mathematically these are the reference's equations term by term (the
 * identity C1[i,j,k]+C1[j,i,k] = d_k gt_ij is used in one Ricci term, and
 * the trace-free part is split by linearity); the floating-point evaluation
 * order differs from the reference's CSE form, which the 1e-11 relative
 * parity budget allows (the derivative stencils, where the O(1) background
 * cancels, stay bit-identical in the kernels).
 *
 * Symmetric pairs are packed 00,01,02,11,12,22 -> 0..5.
 */
#ifndef DGR_BSSN_STAGED_H
#define DGR_BSSN_STAGED_H

#ifdef __HIPCC__
#define DGR_HD __host__ __device__ __forceinline__
#else
#define DGR_HD inline
#endif

namespace dgr {

// VAR order of BSSN_GR/include/grDef.h:31
enum Var {
  V_ALPHA = 0, V_CHI, V_K, V_GT0, V_GT1, V_GT2, V_BETA0, V_BETA1, V_BETA2, V_B0, V_B1, V_B2,
  V_SG0, V_SG1, V_SG2, V_SG3, V_SG4, V_SG5, V_SA0, V_SA1, V_SA2, V_SA3, V_SA4, V_SA5
};

DGR_HD constexpr int sp(int i, int j) {
  return (i <= j) ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
}
DGR_HD constexpr int spi(int p) { return p < 3 ? 0 : (p < 5 ? 1 : 2); }
DGR_HD constexpr int spj(int p) { return p < 3 ? p : (p < 5 ? p - 2 : 2); }
DGR_HD constexpr double spw(int p) { return (p == 0 || p == 3 || p == 5) ? 1.0 : 2.0; }

struct Params {
  double lambda[4];
  double lambda_f[2];
};

// Storage of the long-lived per-point arrays: plain registers by default;
// the GPU kernel can put them in LDS (one slot per thread) instead.
template <int N>
struct RegArr {
  double v[N];
  DGR_HD double& operator[](int i) { return v[i]; }
  DGR_HD const double& operator[](int i) const { return v[i]; }
};
struct RegStore {
  template <int N>
  using A = RegArr<N>;
};

template <class S>
struct PointT {
  // state at the point (Gt itself enters only through derivatives)
  double alpha, chi, K;
  double gt[6], At[6], beta[3], B[3];
  // derived
  double igt[6];
  double AUU[6];
  // d_k gt_p is kept in the storage of C2 until finish_metric() turns it
  // into the Christoffels: dgt(p,k) = C2[p*3 + k]
  typename S::template A<18> C2;  // C2[l*6 + p]
  double CalGt[3];
  double dchi[3];
  typename S::template A<6> R;
  // right-hand-side accumulators
  typename S::template A<6> gt_rhs, At_rhs;
  // small accumulators: Gt_rhs^i, the -lambda3 beta.adv(Gt^i) part of
  // B_rhs^i, w_j = d_j d_k beta^k, div beta, chi_rhs, K_rhs
  enum { A_GT = 0, A_B = 3, A_WB = 6, A_DIVB = 9, A_CHI = 10, A_K = 11 };
  typename S::template A<12> acc;

  DGR_HD double Gt_out(int k) const { return acc[A_GT + k]; }

  DGR_HD void begin() {
    const double g00 = gt[0], g01 = gt[1], g02 = gt[2], g11 = gt[3], g12 = gt[4], g22 = gt[5];
    const double c00 = g11 * g22 - g12 * g12;
    const double c01 = g02 * g12 - g01 * g22;
    const double c02 = g01 * g12 - g02 * g11;
    const double det = g00 * c00 + g01 * c01 + g02 * c02;
    const double idet = 1.0 / det;
    igt[0] = c00 * idet;
    igt[1] = c01 * idet;
    igt[2] = c02 * idet;
    igt[3] = (g00 * g22 - g02 * g02) * idet;
    igt[4] = (g01 * g02 - g00 * g12) * idet;
    igt[5] = (g00 * g11 - g01 * g01) * idet;
    double P[3][3];  // g^{ik} A_kl
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 3; ++l)
        P[i][l] = igt[sp(i, 0)] * At[sp(0, l)] + igt[sp(i, 1)] * At[sp(1, l)] + igt[sp(i, 2)] * At[sp(2, l)];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      AUU[p] = P[i][0] * igt[sp(0, j)] + P[i][1] * igt[sp(1, j)] + P[i][2] * igt[sp(2, j)];
    }
    for (int p = 0; p < 6; ++p) { R[p] = 0.0; gt_rhs[p] = 0.0; At_rhs[p] = 0.0; }
    for (int i = 0; i < 3; ++i) { acc[A_GT + i] = 0.0; acc[A_WB + i] = 0.0; acc[A_B + i] = 0.0; }
    acc[A_DIVB] = 0.0;
    acc[A_K] = 0.0;
    acc[A_CHI] = 0.0;
  }

  // -1/2 g^{lm} d_l d_m gt_ij (first Rt term of dendro.py compute_ricci)
  DGR_HD void add_gt(int p, const double d1[3], const double d2[6], double adv) {
    C2[p * 3 + 0] = d1[0]; C2[p * 3 + 1] = d1[1]; C2[p * 3 + 2] = d1[2];
    const double s = igt[0] * d2[0] + igt[3] * d2[3] + igt[5] * d2[5] +
                     2.0 * (igt[1] * d2[1] + igt[2] * d2[2] + igt[4] * d2[4]);
    R[p] -= 0.5 * s;
    gt_rhs[p] += adv;
  }

  DGR_HD void finish_metric() {
    double dgt[6][3];
    for (int p = 0; p < 6; ++p)
      for (int k = 0; k < 3; ++k) dgt[p][k] = C2[p * 3 + k];
    // C1[k][(i,j)] = 1/2 (d_j g_ki + d_i g_kj - d_k g_ij)
    double C1[3][6];
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < 6; ++p) {
        const int i = spi(p), j = spj(p);
        C1[k][p] = 0.5 * (dgt[sp(k, i)][j] + dgt[sp(k, j)][i] - dgt[p][k]);
      }
    for (int l = 0; l < 3; ++l)
      for (int p = 0; p < 6; ++p)
        C2[l * 6 + p] = igt[sp(l, 0)] * C1[0][p] + igt[sp(l, 1)] * C1[1][p] + igt[sp(l, 2)] * C1[2][p];
    for (int i = 0; i < 3; ++i)
      CalGt[i] = igt[0] * C2[i * 6 + 0] + igt[3] * C2[i * 6 + 3] + igt[5] * C2[i * 6 + 5] +
                 2.0 * (igt[1] * C2[i * 6 + 1] + igt[2] * C2[i * 6 + 2] + igt[4] * C2[i * 6 + 4]);
    // 1/2 Gamma^k (C1[i,j,k] + C1[j,i,k]) = 1/2 Gamma^k d_k gt_ij
    for (int p = 0; p < 6; ++p)
      R[p] += 0.5 * (CalGt[0] * dgt[p][0] + CalGt[1] * dgt[p][1] + CalGt[2] * dgt[p][2]);
    // g^{lm} (C2[k,l,i] C1[j,k,m] + C2[k,l,j] C1[i,k,m] + C2[k,i,m] C1[k,l,j])
    // with U[a][b][c] = g^{cm} C1[a][(b,m)]
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c) {
          const double U = igt[sp(c, 0)] * C1[a][sp(b, 0)] + igt[sp(c, 1)] * C1[a][sp(b, 1)] +
                           igt[sp(c, 2)] * C1[a][sp(b, 2)];
          // U[a][b][c] appears as U[j][k][l] (a=j,b=k,l=c) in the first two
          // terms and as U[k][j][l] (a=k,b=j,l=c) in the third
          for (int p = 0; p < 6; ++p) {
            const int i = spi(p), j = spj(p);
            double q = 0.0;
            if (a == j) q += C2[b * 6 + sp(c, i)];
            if (a == i) q += C2[b * 6 + sp(c, j)];
            if (b == j) q += C2[a * 6 + sp(i, c)];
            R[p] += q * U;
          }
        }
    // 2 alpha Gamma^i_jk A^{jk} in Gt_rhs
    for (int i = 0; i < 3; ++i) {
      const double s = C2[i * 6 + 0] * AUU[0] + C2[i * 6 + 3] * AUU[3] + C2[i * 6 + 5] * AUU[5] +
                       2.0 * (C2[i * 6 + 1] * AUU[1] + C2[i * 6 + 2] * AUU[2] + C2[i * 6 + 4] * AUU[4]);
      acc[A_GT + i] += 2.0 * alpha * s;
    }
  }

  // Ricci part from the conformal factor (compute_ricci, Rphi)
  DGR_HD void add_chi(const double d1[3], const double d2[6], double adv) {
    dchi[0] = d1[0]; dchi[1] = d1[1]; dchi[2] = d1[2];
    const double ichi = 1.0 / chi;
    const double h = 0.5 * ichi;
    double s = 0.0;
    for (int p = 0; p < 6; ++p) s += spw(p) * igt[p] * (d2[p] - 1.5 * ichi * d1[spi(p)] * d1[spj(p)]);
    s -= CalGt[0] * d1[0] + CalGt[1] * d1[1] + CalGt[2] * d1[2];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      const double c = C2[0 * 6 + p] * d1[0] + C2[1 * 6 + p] * d1[1] + C2[2 * 6 + p] * d1[2];
      R[p] += h * (d2[p] - c) - 0.25 * ichi * ichi * d1[i] * d1[j] + h * gt[p] * s;
    }
    acc[A_CHI] = adv;
  }

  // D_i D_j alpha with the full-metric Christoffels: its trace-free part
  // into At_rhs, its trace into K_rhs, and the alpha / chi terms of Gt_rhs.
  // Returns a_rhs.
  DGR_HD double add_alpha(const Params& P, const double d1[3], const double d2[6], double adv) {
    const double ichi = 1.0 / chi;
    double gca = 0.0;  // g^{lm} d_m chi d_l alpha
    for (int l = 0; l < 3; ++l)
      for (int m = 0; m < 3; ++m) gca += igt[sp(l, m)] * dchi[m] * d1[l];
    double DD[6];
    double lap = 0.0;
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      const double c3 = C2[0 * 6 + p] * d1[0] + C2[1 * 6 + p] * d1[1] + C2[2 * 6 + p] * d1[2] -
                        0.5 * ichi * (d1[i] * dchi[j] + d1[j] * dchi[i] - gt[p] * gca);
      DD[p] = d2[p] - c3;
      lap += spw(p) * igt[p] * DD[p];
    }
    for (int p = 0; p < 6; ++p) At_rhs[p] -= chi * (DD[p] - gt[p] * lap * (1.0 / 3.0));
    acc[A_K] -= chi * lap;
    for (int i = 0; i < 3; ++i) {
      double s = 0.0, t = 0.0;
      for (int j = 0; j < 3; ++j) {
        s += AUU[sp(i, j)] * d1[j];
        t += AUU[sp(i, j)] * dchi[j];
      }
      acc[A_GT + i] -= 2.0 * s + alpha * 3.0 * ichi * t;
    }
    return P.lambda[0] * adv - 2.0 * alpha * K;
  }

  // 1/2 (gt_ki d_j Gt^k + gt_kj d_i Gt^k); after k = 2 the Ricci tensor is
  // complete and its trace-free part enters At_rhs.
  DGR_HD void add_Gt(const Params& P, int k, const double d1[3], double adv) {
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      R[p] += 0.5 * (gt[sp(k, i)] * d1[j] + gt[sp(k, j)] * d1[i]);
    }
    acc[A_GT + k] += adv;
    acc[A_B + k] = -P.lambda[3] * adv;
    if (k == 2) {
      double tr = 0.0;
      for (int p = 0; p < 6; ++p) tr += spw(p) * igt[p] * R[p];
      for (int p = 0; p < 6; ++p) At_rhs[p] += chi * alpha * (R[p] - gt[p] * tr * (1.0 / 3.0));
    }
  }

  // Shift component k.  Returns b_rhs^k; after k = 2, gt_out[6] and
  // *chi_out hold gt_rhs and chi_rhs.
  DGR_HD double add_beta(const Params& P, int k, const double d1[3], const double d2[6], double adv,
                         double* gt_out, double* chi_out) {
    acc[A_DIVB] += d1[k];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      gt_rhs[p] += gt[sp(i, k)] * d1[j] + gt[sp(k, j)] * d1[i];
      At_rhs[p] += At[sp(i, k)] * d1[j] + At[sp(k, j)] * d1[i];
    }
    acc[A_GT + k] += igt[0] * d2[0] + igt[3] * d2[3] + igt[5] * d2[5] +
                 2.0 * (igt[1] * d2[1] + igt[2] * d2[2] + igt[4] * d2[4]);
    for (int j = 0; j < 3; ++j) acc[A_WB + j] += d2[sp(j, k)];
    acc[A_GT + k] -= CalGt[0] * d1[0] + CalGt[1] * d1[1] + CalGt[2] * d1[2];
    if (k == 2) {
      const double w = -2.0 / 3.0;
      for (int p = 0; p < 6; ++p) gt_out[p] = gt_rhs[p] + w * gt[p] * acc[A_DIVB] - 2.0 * alpha * At[p];
      *chi_out = acc[A_CHI] + w * chi * acc[A_DIVB] + (2.0 / 3.0) * (chi * alpha * K);
      for (int i = 0; i < 3; ++i)
        acc[A_GT + i] += (2.0 / 3.0) * CalGt[i] * acc[A_DIVB] +
                         (1.0 / 3.0) * (igt[sp(i, 0)] * acc[A_WB + 0] + igt[sp(i, 1)] * acc[A_WB + 1] +
                                        igt[sp(i, 2)] * acc[A_WB + 2]);
    }
    return 0.75 * (P.lambda_f[0] + P.lambda_f[1] * alpha) * B[k] + P.lambda[1] * adv;
  }

  // Returns K_rhs; Gt_rhs is final afterwards.
  DGR_HD double add_K(const double d1[3], double adv) {
    for (int i = 0; i < 3; ++i)
      acc[A_GT + i] -= alpha * (4.0 / 3.0) * (igt[sp(i, 0)] * d1[0] + igt[sp(i, 1)] * d1[1] + igt[sp(i, 2)] * d1[2]);
    double sqrA = 0.0;
    for (int p = 0; p < 6; ++p) sqrA += spw(p) * At[p] * AUU[p];
    return adv + acc[A_K] + alpha * (K * K * (1.0 / 3.0) + sqrA);
  }

  DGR_HD double add_B(const Params& P, int k, double adv, double eta) const {
    return acc[A_GT + k] - eta * B[k] + P.lambda[2] * adv + acc[A_B + k];
  }

  DGR_HD double add_At(int p, double adv) const {
    const int i = spi(p), j = spj(p);
    // A_ik g^{kl} A_lj
    double AA = 0.0;
    for (int k = 0; k < 3; ++k) {
      const double Pkj = igt[sp(k, 0)] * At[sp(0, j)] + igt[sp(k, 1)] * At[sp(1, j)] + igt[sp(k, 2)] * At[sp(2, j)];
      AA += At[sp(i, k)] * Pkj;
    }
    return adv + At_rhs[p] + (-2.0 / 3.0) * At[p] * acc[A_DIVB] + alpha * (K * At[p] - 2.0 * AA);
  }
};

using Point = PointT<RegStore>;

}  // namespace dgr

#endif

/*
 * Staged BSSN point equations (host + device).
 *
 * The 24 right-hand sides of the constant-eta, standard-gauge BSSN system
 * that the reference evaluates per point in bssnrhs (BSSN_GR/src/rhs.cpp:
 * 210-262, equations stated in CodeGen/bssn.py:60-88 with the helpers of
 * GR/rhs_scripts/bssn/dendro.py), rewritten so that a GPU thread can consume
 * each variable's derivatives as soon as they are computed instead of
 * holding all 207 point inputs at once:
 *
 *   begin()          state at the point -> inverse metric, A^{ij}, A_ik A^k_j
 *   add_gt(ij,...)   d gt_ij (kept) and d2 gt_ij (contracted into Ricci)
 *   finish_metric()  Christoffels of the first/second kind, Gamma^i (CalGt),
 *                    the quadratic Christoffel part of the Ricci tensor
 *   add_chi(...)     conformal-factor part of the Ricci tensor
 *   add_Gt(k,...)    d Gt^k part of the Ricci tensor
 *   add_alpha(...)   D_i D_j alpha, trace-free At source, K laplacian
 *   add_beta(k,...)  Lie-derivative and shift terms
 *   add_K(...)       d K in Gt_rhs
 *   add_adv(v,a)     beta^k (upwinded d_k var v), a per-variable advection sum
 *   finish(out)      the 24 right-hand sides in VAR order
 *
 * Mathematically these are the reference's equations term by term (the
 * identity C1[i,j,k]+C1[j,i,k] = d_k gt_ij is used in the Ricci term);
 * the floating-point evaluation order differs from the reference's CSE
 * form, which is allowed by the 1e-11 relative parity budget (the
 * derivative stencils, where cancellation is severe, stay bit-identical).
 *
 * Indices: symmetric pairs are packed 00,01,02,11,12,22 -> 0..5.
 * VAR order (BSSN_GR/include/grDef.h:31): alpha, chi, K, Gt0-2, beta0-2,
 * B0-2, gt0-5, At0-5.
 */
#ifndef DGR_BSSN_STAGED_H
#define DGR_BSSN_STAGED_H

#ifdef __HIPCC__
#define DGR_HD __host__ __device__ __forceinline__
#else
#define DGR_HD inline
#endif

namespace dgr {

enum Var {
  V_ALPHA = 0, V_CHI, V_K, V_GT0, V_GT1, V_GT2, V_BETA0, V_BETA1, V_BETA2, V_B0, V_B1, V_B2,
  V_SG0, V_SG1, V_SG2, V_SG3, V_SG4, V_SG5, V_SA0, V_SA1, V_SA2, V_SA3, V_SA4, V_SA5
};

// packed symmetric index of (i,j)
DGR_HD constexpr int sp(int i, int j) {
  return (i <= j) ? (i == 0 ? j : (i == 1 ? 2 + j : 5)) : (j == 0 ? i : (j == 1 ? 2 + i : 5));
}
// row/column of packed index p
DGR_HD constexpr int spi(int p) { return p < 3 ? 0 : (p < 5 ? 1 : 2); }
DGR_HD constexpr int spj(int p) { return p < 3 ? p : (p < 5 ? p - 2 : 2); }

struct Params {
  double lambda[4];
  double lambda_f[2];
};

struct Point {
  // state
  double alpha, chi, K;
  double gt[6], At[6], beta[3], Gt[3], B[3];
  // derived
  double igt[6];
  double AUU[6];   // A^{ij}
  double AA[6];    // A_ik g^{kl} A_lj
  double dgt[6][3];
  double C2[3][6];
  double CalGt[3];
  double dchi[3];
  double R[6];     // Ricci tensor accumulator
  // right-hand side accumulators
  double gt_rhs[6], At_rhs[6], Gt_rhs[3], B_rhs[3], b_rhs[3];
  double a_rhs, chi_rhs, K_rhs;
  double divb;
  double wb[3];    // sum_k d_j d_k beta^k
  double dbeta[3][3];  // dbeta[k][j] = d_j beta^k
  double adv[24];

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
    // P[i][l] = g^{ik} A_kl
    double P[3][3];
    for (int i = 0; i < 3; ++i)
      for (int l = 0; l < 3; ++l)
        P[i][l] = igt[sp(i, 0)] * At[sp(0, l)] + igt[sp(i, 1)] * At[sp(1, l)] + igt[sp(i, 2)] * At[sp(2, l)];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      AUU[p] = P[i][0] * igt[sp(0, j)] + P[i][1] * igt[sp(1, j)] + P[i][2] * igt[sp(2, j)];
      AA[p] = At[sp(i, 0)] * P[0][j] + At[sp(i, 1)] * P[1][j] + At[sp(i, 2)] * P[2][j];
    }
    for (int p = 0; p < 6; ++p) { R[p] = 0.0; gt_rhs[p] = 0.0; At_rhs[p] = 0.0; }
    for (int i = 0; i < 3; ++i) { Gt_rhs[i] = 0.0; wb[i] = 0.0; }
    divb = 0.0;
    K_rhs = 0.0;
  }

  // d gt_p and d2 gt_p (d2 packed 00,01,02,11,12,22): Ricci term
  // -1/2 g^{lm} d_l d_m gt_ij (dendro.py compute_ricci, first Rt term)
  DGR_HD void add_gt(int p, const double d1[3], const double d2[6]) {
    dgt[p][0] = d1[0]; dgt[p][1] = d1[1]; dgt[p][2] = d1[2];
    const double s = igt[0] * d2[0] + igt[3] * d2[3] + igt[5] * d2[5] +
                     2.0 * (igt[1] * d2[1] + igt[2] * d2[2] + igt[4] * d2[4]);
    R[p] -= 0.5 * s;
  }

  DGR_HD void finish_metric() {
    // C1[k][(i,j)] = 1/2 (d_j g_ki + d_i g_kj - d_k g_ij)
    double C1[3][6];
    for (int k = 0; k < 3; ++k)
      for (int p = 0; p < 6; ++p) {
        const int i = spi(p), j = spj(p);
        C1[k][p] = 0.5 * (dgt[sp(k, i)][j] + dgt[sp(k, j)][i] - dgt[p][k]);
      }
    // C2[l][(i,j)] = g^{lk} C1[k][(i,j)]
    for (int l = 0; l < 3; ++l)
      for (int p = 0; p < 6; ++p)
        C2[l][p] = igt[sp(l, 0)] * C1[0][p] + igt[sp(l, 1)] * C1[1][p] + igt[sp(l, 2)] * C1[2][p];
    // Gamma^i = g^{kl} C2[i][(k,l)]
    for (int i = 0; i < 3; ++i)
      CalGt[i] = igt[0] * C2[i][0] + igt[3] * C2[i][3] + igt[5] * C2[i][5] +
                 2.0 * (igt[1] * C2[i][1] + igt[2] * C2[i][2] + igt[4] * C2[i][4]);
    // U[a][b][c] = g^{cm} C1[a][(b,m)]
    double U[3][3][3];
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b)
        for (int c = 0; c < 3; ++c)
          U[a][b][c] = igt[sp(c, 0)] * C1[a][sp(b, 0)] + igt[sp(c, 1)] * C1[a][sp(b, 1)] +
                       igt[sp(c, 2)] * C1[a][sp(b, 2)];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      // 1/2 Gamma^k (C1[i,j,k] + C1[j,i,k]) = 1/2 Gamma^k d_k gt_ij
      double t = 0.5 * (CalGt[0] * dgt[p][0] + CalGt[1] * dgt[p][1] + CalGt[2] * dgt[p][2]);
      // g^{lm} (C2[k,l,i] C1[j,k,m] + C2[k,l,j] C1[i,k,m] + C2[k,i,m] C1[k,l,j])
      double q = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          q += C2[k][sp(l, i)] * U[j][k][l] + C2[k][sp(l, j)] * U[i][k][l] + C2[k][sp(i, l)] * U[k][j][l];
      R[p] += t + q;
    }
  }

  // Ricci part from the conformal factor (dendro.py compute_ricci, Rphi)
  DGR_HD void add_chi(const double d1[3], const double d2[6]) {
    dchi[0] = d1[0]; dchi[1] = d1[1]; dchi[2] = d1[2];
    const double ichi = 1.0 / chi;
    const double h = 0.5 * ichi;
    double s = 0.0;
    for (int p = 0; p < 6; ++p) {
      const int k = spi(p), l = spj(p);
      const double w = (k == l) ? 1.0 : 2.0;
      s += w * igt[p] * (d2[p] - 1.5 * ichi * d1[k] * d1[l]);
    }
    s -= CalGt[0] * d1[0] + CalGt[1] * d1[1] + CalGt[2] * d1[2];
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      const double c = C2[0][p] * d1[0] + C2[1][p] * d1[1] + C2[2][p] * d1[2];
      R[p] += h * (d2[p] - c) - 0.25 * ichi * ichi * d1[i] * d1[j] + h * gt[p] * s;
    }
  }

  // Ricci part 1/2 (gt_ki d_j Gt^k + gt_kj d_i Gt^k) for one k
  DGR_HD void add_Gt(int k, const double d1[3]) {
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      R[p] += 0.5 * (gt[sp(k, i)] * d1[j] + gt[sp(k, j)] * d1[i]);
    }
  }

  // D_i D_j alpha with the full-metric Christoffels, the At source, the
  // K laplacian and the alpha terms of Gt_rhs
  DGR_HD void add_alpha(const double d1[3], const double d2[6]) {
    const double ichi = 1.0 / chi;
    // g^{lm} d_m chi d_l alpha
    double gca = 0.0;
    for (int l = 0; l < 3; ++l)
      for (int m = 0; m < 3; ++m) gca += igt[sp(l, m)] * dchi[m] * d1[l];
    double M[6];
    double lap = 0.0;
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      const double c3 = C2[0][p] * d1[0] + C2[1][p] * d1[1] + C2[2][p] * d1[2] -
                        0.5 * ichi * (d1[i] * dchi[j] + d1[j] * dchi[i] - gt[p] * gca);
      const double DD = d2[p] - c3;
      M[p] = alpha * R[p] - DD;
      lap += ((i == j) ? 1.0 : 2.0) * igt[p] * DD;
    }
    double tr = 0.0;
    for (int p = 0; p < 6; ++p) tr += ((spi(p) == spj(p)) ? 1.0 : 2.0) * igt[p] * M[p];
    for (int p = 0; p < 6; ++p) At_rhs[p] += chi * (M[p] - gt[p] * tr * (1.0 / 3.0));
    K_rhs -= chi * lap;
    for (int i = 0; i < 3; ++i) {
      double s = 0.0, t = 0.0;
      for (int j = 0; j < 3; ++j) {
        s += AUU[sp(i, j)] * d1[j];
        t += AUU[sp(i, j)] * dchi[j];
      }
      Gt_rhs[i] -= 2.0 * s + alpha * 3.0 * ichi * t;
    }
    // 2 alpha Gamma^i_jk A^{jk}
    for (int i = 0; i < 3; ++i) {
      const double s = C2[i][0] * AUU[0] + C2[i][3] * AUU[3] + C2[i][5] * AUU[5] +
                       2.0 * (C2[i][1] * AUU[1] + C2[i][2] * AUU[2] + C2[i][4] * AUU[4]);
      Gt_rhs[i] += 2.0 * alpha * s;
    }
  }

  // shift component k: first derivatives d_j beta^k and second d2 beta^k
  DGR_HD void add_beta(int k, const double d1[3], const double d2[6]) {
    for (int j = 0; j < 3; ++j) dbeta[k][j] = d1[j];
    divb += d1[k];
    // Lie terms A_ik d_j beta^k + A_kj d_i beta^k
    for (int p = 0; p < 6; ++p) {
      const int i = spi(p), j = spj(p);
      gt_rhs[p] += gt[sp(i, k)] * d1[j] + gt[sp(k, j)] * d1[i];
      At_rhs[p] += At[sp(i, k)] * d1[j] + At[sp(k, j)] * d1[i];
    }
    // g^{jl} d_j d_l beta^k
    Gt_rhs[k] += igt[0] * d2[0] + igt[3] * d2[3] + igt[5] * d2[5] +
                 2.0 * (igt[1] * d2[1] + igt[2] * d2[2] + igt[4] * d2[4]);
    // w_j += d_j d_k beta^k
    for (int j = 0; j < 3; ++j) wb[j] += d2[sp(j, k)];
    // -Gamma^j d_j beta^k
    Gt_rhs[k] -= CalGt[0] * d1[0] + CalGt[1] * d1[1] + CalGt[2] * d1[2];
  }

  DGR_HD void add_K(const double d1[3]) {
    for (int i = 0; i < 3; ++i)
      Gt_rhs[i] -= alpha * (4.0 / 3.0) * (igt[sp(i, 0)] * d1[0] + igt[sp(i, 1)] * d1[1] + igt[sp(i, 2)] * d1[2]);
  }

  DGR_HD void finish(const Params& P, double eta, double* out) const {
    const double w = -2.0 / 3.0;
    double g[3];
    for (int i = 0; i < 3; ++i)
      g[i] = Gt_rhs[i] + (2.0 / 3.0) * CalGt[i] * divb +
             (1.0 / 3.0) * (igt[sp(i, 0)] * wb[0] + igt[sp(i, 1)] * wb[1] + igt[sp(i, 2)] * wb[2]) + adv[V_GT0 + i];
    out[V_ALPHA] = P.lambda[0] * adv[V_ALPHA] - 2.0 * alpha * K;
    out[V_CHI] = adv[V_CHI] + w * chi * divb + (2.0 / 3.0) * (chi * alpha * K);
    double sqrA = 0.0;
    for (int p = 0; p < 6; ++p) sqrA += ((spi(p) == spj(p)) ? 1.0 : 2.0) * At[p] * AUU[p];
    out[V_K] = adv[V_K] + K_rhs + alpha * (K * K * (1.0 / 3.0) + sqrA);
    for (int i = 0; i < 3; ++i) {
      out[V_GT0 + i] = g[i];
      out[V_BETA0 + i] = 0.75 * (P.lambda_f[0] + P.lambda_f[1] * alpha) * B[i] + P.lambda[1] * adv[V_BETA0 + i];
      out[V_B0 + i] = g[i] - eta * B[i] + P.lambda[2] * adv[V_B0 + i] - P.lambda[3] * adv[V_GT0 + i];
    }
    for (int p = 0; p < 6; ++p) {
      out[V_SG0 + p] = adv[V_SG0 + p] + gt_rhs[p] + w * gt[p] * divb - 2.0 * alpha * At[p];
      out[V_SA0 + p] = adv[V_SA0 + p] + At_rhs[p] + w * At[p] * divb + alpha * (K * At[p] - 2.0 * AA[p]);
    }
  }
};

}  // namespace dgr

#endif

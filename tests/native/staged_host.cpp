// Test harness: evaluates the product's staged BSSN point equations
// (dendro-5.01_amd/csrc/bssn_staged.h) on the host, in the call order the
// GPU kernel uses, fed from the oracle's 207-input point layout, so tests
// can compare them with the oracle's CSE-form equations and the
// reference's generated equations.
#include <cstddef>
#include "../../dendro-5.01_amd/csrc/bssn_staged.h"

extern "C" void staged_point_eqs(long n, const double* in, const double* eta, const double* lam,
                                 const double* lamf, double* out) {
  using namespace dgr;
  auto I = [&](int k, long pp) { return in[(size_t)k * n + pp]; };
  // grad block (15 vars x 3): alpha, beta0-2, chi, Gt0-2, K, gt0-5
  const int G_ALPHA = 24, G_BETA = 27, G_CHI = 36, G_GT = 39, G_K = 48, G_SG = 51;
  // grad2 block (11 vars x 6): gt0-5, chi, alpha, beta0-2
  const int H_SG = 69, H_CHI = 105, H_ALPHA = 111, H_BETA = 117;
  const int A0 = 135;  // agrad: 24 vars x 3 in VAR order
  Params P;
  for (int i = 0; i < 4; ++i) P.lambda[i] = lam[i];
  P.lambda_f[0] = lamf[0]; P.lambda_f[1] = lamf[1];
  for (long pp = 0; pp < n; ++pp) {
    Point s;
    s.alpha = I(0, pp); s.chi = I(1, pp); s.K = I(2, pp);
    for (int i = 0; i < 3; ++i) { s.beta[i] = I(6 + i, pp); s.B[i] = I(9 + i, pp); }
    for (int p = 0; p < 6; ++p) { s.gt[p] = I(12 + p, pp); s.At[p] = I(18 + p, pp); }
    auto adv = [&](int v) {
      return s.beta[0] * I(A0 + 3 * v, pp) + s.beta[1] * I(A0 + 3 * v + 1, pp) + s.beta[2] * I(A0 + 3 * v + 2, pp);
    };
    double o[24];
    double d1[3], d2[6];
    s.begin();
    for (int p = 0; p < 6; ++p) {
      for (int d = 0; d < 3; ++d) d1[d] = I(G_SG + 3 * p + d, pp);
      for (int q = 0; q < 6; ++q) d2[q] = I(H_SG + 6 * p + q, pp);
      s.add_gt(p, d1, d2, adv(V_SG0 + p));
    }
    s.finish_metric();
    for (int d = 0; d < 3; ++d) d1[d] = I(G_CHI + d, pp);
    for (int q = 0; q < 6; ++q) d2[q] = I(H_CHI + q, pp);
    s.add_chi(d1, d2, adv(V_CHI));
    for (int d = 0; d < 3; ++d) d1[d] = I(G_ALPHA + d, pp);
    for (int q = 0; q < 6; ++q) d2[q] = I(H_ALPHA + q, pp);
    o[V_ALPHA] = s.add_alpha(P, d1, d2, adv(V_ALPHA));
    for (int k = 0; k < 3; ++k) {
      for (int d = 0; d < 3; ++d) d1[d] = I(G_GT + 3 * k + d, pp);
      s.add_Gt(P, k, d1, adv(V_GT0 + k));
    }
    for (int k = 0; k < 3; ++k) {
      for (int d = 0; d < 3; ++d) d1[d] = I(G_BETA + 3 * k + d, pp);
      for (int q = 0; q < 6; ++q) d2[q] = I(H_BETA + 6 * k + q, pp);
      o[V_BETA0 + k] = s.add_beta(P, k, d1, d2, adv(V_BETA0 + k), o + V_SG0, o + V_CHI);
    }
    for (int d = 0; d < 3; ++d) d1[d] = I(G_K + d, pp);
    o[V_K] = s.add_K(d1, adv(V_K));
    for (int k = 0; k < 3; ++k) o[V_GT0 + k] = s.Gt_out(k);
    for (int k = 0; k < 3; ++k) o[V_B0 + k] = s.add_B(P, k, adv(V_B0 + k), eta[pp]);
    for (int p = 0; p < 6; ++p) o[V_SA0 + p] = s.add_At(p, adv(V_SA0 + p));
    for (int v = 0; v < 24; ++v) out[(size_t)v * n + pp] = o[v];
  }
}

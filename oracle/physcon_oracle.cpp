/*
 * CPU ORACLE -- test infrastructure only (see bssn_oracle.h).
 *
 * Restatement of the reference's constraint monitor on unzipped blocks,
 * written from its formulas, not copied:
 *
 *   orc_physcon          BSSN_GR/src/physcon.cpp:12-96 per block: deriv42
 *                        first / second derivatives of BSSN_GR/scripts/
 *                        constraint_derivs.h (mixed ones by composition),
 *                        then the point equations on interior points
 *   orc_physcon_point    the point equations: Hamiltonian and momentum
 *                        constraints and the Weyl scalar psi4 as stated in
 *                        GR/rhs_scripts/bssn/constraints.py (tetrad by
 *                        Gram-Schmidt with the physical metric gt/chi),
 *                        Christoffels / Ricci of GR/rhs_scripts/bssn/
 *                        dendro.py:393-527, psi4 = 0 within 1e-7 of the z
 *                        axis as BSSN_GR/src/physconeqs.cpp:1039-1044
 *
 * Pinning: golden vectors written by the reference's own generated
 * equations (BSSN_GR/src/physconeqs.cpp compiled by oracle/_ref, fixtures in
 * tests/golden/physcon_*.npz); closed forms in tests/test_oracle_physcon.py.
 * Compiled with -ffp-contract=off.
 */
#include "bssn_oracle.h"
#include "physcon_point_io.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace {

// state indices inside the physcon input order
enum { P_CHI = 0, P_K, P_GT0, P_G0 = 5, P_A0 = 11, P_X = 17 };

inline int sp(int i, int j) {
  static const int t[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  return t[i][j];
}

// Point equations from named inputs v[DGR_PHYSCON_NIN] (physcon_point_io.h).
void physcon_point(const double* v, double out[6]) {
  const double chi = v[P_CHI], K = v[P_K];
  const double x = v[P_X], y = v[P_X + 1], z = v[P_X + 2];
  double g[3][3], A[3][3], Gt[3];
  for (int i = 0; i < 3; ++i) {
    Gt[i] = v[P_GT0 + i];
    for (int j = 0; j < 3; ++j) {
      g[i][j] = v[P_G0 + sp(i, j)];
      A[i][j] = v[P_A0 + sp(i, j)];
    }
  }
  // first derivatives: grad_d_<var> at 20 + 3*var + d (state order)
  auto d1 = [&](int var, int d) { return v[20 + 3 * var + d]; };
  // second derivatives: grad2_<pair>_<var> at 71 + 6*k + pair, k = 0 chi, 1..6 gt
  auto d2 = [&](int k, int a, int b) { return v[71 + 6 * k + sp(a, b)]; };
  double dchi[3], dK[3], dGt[3][3], dg[3][3][3], dA[3][3][3], ddg[3][3][3][3], ddchi[3][3];
  for (int d = 0; d < 3; ++d) {
    dchi[d] = d1(P_CHI, d);
    dK[d] = d1(P_K, d);
    for (int k = 0; k < 3; ++k) dGt[d][k] = d1(P_GT0 + k, d);  // dGt[d][k] = d_d Gt^k
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        dg[d][i][j] = d1(P_G0 + sp(i, j), d);
        dA[d][i][j] = d1(P_A0 + sp(i, j), d);
      }
  }
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      ddchi[a][b] = d2(0, a, b);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) ddg[a][b][i][j] = d2(1 + sp(i, j), a, b);
    }

  // inverse conformal metric (adjugate / determinant)
  const double det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) -
                     g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
                     g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
  double ig[3][3];
  ig[0][0] = (g[1][1] * g[2][2] - g[1][2] * g[2][1]) / det;
  ig[0][1] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) / det;
  ig[0][2] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) / det;
  ig[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) / det;
  ig[1][2] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) / det;
  ig[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) / det;
  ig[1][0] = ig[0][1];
  ig[2][0] = ig[0][2];
  ig[2][1] = ig[1][2];

  // Christoffels: C1[k][i][j] = 1/2 (d_j g_ki + d_i g_kj - d_k g_ij),
  // C2[i][j][k] = g^il C1[l][j][k] (dendro.py:393-436)
  double C1[3][3][3], C2[3][3][3];
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) C1[k][i][j] = 0.5 * (dg[j][k][i] + dg[i][k][j] - dg[k][i][j]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double s = 0.0;
        for (int l = 0; l < 3; ++l) s += ig[i][l] * C1[l][j][k];
        C2[i][j][k] = s;
      }
  double CalGt[3];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int k = 0; k < 3; ++k)
      for (int l = 0; l < 3; ++l) s += ig[k][l] * C2[i][k][l];
    CalGt[i] = s;
  }
  // Ricci tensor R = Rt + Rphi (dendro.py:468-527)
  double R[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0;
      for (int l = 0; l < 3; ++l)
        for (int m = 0; m < 3; ++m) t1 += ig[l][m] * ddg[l][m][i][j];
      for (int k = 0; k < 3; ++k) {
        t2 += g[k][i] * dGt[j][k] + g[k][j] * dGt[i][k];
        t3 += CalGt[k] * (C1[i][j][k] + C1[j][i][k]);
        for (int l = 0; l < 3; ++l)
          for (int m = 0; m < 3; ++m)
            t4 += ig[l][m] * (C2[k][l][i] * C1[j][k][m] + C2[k][l][j] * C1[i][k][m] + C2[k][i][m] * C1[k][l][j]);
      }
      const double Rt = -0.5 * t1 + 0.5 * t2 + 0.5 * t3 + t4;
      double c = 0.0;
      for (int k = 0; k < 3; ++k) c += C2[k][j][i] * dchi[k];
      const double xr = 1.0 / (2.0 * chi) * (ddchi[i][j] - c) - 1.0 / (4.0 * chi * chi) * dchi[i] * dchi[j];
      double lap = 0.0, cd = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) lap += ig[k][l] * (ddchi[k][l] - 3.0 / (2.0 * chi) * dchi[k] * dchi[l]);
      for (int m = 0; m < 3; ++m) cd += CalGt[m] * dchi[m];
      R[i][j] = Rt + xr + 1.0 / (2.0 * chi) * g[i][j] * (lap - cd);
    }

  // Hamiltonian constraint (constraints.py: ham)
  double trR = 0.0, sqrA = 0.0;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) trR += chi * ig[j][k] * R[j][k];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) s += ig[i][k] * ig[j][l] * A[k][l];
      sqrA += A[i][j] * s;
    }
  out[0] = trR - sqrA + (2.0 / 3.0) * K * K;

  // momentum constraints (constraints.py: mom)
  for (int i = 0; i < 3; ++i) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 3; ++k) {
        double cc = 0.0;
        for (int m = 0; m < 3; ++m) cc += C2[m][k][i] * A[j][m];
        a += ig[j][k] * (dA[k][i][j] - cc);
        c += ig[j][k] * A[k][i] * dchi[j] / chi;
      }
    for (int j = 0; j < 3; ++j) b += Gt[j] * A[i][j];
    out[1 + i] = a - b - 1.5 * c - (2.0 / 3.0) * dK[i];
  }

  // psi4 (constraints.py: tetrad, Psi4); zero near the z axis
  if (std::fabs(x) <= 1e-7 && std::fabs(y) <= 1e-7) {
    out[4] = 0.0;
    out[5] = 0.0;
    return;
  }
  const double ichi = 1.0 / chi;
  double gd[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) gd[i][j] = g[i][j] * ichi;
  auto dot = [&](const double* a, const double* b) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) s += gd[i][j] * a[i] * b[j];
    return s;
  };
  double r[3] = {x, y, z};
  double th[3] = {x * z, y * z, -(x * x + y * y)};
  double ph[3] = {-y, x, 0.0};
  {
    const double n = std::sqrt(dot(r, r));
    for (int i = 0; i < 3; ++i) r[i] /= n;
  }
  {
    const double p1 = dot(th, th), p2 = dot(r, th);
    for (int i = 0; i < 3; ++i) th[i] -= p2 * r[i];
    const double n = std::sqrt(p1 - p2 * p2);
    for (int i = 0; i < 3; ++i) th[i] /= n;
  }
  {
    const double p1 = dot(ph, ph), p2 = dot(r, ph), p3 = dot(th, ph);
    for (int i = 0; i < 3; ++i) ph[i] -= p2 * r[i] + p3 * th[i];
    const double n = std::sqrt(p1 - p2 * p2 - p3 * p3);
    for (int i = 0; i < 3; ++i) ph[i] /= n;
  }
  const double is2 = 0.7071067811865475244;
  double mr[3], mi[3];
  for (int i = 0; i < 3; ++i) {
    mr[i] = th[i] * is2;
    mi[i] = ph[i] * is2;
  }
  double MM[3][3], NN[3][3], MR[3][3], NR[3][3], U[3][3], V[3][3], Av[3];
  for (int i = 0; i < 3; ++i) {
    Av[i] = 0.0;
    for (int j = 0; j < 3; ++j) Av[i] += A[i][j] * r[j];
    for (int j = 0; j < 3; ++j) {
      MM[i][j] = mr[i] * mr[j] - mi[i] * mi[j];
      NN[i][j] = mr[i] * mi[j] + mr[j] * mi[i];
      MR[i][j] = mr[i] * r[j] - mr[j] * r[i];
      NR[i][j] = mi[i] * r[j] - mi[j] * r[i];
      double u = 0.0, w = 0.0;
      for (int k = 0; k < 3; ++k) {
        double cc = 0.0;
        for (int m = 0; m < 3; ++m) cc += C2[m][k][i] * A[m][j];
        u += mr[k] * (dA[j][k][i] + cc);
        w += mi[k] * (dA[j][k][i] + cc);
      }
      U[i][j] = u;
      V[i][j] = w;
    }
  }
  double rdchi = 0.0, rA = 0.0, mrdchi = 0.0, midchi = 0.0, mrA = 0.0, miA = 0.0;
  for (int i = 0; i < 3; ++i) {
    rdchi += r[i] * dchi[i];
    rA += Av[i] * r[i];
    mrdchi += mr[i] * dchi[i];
    midchi += mi[i] * dchi[i];
    mrA += mr[i] * Av[i];
    miA += mi[i] * Av[i];
  }
  const double At_ = ichi * ichi * (rA + K * chi / 3.0 + 0.5 * rdchi);
  double p1r = 0.0, p1i = 0.0, p2r = 0.0, p2i = 0.0, p3r = 0.0, p3i = 0.0;
  for (int i = 0; i < 3; ++i) {
    p1r += R[i][i] * MM[i][i];
    p1i += R[i][i] * NN[i][i];
    p2r += A[i][i] * MM[i][i];
    p2i += A[i][i] * NN[i][i];
    for (int j = i + 1; j < 3; ++j) {
      p1r += 2.0 * R[i][j] * MM[i][j];
      p1i += 2.0 * R[i][j] * NN[i][j];
      p2r += 2.0 * A[i][j] * MM[i][j];
      p2i += 2.0 * A[i][j] * NN[i][j];
    }
    for (int j = 0; j < 3; ++j) {
      p3r += MR[i][j] * U[i][j] - NR[i][j] * V[i][j];
      p3i += NR[i][j] * U[i][j] + MR[i][j] * V[i][j];
    }
  }
  p2r *= At_;
  p2i *= At_;
  p3r *= ichi;
  p3i *= ichi;
  const double p4r = ichi * ichi * (mrA * (mrA + 0.5 * mrdchi) - miA * (miA + 0.5 * midchi));
  const double p4i = ichi * ichi * (mrA * (miA - 0.5 * midchi) + miA * (mrA - 0.5 * mrdchi));
  out[4] = p1r + p2r - p3r - p4r;
  out[5] = -(p1i + p2i - p3i - p4i);
}

}  // namespace

extern "C" int orc_num_physcon_inputs(void) { return DGR_PHYSCON_NIN; }

extern "C" void orc_physcon_point(long n, const double* in, double* out) {
  std::vector<double> v(DGR_PHYSCON_NIN);
  for (long pp = 0; pp < n; ++pp) {
    for (int k = 0; k < DGR_PHYSCON_NIN; ++k) v[k] = in[(size_t)k * n + pp];
    double o[6];
    physcon_point(v.data(), o);
    for (int c = 0; c < 6; ++c) out[(size_t)c * n + pp] = o[c];
  }
}

extern "C" void orc_physcon(double* const* con, const double* const* uz, uint64_t offset, const double* pmin,
                            const double* pmax, const unsigned* sz, unsigned bflag) {
  // VAR order (grDef.h:31) of the 17 state inputs
  static const int kVars[17] = {1, 2, 3, 4, 5, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23};
  const int nx = sz[0], ny = sz[1], nz = sz[2];
  const long n = (long)nx * ny * nz;
  const double hx = (pmax[0] - pmin[0]) / (nx - 1);
  const double hy = (pmax[1] - pmin[1]) / (ny - 1);
  const double hz = (pmax[2] - pmin[2]) / (nz - 1);
  const double h[3] = {hx, hy, hz};
  const double* u[17];
  for (int a = 0; a < 17; ++a) u[a] = uz[kVars[a]] + offset;
  // scratch (constraint_memalloc.h): grad of 17 variables, grad2 of 7
  std::vector<double> buf((size_t)n * (51 + 42), 0.0);
  auto G = [&](int a, int d) { return buf.data() + (size_t)n * (3 * a + d); };
  auto G2 = [&](int k, int pair) { return buf.data() + (size_t)n * (51 + 6 * k + pair); };
  // chi and gt0..5 take second derivatives: state slots 0 and 5..10
  const int g2var[7] = {0, 5, 6, 7, 8, 9, 10};
  // constraint_derivs.h: deriv42 first and pure second derivatives ...
  for (int d = 0; d < 3; ++d)
    for (int a = 0; a < 17; ++a) orc_deriv(0, d, G(a, d), u[a], h[d], sz, bflag, nullptr);
  for (int k = 0; k < 7; ++k) {
    orc_deriv(1, 0, G2(k, 0), u[g2var[k]], hx, sz, bflag, nullptr);
    orc_deriv(1, 1, G2(k, 3), u[g2var[k]], hy, sz, bflag, nullptr);
    orc_deriv(1, 2, G2(k, 5), u[g2var[k]], hz, sz, bflag, nullptr);
    // ... and the mixed ones by composition: d_y(d_x f), d_z(d_x f), d_z(d_y f)
    orc_deriv(0, 1, G2(k, 1), G(g2var[k], 0), hy, sz, bflag, nullptr);
    orc_deriv(0, 2, G2(k, 2), G(g2var[k], 0), hz, sz, bflag, nullptr);
    orc_deriv(0, 2, G2(k, 4), G(g2var[k], 1), hz, sz, bflag, nullptr);
  }
  std::vector<double> v(DGR_PHYSCON_NIN);
  for (int k = 3; k < nz - 3; ++k) {
    const double zc = pmin[2] + k * hz;
    for (int j = 3; j < ny - 3; ++j) {
      const double yc = pmin[1] + j * hy;
      for (int i = 3; i < nx - 3; ++i) {
        const double xc = pmin[0] + i * hx;
        const long pp = i + (long)nx * (j + (long)ny * k);
        int q = 0;
        for (int a = 0; a < 17; ++a) v[q++] = u[a][pp];
        v[q++] = xc;
        v[q++] = yc;
        v[q++] = zc;
        for (int a = 0; a < 17; ++a)
          for (int d = 0; d < 3; ++d) v[q++] = G(a, d)[pp];
        for (int kk = 0; kk < 7; ++kk)
          for (int pair = 0; pair < 6; ++pair) v[q++] = G2(kk, pair)[pp];
        double o[6];
        physcon_point(v.data(), o);
        for (int c = 0; c < 6; ++c) con[c][offset + pp] = o[c];
      }
    }
  }
}

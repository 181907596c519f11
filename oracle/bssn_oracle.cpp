/*
 * CPU ORACLE -- test infrastructure only (see bssn_oracle.h).
 *
 * Restatement of the reference's per-block BSSN RHS, written from the
 * reference's loop structure and formulas, not copied:
 *
 *   orc_deriv          BSSN_GR/src/derivs.cpp:11-1411 (deriv42_*, deriv42adv_*,
 *                      ko_deriv42_*), QuadGrav/src/derivs.cpp:2141-2952 (deriv64_*)
 *   orc_bssnrhs        BSSN_GR/src/rhs.cpp:86-392 with the derivative lists of
 *                      BSSN_GR/scripts/bssnrhs_{derivs,derivs_adv,ko_derivs}.h
 *   bcs (static)       BSSN_GR/src/rhs.cpp:838-982 (bssn_bcs) and the
 *                      (falloff, asymptote) table at rhs.cpp:265-324
 *   orc_point_eqs      the 24 point equations; the body is bssn_eqs_cse.inc,
 *                      emitted by dendro-5.01_amd/codegen/bssn_sym.py
 *   orc_enforce_constraints  BSSN_GR/src/bssn_constraints.cpp:12-172
 *   orc_rk4_*          BSSN_GR/src/rkBSSN.cpp:895-905, 996-1011; RK4_C/U at
 *                      BSSN_GR/include/rkBSSN.h:55-58
 *
 * Pinning (tests/test_oracle_*.py):
 *   point equations   golden vectors from the reference's generated
 *                     equations (tests/golden/, made by oracle/_ref)
 *   stencils          polynomial exactness of every interior/closure formula,
 *                     exact differences of monomials that fix each
 *                     normalisation constant, axis-permutation symmetry
 *   Sommerfeld        analytic f(x,y,z) with known derivatives
 *   constraints       det(gt)=1 and tr(A)=0 after projection, floors, and a
 *                     conformally scaled metric whose projection is known
 *   RK4               exact integration of du/dt = k with constant k and of a
 *                     linear ODE against the RK4 stability polynomial
 *
 * Floating point: this file is compiled with -ffp-contract=off so that each
 * operation rounds like the reference's scalar code.
 */
#include "bssn_oracle.h"
#include "bssn_point_io.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct Line {
  // Access along one axis of a padded block: at(q) is the value at axis
  // coordinate q on the line through the current point.
  const double* base;  // pointer to the point with axis coordinate 0
  long s;
  double at(int q) const { return base[(long)q * s]; }
};

enum Kind { FIRST = 0, SECOND = 1, ADV = 2, KO = 3, FIRST6 = 4, SECOND6 = 5, ADV6 = 6 };

// Loop ranges of the reference operators: deriv42_x covers i in [3,n-3),
// j,k in [1,n-1) (derivs.cpp:19-27); deriv42_y covers i,j in [3,n-3),
// k in [1,n-1) (:96-104); everything else the interior [3,n-3)^3.
// deriv64_x/y widen to [0,n) (QuadGrav/src/derivs.cpp:2147-2156).
void ranges(int kind, int axis, const unsigned* sz, int lo[3], int hi[3]) {
  for (int d = 0; d < 3; ++d) { lo[d] = 3; hi[d] = (int)sz[d] - 3; }
  int w = (kind == FIRST) ? 1 : (kind == FIRST6 ? 0 : -1);
  if (w >= 0) {
    if (axis == 0) { lo[1] = w; hi[1] = sz[1] - w; lo[2] = w; hi[2] = sz[2] - w; }
    if (axis == 1) { lo[2] = w; hi[2] = sz[2] - w; }
  }
}

// deriv42_{x,y,z} at axis coordinate q (derivs.cpp:29-65).
inline double first42(const Line& L, int q, int e, bool lf, bool hf, double idx_by_2, double idx_by_12) {
  if (hf && q == e - 1) return (L.at(e - 3) - 4.0 * L.at(e - 2) + 3.0 * L.at(e - 1)) * idx_by_2;
  if (hf && q == e - 2) return (-L.at(e - 3) + L.at(e - 1)) * idx_by_2;
  if (lf && q == 4) return (-L.at(3) + L.at(5)) * idx_by_2;
  if (lf && q == 3) return (-3.0 * L.at(3) + 4.0 * L.at(4) - L.at(5)) * idx_by_2;
  return (L.at(q - 2) - 8.0 * L.at(q - 1) + 8.0 * L.at(q + 1) - L.at(q + 2)) * idx_by_12;
}

// deriv42_{xx,yy,zz} (derivs.cpp:244-326).
inline double second42(const Line& L, int q, int e, bool lf, bool hf, double isq, double isq_by_12) {
  if (hf && q == e - 1) return (-L.at(e - 4) + 4.0 * L.at(e - 3) - 5.0 * L.at(e - 2) + 2.0 * L.at(e - 1)) * isq;
  if (hf && q == e - 2) return (L.at(e - 3) - 2.0 * L.at(e - 2) + L.at(e - 1)) * isq;
  if (lf && q == 4) return (L.at(3) - 2.0 * L.at(4) + L.at(5)) * isq;
  if (lf && q == 3) return (2.0 * L.at(3) - 5.0 * L.at(4) + 4.0 * L.at(5) - L.at(6)) * isq;
  return (-L.at(q - 2) + 16.0 * L.at(q - 1) - 30.0 * L.at(q) + 16.0 * L.at(q + 1) - L.at(q + 2)) * isq_by_12;
}

// deriv42adv_{x,y,z}: upwind side chosen by the sign of beta at the point
// (derivs.cpp:494-632).
inline double adv42(const Line& L, int q, int e, bool lf, bool hf, double b, double idx_by_2, double idx_by_12) {
  const bool pos = b > 0.0;
  if (hf && q == e - 1) return (L.at(e - 3) - 4.0 * L.at(e - 2) + 3.0 * L.at(e - 1)) * idx_by_2;
  if (hf && q == e - 2) {
    if (pos) return (-L.at(e - 3) + L.at(e - 1)) * idx_by_2;
    return (L.at(e - 4) - 4.0 * L.at(e - 3) + 3.0 * L.at(e - 2)) * idx_by_2;
  }
  if (hf && q == e - 3) {
    if (pos) return (-3.0 * L.at(e - 3) + 4.0 * L.at(e - 2) - L.at(e - 1)) * idx_by_2;
    return (-L.at(e - 6) + 6.0 * L.at(e - 5) - 18.0 * L.at(e - 4) + 10.0 * L.at(e - 3) + 3.0 * L.at(e - 2)) * idx_by_12;
  }
  if (lf && q == 3) return (-3.0 * L.at(3) + 4.0 * L.at(4) - L.at(5)) * idx_by_2;
  if (lf && q == 4) {
    if (pos) return (-3.0 * L.at(4) + 4.0 * L.at(5) - L.at(6)) * idx_by_2;
    return (-L.at(3) + L.at(5)) * idx_by_2;
  }
  if (lf && q == 5) {
    if (pos) return (-3.0 * L.at(4) - 10.0 * L.at(5) + 18.0 * L.at(6) - 6.0 * L.at(7) + L.at(8)) * idx_by_12;
    return (L.at(3) - 4.0 * L.at(4) + 3.0 * L.at(5)) * idx_by_2;
  }
  if (pos) return (-3.0 * L.at(q - 1) - 10.0 * L.at(q) + 18.0 * L.at(q + 1) - 6.0 * L.at(q + 2) + L.at(q + 3)) * idx_by_12;
  return (-L.at(q - 3) + 6.0 * L.at(q - 2) - 18.0 * L.at(q - 1) + 10.0 * L.at(q) + 3.0 * L.at(q + 1)) * idx_by_12;
}

// ko_deriv42_{x,y,z} (derivs.cpp:914-1077).
inline double ko42(const Line& L, int q, int e, bool lf, bool hf, double pre, double smr1, double smr2, double smr3) {
  if (hf && q == e - 1) return (L.at(e - 4) - 3.0 * L.at(e - 3) + 3.0 * L.at(e - 2) - L.at(e - 1)) / smr3;
  if (hf && q == e - 2) return (L.at(e - 5) - 6.0 * L.at(e - 4) + 12.0 * L.at(e - 3) - 10.0 * L.at(e - 2) + 3.0 * L.at(e - 1)) / smr2;
  if (hf && q == e - 3) return (L.at(e - 6) - 6.0 * L.at(e - 5) + 15.0 * L.at(e - 4) - 19.0 * L.at(e - 3) + 12.0 * L.at(e - 2) - 3.0 * L.at(e - 1)) / smr1;
  if (lf && q == 3) return (L.at(6) - 3.0 * L.at(5) + 3.0 * L.at(4) - L.at(3)) / smr3;
  if (lf && q == 4) return (L.at(7) - 6.0 * L.at(6) + 12.0 * L.at(5) - 10.0 * L.at(4) + 3.0 * L.at(3)) / smr2;
  if (lf && q == 5) return (L.at(8) - 6.0 * L.at(7) + 15.0 * L.at(6) - 19.0 * L.at(5) + 12.0 * L.at(4) - 3.0 * L.at(3)) / smr1;
  return pre * (-L.at(q - 3) + 6.0 * L.at(q - 2) - 15.0 * L.at(q - 1) + 20.0 * L.at(q) - 15.0 * L.at(q + 1) + 6.0 * L.at(q + 2) - L.at(q + 3));
}

// deriv64_{x,y,z} (QuadGrav/src/derivs.cpp:2141-2280); deriv64adv_* is the
// same centred operator (:2913-2952).
inline double first64(const Line& L, int q, int e, bool lf, bool hf, double idx_by_12, double idx_by_60) {
  if (hf && q == e - 1) return (3.0 * L.at(e - 5) - 16.0 * L.at(e - 4) + 36.0 * L.at(e - 3) - 48.0 * L.at(e - 2) + 25.0 * L.at(e - 1)) * idx_by_12;
  if (hf && q == e - 2) return (-1.0 * L.at(e - 5) + 6.0 * L.at(e - 4) - 18.0 * L.at(e - 3) + 10.0 * L.at(e - 2) + 3.0 * L.at(e - 1)) * idx_by_12;
  if (hf && q == e - 3) return (1.0 * L.at(e - 5) - 8.0 * L.at(e - 4) + 8.0 * L.at(e - 2) - 1.0 * L.at(e - 1)) * idx_by_12;
  if (lf && q == 3) return (-25.0 * L.at(3) + 48.0 * L.at(4) - 36.0 * L.at(5) + 16.0 * L.at(6) - 3.0 * L.at(7)) * idx_by_12;
  if (lf && q == 4) return (-3.0 * L.at(3) - 10.0 * L.at(4) + 18.0 * L.at(5) - 6.0 * L.at(6) + 1.0 * L.at(7)) * idx_by_12;
  if (lf && q == 5) return (1.0 * L.at(3) - 8.0 * L.at(4) + 8.0 * L.at(6) - 1.0 * L.at(7)) * idx_by_12;
  return (-1.0 * L.at(q - 3) + 9.0 * L.at(q - 2) - 45.0 * L.at(q - 1) + 45.0 * L.at(q + 1) - 9.0 * L.at(q + 2) + 1.0 * L.at(q + 3)) * idx_by_60;
}

// deriv64_{xx,yy,zz} (QuadGrav/src/derivs.cpp:2569-2680).
inline double second64(const Line& L, int q, int e, bool lf, bool hf, double isq_by_12, double isq_by_180) {
  if (hf && q == e - 1) return (11.0 * L.at(e - 5) - 56.0 * L.at(e - 4) + 114.0 * L.at(e - 3) - 104.0 * L.at(e - 2) + 35.0 * L.at(e - 1)) * isq_by_12;
  if (hf && q == e - 2) return (-1.0 * L.at(e - 5) + 4.0 * L.at(e - 4) + 6.0 * L.at(e - 3) - 20.0 * L.at(e - 2) + 11.0 * L.at(e - 1)) * isq_by_12;
  if (hf && q == e - 3) return (-1.0 * L.at(e - 5) + 16.0 * L.at(e - 4) - 30.0 * L.at(e - 3) + 16.0 * L.at(e - 2) - 1.0 * L.at(e - 1)) * isq_by_12;
  if (lf && q == 3) return (35.0 * L.at(3) - 104.0 * L.at(4) + 114.0 * L.at(5) - 56.0 * L.at(6) + 11.0 * L.at(7)) * isq_by_12;
  if (lf && q == 4) return (11.0 * L.at(3) - 20.0 * L.at(4) + 6.0 * L.at(5) + 4.0 * L.at(6) - 1.0 * L.at(7)) * isq_by_12;
  if (lf && q == 5) return (-1.0 * L.at(3) + 16.0 * L.at(4) - 30.0 * L.at(5) + 16.0 * L.at(6) - 1.0 * L.at(7)) * isq_by_12;
  return (2.0 * L.at(q - 3) - 27.0 * L.at(q - 2) + 270.0 * L.at(q - 1) - 490.0 * L.at(q) + 270.0 * L.at(q + 1) - 27.0 * L.at(q + 2) + 2.0 * L.at(q + 3)) * isq_by_180;
}

}  // namespace

extern "C" void orc_deriv(int kind, int axis, double* Du, const double* u, double h,
                          const unsigned* sz, unsigned bflag, const double* beta) {
  const long s3[3] = {1, (long)sz[0], (long)sz[0] * (long)sz[1]};
  const long s = s3[axis];
  const int e = (int)sz[axis] - 3;
  const bool lf = bflag & (1u << (2 * axis));
  const bool hf = bflag & (1u << (2 * axis + 1));
  int lo[3], hi[3];
  ranges(kind, axis, sz, lo, hi);

  const double idx = 1.0 / h;
  const double idx_by_2 = 0.5 * idx;
  const double idx_by_12 = idx / 12.0;
  const double idx_by_60 = idx / 60.0;
  const double isq = 1.0 / (h * h);
  const double isq_by_12 = isq / 12.0;
  const double isq_by_180 = isq / 180.0;
  const double pre = -1.0 / 64.0 / h;
  const double smr3 = 59.0 / 48.0 * 64 * h;
  const double smr2 = 43.0 / 48.0 * 64 * h;
  const double smr1 = 49.0 / 48.0 * 64 * h;

  for (int k = lo[2]; k < hi[2]; ++k)
    for (int j = lo[1]; j < hi[1]; ++j)
      for (int i = lo[0]; i < hi[0]; ++i) {
        const long pp = i + (long)sz[0] * (j + (long)sz[1] * k);
        const int c[3] = {i, j, k};
        const int q = c[axis];
        Line L{u + pp - (long)q * s, s};
        double v;
        switch (kind) {
          case FIRST: v = first42(L, q, e, lf, hf, idx_by_2, idx_by_12); break;
          case SECOND: v = second42(L, q, e, lf, hf, isq, isq_by_12); break;
          case ADV: v = adv42(L, q, e, lf, hf, beta[pp], idx_by_2, idx_by_12); break;
          case KO: v = ko42(L, q, e, lf, hf, pre, smr1, smr2, smr3); break;
          case FIRST6: case ADV6: v = first64(L, q, e, lf, hf, idx_by_12, idx_by_60); break;
          case SECOND6: v = second64(L, q, e, lf, hf, isq_by_12, isq_by_180); break;
          default: v = 0.0;
        }
        Du[pp] = v;
      }
}

namespace {

// Point equations: read the inputs named in bssn_point_io.h from per-input
// arrays D[k][pp] and evaluate the generated body.
inline void point_eqs(const double* const* D, long pp, double eta, const orc_params* p, double* out) {
#define ORC_DECL(name, k) const double name = D[k][pp];
  DGR_POINT_INPUTS(ORC_DECL)
#undef ORC_DECL
  const double lambda0 = p->lambda[0], lambda1 = p->lambda[1], lambda2 = p->lambda[2], lambda3 = p->lambda[3];
  const double lambda_f0 = p->lambda_f[0], lambda_f1 = p->lambda_f[1];
  (void)lambda0; (void)lambda1; (void)lambda2; (void)lambda3; (void)lambda_f0; (void)lambda_f1;
#include "bssn_eqs_cse.inc"
}

// Sommerfeld outgoing-wave terms on flagged faces (rhs.cpp:838-982):
// rhs = -(1/r)(x f_x + y f_y + z f_z + falloff (f - f_inf)) on the first /
// last interior plane of each flagged face; faces in the order
// left, right, down, up, back, front.
void bcs(double* f_rhs, const double* f, const double* dxf, const double* dyf, const double* dzf,
         const double* pmin, const double* pmax, double falloff, double asym, const unsigned* sz, unsigned bflag) {
  const int nx = sz[0], ny = sz[1], nz = sz[2];
  const double h[3] = {(pmax[0] - pmin[0]) / (nx - 1), (pmax[1] - pmin[1]) / (ny - 1), (pmax[2] - pmin[2]) / (nz - 1)};
  const int b[3] = {3, 3, 3};
  const int e[3] = {nx - 3, ny - 3, nz - 3};
  for (int face = 0; face < 6; ++face) {
    if (!(bflag & (1u << face))) continue;
    const int axis = face / 2;
    const int fixed = (face & 1) ? e[axis] - 1 : b[axis];
    for (int k = b[2]; k < e[2]; ++k)
      for (int j = b[1]; j < e[1]; ++j)
        for (int i = b[0]; i < e[0]; ++i) {
          const int c[3] = {i, j, k};
          if (c[axis] != fixed) continue;
          const double x = pmin[0] + i * h[0];
          const double y = pmin[1] + j * h[1];
          const double z = pmin[2] + k * h[2];
          const long pp = i + (long)nx * (j + (long)ny * k);
          const double inv_r = 1.0 / std::sqrt(x * x + y * y + z * z);
          f_rhs[pp] = -inv_r * (x * dxf[pp] + y * dyf[pp] + z * dzf[pp] + falloff * (f[pp] - asym));
        }
  }
}

// (falloff, asymptotic value) per variable in VAR order (rhs.cpp:267-320).
const double kFalloff[24] = {1, 1, 1, 2, 2, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2};
const double kAsym[24] = {1, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 1, 0, 1, 0, 0, 0, 0, 0, 0};

// VAR indices (grDef.h:31)
enum { ALPHA = 0, CHI, KK, GT0, GT1, GT2, BETA0, BETA1, BETA2, B0, B1, B2, SG0, SG1, SG2, SG3, SG4, SG5, SA0, SA1, SA2, SA3, SA4, SA5 };

}  // namespace

extern "C" int orc_num_point_inputs(void) { return DGR_POINT_NIN; }

extern "C" void orc_bssnrhs(double* const* rhs, const double* const* uz, uint64_t offset,
                            const double* pmin, const double* pmax, const unsigned* sz,
                            unsigned bflag, const orc_params* p) {
  const int nx = sz[0], ny = sz[1], nz = sz[2];
  const long n = (long)nx * ny * nz;
  const double hx = (pmax[0] - pmin[0]) / (nx - 1);
  const double hy = (pmax[1] - pmin[1]) / (ny - 1);
  const double hz = (pmax[2] - pmin[2]) / (nz - 1);
  const double h[3] = {hx, hy, hz};

  const double* u[24];
  double* r[24];
  for (int v = 0; v < 24; ++v) { u[v] = uz[v] + offset; r[v] = rhs[v] + offset; }

  // scratch: grad for all 24 vars (BCs need them), grad2 for the 11 vars
  // with second derivatives, agrad for all 24 (bssnrhs_memalloc*.h)
  std::vector<double> buf((size_t)n * (72 + 66 + 72), 0.0);
  auto G = [&](int d, int v) { return buf.data() + (size_t)n * (d * 24 + v); };
  auto G2 = [&](int pair, int k) { return buf.data() + (size_t)n * (72 + pair * 11 + k); };
  auto AG = [&](int d, int v) { return buf.data() + (size_t)n * (72 + 66 + d * 24 + v); };
  static const int kGrad2Vars[11] = {SG0, SG1, SG2, SG3, SG4, SG5, CHI, ALPHA, BETA0, BETA1, BETA2};
  int g2slot[24];
  for (int v = 0; v < 24; ++v) g2slot[v] = -1;
  for (int k = 0; k < 11; ++k) g2slot[kGrad2Vars[k]] = k;

  // operator set: deriv42 (BSSN_GR/include/rhs.h:23-37) or the 6th-order
  // option deriv64 / deriv64adv (= centred deriv64) with ko_deriv42
  // (QuadGrav/include/derivs.h:80-96)
  const bool six = p->deriv_order == 64;
  const int kFirst = six ? FIRST6 : FIRST, kSecond = six ? SECOND6 : SECOND, kAdv = six ? ADV6 : ADV;
  // first and pure second derivatives (bssnrhs_derivs.h:1-105)
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 24; ++v) {
      orc_deriv(kFirst, d, G(d, v), u[v], h[d], sz, bflag, nullptr);
      if (g2slot[v] >= 0) {
        const int pair = d == 0 ? 0 : (d == 1 ? 3 : 5);
        orc_deriv(kSecond, d, G2(pair, g2slot[v]), u[v], h[d], sz, bflag, nullptr);
      }
    }
  // mixed derivatives by composition (bssnrhs_derivs.h:106-138):
  // d01 = d_y(d_x f), d02 = d_z(d_x f), d12 = d_z(d_y f)
  for (int k = 0; k < 11; ++k) {
    const int v = kGrad2Vars[k];
    orc_deriv(kFirst, 1, G2(1, k), G(0, v), hy, sz, bflag, nullptr);
    orc_deriv(kFirst, 2, G2(2, k), G(0, v), hz, sz, bflag, nullptr);
    orc_deriv(kFirst, 2, G2(4, k), G(1, v), hz, sz, bflag, nullptr);
  }
  // upwinded derivatives (bssnrhs_derivs_adv.h)
  const double* beta[3] = {u[BETA0], u[BETA1], u[BETA2]};
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 24; ++v) orc_deriv(kAdv, d, AG(d, v), u[v], h[d], sz, bflag, beta[d]);

  // per-input pointer table in bssn_point_io.h order
  const double* D[DGR_POINT_NIN];
  {
    int k = 0;
    for (int v = 0; v < 24; ++v) D[k++] = u[v];
    static const int kGradVars[15] = {ALPHA, BETA0, BETA1, BETA2, CHI, GT0, GT1, GT2, KK, SG0, SG1, SG2, SG3, SG4, SG5};
    for (int a = 0; a < 15; ++a)
      for (int d = 0; d < 3; ++d) D[k++] = G(d, kGradVars[a]);
    for (int a = 0; a < 11; ++a)
      for (int pair = 0; pair < 6; ++pair) D[k++] = G2(pair, a);
    for (int v = 0; v < 24; ++v)
      for (int d = 0; d < 3; ++d) D[k++] = AG(d, v);
  }

  // point loop over the interior (rhs.cpp:210-262)
  for (int k = 3; k < nz - 3; ++k) {
    const double z = pmin[2] + k * hz;
    for (int j = 3; j < ny - 3; ++j) {
      const double y = pmin[1] + j * hy;
      for (int i = 3; i < nx - 3; ++i) {
        const double x = pmin[0] + i * hx;
        const long pp = i + (long)nx * (j + (long)ny * k);
        const double rc = std::sqrt(x * x + y * y + z * z);
        double eta = p->eta_const;
        if (rc >= p->eta_r0) eta *= std::pow(p->eta_r0 / rc, p->eta_damping_exp);
        double out[24];
        point_eqs(D, pp, eta, p, out);
        for (int v = 0; v < 24; ++v) r[v][pp] = out[v];
      }
    }
  }

  // Sommerfeld terms on flagged faces (rhs.cpp:265-324)
  if (bflag != 0)
    for (int v = 0; v < 24; ++v)
      bcs(r[v], u[v], G(0, v), G(1, v), G(2, v), pmin, pmax, kFalloff[v], kAsym[v], sz, bflag);

  // Kreiss-Oliger terms reuse the first-derivative scratch
  // (bssnrhs_ko_derivs.h) and are added on every interior point (rhs.cpp:331-373)
  for (int d = 0; d < 3; ++d)
    for (int v = 0; v < 24; ++v) orc_deriv(KO, d, G(d, v), u[v], h[d], sz, bflag, nullptr);
  const double sigma = p->ko_sigma;
  for (int k = 3; k < nz - 3; ++k)
    for (int j = 3; j < ny - 3; ++j)
      for (int i = 3; i < nx - 3; ++i) {
        const long pp = i + (long)nx * (j + (long)ny * k);
        for (int v = 0; v < 24; ++v) r[v][pp] += sigma * (G(0, v)[pp] + G(1, v)[pp] + G(2, v)[pp]);
      }
}

extern "C" void orc_bssnrhs_blocks(double* const* rhs, const double* const* u, uint32_t nblocks,
                                   const uint64_t* offsets, const unsigned* sz3,
                                   const unsigned* bflags, const double* pmin3, const double* pmax3,
                                   const orc_params* p, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (long b = 0; b < (long)nblocks; ++b)
    orc_bssnrhs(rhs, u, offsets[b], pmin3 + 3 * b, pmax3 + 3 * b, sz3 + 3 * b, bflags[b], p);
  (void)nthreads;
}

extern "C" void orc_point_eqs(long n, const double* in, const double* eta, const orc_params* p, double* out) {
  const double* D[DGR_POINT_NIN];
  for (int k = 0; k < DGR_POINT_NIN; ++k) D[k] = in + (size_t)k * n;
  for (long pp = 0; pp < n; ++pp) {
    double o[24];
    point_eqs(D, pp, eta[pp], p, o);
    for (int v = 0; v < 24; ++v) out[(size_t)v * n + pp] = o[v];
  }
}

extern "C" int orc_enforce_constraints(double* const* u, long n, double chi_floor) {
  int bad = 0;
  const double one_third = 1.0 / 3.0;
  for (long node = 0; node < n; ++node) {
    double g[3][3], A[3][3];
    g[0][0] = u[SG0][node]; g[0][1] = u[SG1][node]; g[0][2] = u[SG2][node];
    g[1][1] = u[SG3][node]; g[1][2] = u[SG4][node]; g[2][2] = u[SG5][node];
    g[1][0] = g[0][1]; g[2][0] = g[0][2]; g[2][1] = g[1][2];
    A[0][0] = u[SA0][node]; A[0][1] = u[SA1][node]; A[0][2] = u[SA2][node];
    A[1][1] = u[SA3][node]; A[1][2] = u[SA4][node]; A[2][2] = u[SA5][node];
    A[1][0] = A[0][1]; A[2][0] = A[0][2]; A[2][1] = A[1][2];
    // det(gt), rescale to unit determinant (bssn_constraints.cpp:36-63)
    double det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[1][2]) - g[0][1] * g[0][1] * g[2][2] +
                 2.0 * g[0][1] * g[0][2] * g[1][2] - g[0][2] * g[0][2] * g[1][1];
    if (det < 0.0) { bad = 1; continue; }
    const double s = 1.0 / std::pow(det, one_third);
    for (int j = 0; j < 3; ++j)
      for (int i = 0; i < 3; ++i) g[i][j] *= s;
    det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[1][2]) - g[0][1] * g[0][1] * g[2][2] +
          2.0 * g[0][1] * g[0][2] * g[1][2] - g[0][2] * g[0][2] * g[1][1];
    if (std::fabs(det - 1.0) > 1.0e-6) bad = 1;
    // inverse metric and trace-free projection of At (:84-131)
    const double idet = 1.0 / det;
    double gu[3][3];
    gu[0][0] = idet * (g[1][1] * g[2][2] - g[1][2] * g[1][2]);
    gu[0][1] = idet * (-g[0][1] * g[2][2] + g[0][2] * g[1][2]);
    gu[0][2] = idet * (g[0][1] * g[1][2] - g[0][2] * g[1][1]);
    gu[1][1] = idet * (g[0][0] * g[2][2] - g[0][2] * g[0][2]);
    gu[1][2] = idet * (-g[0][0] * g[1][2] + g[0][1] * g[0][2]);
    gu[2][2] = idet * (g[0][0] * g[1][1] - g[0][1] * g[0][1]);
    const double t3 = one_third * (A[0][0] * gu[0][0] + A[1][1] * gu[1][1] + A[2][2] * gu[2][2] +
                                   2.0 * (A[0][1] * gu[0][1] + A[0][2] * gu[0][2] + A[1][2] * gu[1][2]));
    A[0][0] -= t3 * g[0][0]; A[0][1] -= t3 * g[0][1]; A[0][2] -= t3 * g[0][2];
    A[1][1] -= t3 * g[1][1]; A[1][2] -= t3 * g[1][2]; A[2][2] -= t3 * g[2][2];
    const double trA = A[0][0] * gu[0][0] + A[1][1] * gu[1][1] + A[2][2] * gu[2][2] +
                       2.0 * (A[0][1] * gu[0][1] + A[0][2] * gu[0][2] + A[1][2] * gu[1][2]);
    if (std::fabs(trA) > 1.0e-6) bad = 1;
    u[SA0][node] = A[0][0]; u[SA1][node] = A[0][1]; u[SA2][node] = A[0][2];
    u[SA3][node] = A[1][1]; u[SA4][node] = A[1][2]; u[SA5][node] = A[2][2];
    u[SG0][node] = g[0][0]; u[SG1][node] = g[0][1]; u[SG2][node] = g[0][2];
    u[SG3][node] = g[1][1]; u[SG4][node] = g[1][2]; u[SG5][node] = g[2][2];
    // floors (:150-171)
    if (u[CHI][node] < chi_floor) u[CHI][node] = chi_floor;
    u[ALPHA][node] = std::max(u[ALPHA][node], chi_floor);
  }
  return bad;
}

extern "C" void orc_rk4_stage(double* u_out, const double* u_prev, const double* k, double cdt, long n) {
  for (long i = 0; i < n; ++i) u_out[i] = u_prev[i] + cdt * k[i];
}

extern "C" void orc_rk4_final(double* u, const double* u_prev, const double* const* k4, double dt, long n) {
  static const double C[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  for (long i = 0; i < n; ++i) {
    double acc = u_prev[i];
    for (int s = 0; s < 4; ++s) acc += C[s] * dt * k4[s][i];
    u[i] = acc;
  }
}

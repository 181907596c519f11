/*
 * Finite-difference stencils near flagged faces as coefficient tables.
 *
 * Every operator at block coordinate q along an axis reads only points of
 * the window q-4 .. q+4 (deriv42 and ko_deriv42 stay within q-3 .. q+3; the
 * 4th-order one-sided rows of deriv64 reach one further), so each operator at
 * each position is a row of nine coefficients over that window plus a final
 * scale:
 *
 *   deriv42    deriv42_x/xx, deriv42adv_x, ko_deriv42_x
 *              (BSSN_GR/src/derivs.cpp:11-88, 244-328, 494-634, 914-1079)
 *   deriv64    deriv64_x/xx (QuadGrav/src/derivs.cpp:2141-2280, 2569-2680);
 *              deriv64adv_x is deriv64_x (:2913-2928); KO stays ko_deriv42
 *   closures   the i == 3, 4, 5 and ie-1, ie-2, ie-3 branches of the same
 *              functions, taken when the face's bflag bit is set.
 *
 * A row is summed in the reference's term order with FP contraction off:
 * ascending window position for all forms except the ko_deriv42 low-face
 * closures, which the reference writes from the highest point down
 * (derivs.cpp:1002-1020).  Window points with a zero coefficient are skipped
 * by a select, so their values (possibly unset padding) never enter, and the
 * sum starts from +0, so each result equals the reference expression
 * (a leading +0 and skipped terms change at most the sign of a zero).  The
 * final step multiplies by the axis constant the reference multiplies by, or
 * divides by smr1/2/3 where it divides.
 *
 * Used by the boundary-tile kernel in place of per-case branches: one code
 * path per operator instead of up to seven divergent ones.
 */
#ifndef DGR_STENCIL_FORMS_H
#define DGR_STENCIL_FORMS_H

#undef DGR_HD  // (re)defined for what follows; bssn_staged.h defines its own
#ifdef __HIPCC__
#define DGR_HD __host__ __device__
#else
#define DGR_HD
#endif

namespace dgr_forms {

// Scale selectors: the first ten doubles of dgr_internal::AxisC in order.
enum Scale { S_IH2 = 0, S_IH12, S_ISQ, S_ISQ12, S_PRE, S_SMR1, S_SMR2, S_SMR3, S_IH60, S_ISQ180 };

constexpr int W = 9;  // window q-4 .. q+4

struct Form {
  double c[W];  // coefficients of points q-4 .. q+4
  int sel;      // Scale; S_SMR1..3 divide, the others multiply
  int desc;     // 1: sum from q+4 down to q-4
};

enum FormId {
  // deriv42_x (derivs.cpp:11-88)
  F_D1 = 0, F_D1_L3, F_D1_L4, F_D1_H2, F_D1_H1,
  // deriv42_xx (derivs.cpp:244-328)
  F_D2, F_D2_L3, F_D2_L4, F_D2_H2, F_D2_H1,
  // deriv42adv_x (derivs.cpp:494-634): P = beta > 0, N = otherwise
  F_ADV_P, F_ADV_N, F_ADV_L3, F_ADV_L4P, F_ADV_L4N, F_ADV_L5P, F_ADV_L5N,
  F_ADV_H1, F_ADV_H2P, F_ADV_H2N, F_ADV_H3P, F_ADV_H3N,
  // ko_deriv42_x (derivs.cpp:914-1079)
  F_KO, F_KO_L3, F_KO_L4, F_KO_L5, F_KO_H3, F_KO_H2, F_KO_H1,
  F_NFORMS
};

// Rows of the 6th-order option: deriv64 first (also the advective
// operator) and second derivatives; the KO rows are those of the deriv42
// table at the same ids.
enum FormId64 {
  G_D1 = 0, G_D1_L3, G_D1_L4, G_D1_L5, G_D1_H3, G_D1_H2, G_D1_H1,
  G_D2, G_D2_L3, G_D2_L4, G_D2_L5, G_D2_H3, G_D2_H2, G_D2_H1,
  G_KO, G_KO_L3, G_KO_L4, G_KO_L5, G_KO_H3, G_KO_H2, G_KO_H1,
  G_NFORMS
};

// Window index w: 0 = q-3, 3 = q, 6 = q+3.
// Window index: 0 = q-4, 4 = q, 8 = q+4.
#define DGR_FORMS_TABLE                                                                   \
  {                                                                                       \
    /* F_D1    (u[-2] - 8u[-1] + 8u[+1] - u[+2]) ih12 */                                  \
    {{0, 0, 1, -8, 0, 8, -1, 0, 0}, ::dgr_forms::S_IH12, 0},                              \
    /* F_D1_L3 (-3u3 + 4u4 - u5) ih2, q = 3 */                                            \
    {{0, 0, 0, 0, -3, 4, -1, 0, 0}, ::dgr_forms::S_IH2, 0},                               \
    /* F_D1_L4 (-u3 + u5) ih2, q = 4 */                                                   \
    {{0, 0, 0, -1, 0, 1, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_D1_H2 (-u[e-3] + u[e-1]) ih2, q = e-2 */                                         \
    {{0, 0, 0, -1, 0, 1, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_D1_H1 (u[e-3] - 4u[e-2] + 3u[e-1]) ih2, q = e-1 */                               \
    {{0, 0, 1, -4, 3, 0, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_D2    (-u[-2] + 16u[-1] - 30u0 + 16u[+1] - u[+2]) isq12 */                       \
    {{0, 0, -1, 16, -30, 16, -1, 0, 0}, ::dgr_forms::S_ISQ12, 0},                         \
    /* F_D2_L3 (2u3 - 5u4 + 4u5 - u6) isq */                                              \
    {{0, 0, 0, 0, 2, -5, 4, -1, 0}, ::dgr_forms::S_ISQ, 0},                               \
    /* F_D2_L4 (u3 - 2u4 + u5) isq */                                                     \
    {{0, 0, 0, 1, -2, 1, 0, 0, 0}, ::dgr_forms::S_ISQ, 0},                                \
    /* F_D2_H2 (u[e-3] - 2u[e-2] + u[e-1]) isq */                                         \
    {{0, 0, 0, 1, -2, 1, 0, 0, 0}, ::dgr_forms::S_ISQ, 0},                                \
    /* F_D2_H1 (-u[e-4] + 4u[e-3] - 5u[e-2] + 2u[e-1]) isq */                             \
    {{0, -1, 4, -5, 2, 0, 0, 0, 0}, ::dgr_forms::S_ISQ, 0},                               \
    /* F_ADV_P (-3u[-1] - 10u0 + 18u[+1] - 6u[+2] + u[+3]) ih12 */                        \
    {{0, 0, 0, -3, -10, 18, -6, 1, 0}, ::dgr_forms::S_IH12, 0},                           \
    /* F_ADV_N (-u[-3] + 6u[-2] - 18u[-1] + 10u0 + 3u[+1]) ih12 */                        \
    {{0, -1, 6, -18, 10, 3, 0, 0, 0}, ::dgr_forms::S_IH12, 0},                            \
    /* F_ADV_L3 (-3u3 + 4u4 - u5) ih2 */                                                  \
    {{0, 0, 0, 0, -3, 4, -1, 0, 0}, ::dgr_forms::S_IH2, 0},                               \
    /* F_ADV_L4P (-3u4 + 4u5 - u6) ih2 */                                                 \
    {{0, 0, 0, 0, -3, 4, -1, 0, 0}, ::dgr_forms::S_IH2, 0},                               \
    /* F_ADV_L4N (-u3 + u5) ih2 */                                                        \
    {{0, 0, 0, -1, 0, 1, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_ADV_L5P (-3u4 - 10u5 + 18u6 - 6u7 + u8) ih12 */                                  \
    {{0, 0, 0, -3, -10, 18, -6, 1, 0}, ::dgr_forms::S_IH12, 0},                           \
    /* F_ADV_L5N (u3 - 4u4 + 3u5) ih2 */                                                  \
    {{0, 0, 1, -4, 3, 0, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_ADV_H1 (u[e-3] - 4u[e-2] + 3u[e-1]) ih2 */                                       \
    {{0, 0, 1, -4, 3, 0, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_ADV_H2P (-u[e-3] + u[e-1]) ih2 */                                                \
    {{0, 0, 0, -1, 0, 1, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_ADV_H2N (u[e-4] - 4u[e-3] + 3u[e-2]) ih2 */                                      \
    {{0, 0, 1, -4, 3, 0, 0, 0, 0}, ::dgr_forms::S_IH2, 0},                                \
    /* F_ADV_H3P (-3u[e-3] + 4u[e-2] - u[e-1]) ih2 */                                     \
    {{0, 0, 0, 0, -3, 4, -1, 0, 0}, ::dgr_forms::S_IH2, 0},                               \
    /* F_ADV_H3N (-u[e-6] + 6u[e-5] - 18u[e-4] + 10u[e-3] + 3u[e-2]) ih12 */              \
    {{0, -1, 6, -18, 10, 3, 0, 0, 0}, ::dgr_forms::S_IH12, 0},                            \
    /* F_KO    pre (-u[-3] + 6u[-2] - 15u[-1] + 20u0 - 15u[+1] + 6u[+2] - u[+3]) */       \
    {{0, -1, 6, -15, 20, -15, 6, -1, 0}, ::dgr_forms::S_PRE, 0},                          \
    /* F_KO_L3 (u6 - 3u5 + 3u4 - u3) / smr3, summed downwards */                          \
    {{0, 0, 0, 0, -1, 3, -3, 1, 0}, ::dgr_forms::S_SMR3, 1},                              \
    /* F_KO_L4 (u7 - 6u6 + 12u5 - 10u4 + 3u3) / smr2, downwards */                        \
    {{0, 0, 0, 3, -10, 12, -6, 1, 0}, ::dgr_forms::S_SMR2, 1},                            \
    /* F_KO_L5 (u8 - 6u7 + 15u6 - 19u5 + 12u4 - 3u3) / smr1, downwards */                 \
    {{0, 0, -3, 12, -19, 15, -6, 1, 0}, ::dgr_forms::S_SMR1, 1},                          \
    /* F_KO_H3 (u[e-6] - 6u[e-5] + 15u[e-4] - 19u[e-3] + 12u[e-2] - 3u[e-1]) / smr1 */    \
    {{0, 1, -6, 15, -19, 12, -3, 0, 0}, ::dgr_forms::S_SMR1, 0},                          \
    /* F_KO_H2 (u[e-5] - 6u[e-4] + 12u[e-3] - 10u[e-2] + 3u[e-1]) / smr2 */               \
    {{0, 1, -6, 12, -10, 3, 0, 0, 0}, ::dgr_forms::S_SMR2, 0},                            \
    /* F_KO_H1 (u[e-4] - 3u[e-3] + 3u[e-2] - u[e-1]) / smr3 */                            \
    {{0, 1, -3, 3, -1, 0, 0, 0, 0}, ::dgr_forms::S_SMR3, 0},                              \
  }

#define DGR_FORMS64_TABLE                                                                 \
  {                                                                                       \
    /* G_D1    (-u[-3] + 9u[-2] - 45u[-1] + 45u[+1] - 9u[+2] + u[+3]) ih60 */             \
    {{0, -1, 9, -45, 0, 45, -9, 1, 0}, ::dgr_forms::S_IH60, 0},                           \
    /* G_D1_L3 (-25u3 + 48u4 - 36u5 + 16u6 - 3u7) ih12 */                                 \
    {{0, 0, 0, 0, -25, 48, -36, 16, -3}, ::dgr_forms::S_IH12, 0},                         \
    /* G_D1_L4 (-3u3 - 10u4 + 18u5 - 6u6 + u7) ih12 */                                    \
    {{0, 0, 0, -3, -10, 18, -6, 1, 0}, ::dgr_forms::S_IH12, 0},                           \
    /* G_D1_L5 (u3 - 8u4 + 8u6 - u7) ih12 */                                              \
    {{0, 0, 1, -8, 0, 8, -1, 0, 0}, ::dgr_forms::S_IH12, 0},                              \
    /* G_D1_H3 (u[e-5] - 8u[e-4] + 8u[e-2] - u[e-1]) ih12, q = e-3 */                     \
    {{0, 0, 1, -8, 0, 8, -1, 0, 0}, ::dgr_forms::S_IH12, 0},                              \
    /* G_D1_H2 (-u[e-5] + 6u[e-4] - 18u[e-3] + 10u[e-2] + 3u[e-1]) ih12, q = e-2 */       \
    {{0, -1, 6, -18, 10, 3, 0, 0, 0}, ::dgr_forms::S_IH12, 0},                            \
    /* G_D1_H1 (3u[e-5] - 16u[e-4] + 36u[e-3] - 48u[e-2] + 25u[e-1]) ih12, q = e-1 */     \
    {{3, -16, 36, -48, 25, 0, 0, 0, 0}, ::dgr_forms::S_IH12, 0},                          \
    /* G_D2    (2u[-3] - 27u[-2] + 270u[-1] - 490u0 + 270u[+1] - 27u[+2] + 2u[+3]) isq180 */ \
    {{0, 2, -27, 270, -490, 270, -27, 2, 0}, ::dgr_forms::S_ISQ180, 0},                   \
    /* G_D2_L3 (35u3 - 104u4 + 114u5 - 56u6 + 11u7) isq12 */                              \
    {{0, 0, 0, 0, 35, -104, 114, -56, 11}, ::dgr_forms::S_ISQ12, 0},                      \
    /* G_D2_L4 (11u3 - 20u4 + 6u5 + 4u6 - u7) isq12 */                                    \
    {{0, 0, 0, 11, -20, 6, 4, -1, 0}, ::dgr_forms::S_ISQ12, 0},                           \
    /* G_D2_L5 (-u3 + 16u4 - 30u5 + 16u6 - u7) isq12 */                                   \
    {{0, 0, -1, 16, -30, 16, -1, 0, 0}, ::dgr_forms::S_ISQ12, 0},                         \
    /* G_D2_H3 (-u[e-5] + 16u[e-4] - 30u[e-3] + 16u[e-2] - u[e-1]) isq12 */               \
    {{0, 0, -1, 16, -30, 16, -1, 0, 0}, ::dgr_forms::S_ISQ12, 0},                         \
    /* G_D2_H2 (-u[e-5] + 4u[e-4] + 6u[e-3] - 20u[e-2] + 11u[e-1]) isq12 */               \
    {{0, -1, 4, 6, -20, 11, 0, 0, 0}, ::dgr_forms::S_ISQ12, 0},                           \
    /* G_D2_H1 (11u[e-5] - 56u[e-4] + 114u[e-3] - 104u[e-2] + 35u[e-1]) isq12 */          \
    {{11, -56, 114, -104, 35, 0, 0, 0, 0}, ::dgr_forms::S_ISQ12, 0},                      \
    /* G_KO    pre (-u[-3] + 6u[-2] - 15u[-1] + 20u0 - 15u[+1] + 6u[+2] - u[+3]) */       \
    {{0, -1, 6, -15, 20, -15, 6, -1, 0}, ::dgr_forms::S_PRE, 0},                          \
    /* G_KO_L3 (u6 - 3u5 + 3u4 - u3) / smr3, summed downwards */                          \
    {{0, 0, 0, 0, -1, 3, -3, 1, 0}, ::dgr_forms::S_SMR3, 1},                              \
    /* G_KO_L4 (u7 - 6u6 + 12u5 - 10u4 + 3u3) / smr2, downwards */                        \
    {{0, 0, 0, 3, -10, 12, -6, 1, 0}, ::dgr_forms::S_SMR2, 1},                            \
    /* G_KO_L5 (u8 - 6u7 + 15u6 - 19u5 + 12u4 - 3u3) / smr1, downwards */                 \
    {{0, 0, -3, 12, -19, 15, -6, 1, 0}, ::dgr_forms::S_SMR1, 1},                          \
    /* G_KO_H3 (u[e-6] - 6u[e-5] + 15u[e-4] - 19u[e-3] + 12u[e-2] - 3u[e-1]) / smr1 */    \
    {{0, 1, -6, 15, -19, 12, -3, 0, 0}, ::dgr_forms::S_SMR1, 0},                          \
    /* G_KO_H2 (u[e-5] - 6u[e-4] + 12u[e-3] - 10u[e-2] + 3u[e-1]) / smr2 */               \
    {{0, 1, -6, 12, -10, 3, 0, 0, 0}, ::dgr_forms::S_SMR2, 0},                            \
    /* G_KO_H1 (u[e-4] - 3u[e-3] + 3u[e-2] - u[e-1]) / smr3 */                            \
    {{0, 1, -3, 3, -1, 0, 0, 0, 0}, ::dgr_forms::S_SMR3, 0},                              \
  }

// Form of each operator at position q of an axis ending at e = sz-3 (the
// first padding point), with the low / high face flags lf, hf.  The order of
// the tests follows the reference's branches.
DGR_HD inline int form_d1(int q, int e, bool lf, bool hf) {
  if (hf && q == e - 1) return F_D1_H1;
  if (hf && q == e - 2) return F_D1_H2;
  if (lf && q == 4) return F_D1_L4;
  if (lf && q == 3) return F_D1_L3;
  return F_D1;
}
DGR_HD inline int form_d2(int q, int e, bool lf, bool hf) {
  if (hf && q == e - 1) return F_D2_H1;
  if (hf && q == e - 2) return F_D2_H2;
  if (lf && q == 4) return F_D2_L4;
  if (lf && q == 3) return F_D2_L3;
  return F_D2;
}
DGR_HD inline int form_adv(int q, int e, bool lf, bool hf, bool pos) {
  if (hf && q == e - 1) return F_ADV_H1;
  if (hf && q == e - 2) return pos ? F_ADV_H2P : F_ADV_H2N;
  if (hf && q == e - 3) return pos ? F_ADV_H3P : F_ADV_H3N;
  if (lf && q == 3) return F_ADV_L3;
  if (lf && q == 4) return pos ? F_ADV_L4P : F_ADV_L4N;
  if (lf && q == 5) return pos ? F_ADV_L5P : F_ADV_L5N;
  return pos ? F_ADV_P : F_ADV_N;
}
DGR_HD inline int form_ko(int q, int e, bool lf, bool hf) {
  if (hf && q == e - 1) return F_KO_H1;
  if (hf && q == e - 2) return F_KO_H2;
  if (hf && q == e - 3) return F_KO_H3;
  if (lf && q == 3) return F_KO_L3;
  if (lf && q == 4) return F_KO_L4;
  if (lf && q == 5) return F_KO_L5;
  return F_KO;
}

// deriv64 rows (QuadGrav/src/derivs.cpp: high-face rows tested first).
DGR_HD inline int form64_d1(int q, int e, bool lf, bool hf) {
  if (hf && q == e - 1) return G_D1_H1;
  if (hf && q == e - 2) return G_D1_H2;
  if (hf && q == e - 3) return G_D1_H3;
  if (lf && q == 3) return G_D1_L3;
  if (lf && q == 4) return G_D1_L4;
  if (lf && q == 5) return G_D1_L5;
  return G_D1;
}
DGR_HD inline int form64_d2(int q, int e, bool lf, bool hf) {
  if (hf && q == e - 1) return G_D2_H1;
  if (hf && q == e - 2) return G_D2_H2;
  if (hf && q == e - 3) return G_D2_H3;
  if (lf && q == 3) return G_D2_L3;
  if (lf && q == 4) return G_D2_L4;
  if (lf && q == 5) return G_D2_L5;
  return G_D2;
}
DGR_HD inline int form64_ko(int q, int e, bool lf, bool hf) { return form_ko(q, e, lf, hf) - F_KO + G_KO; }

// Sum of one form over the window w[0..8] (points q-4 .. q+4) in ascending
// order; the caller takes the descending sum for desc forms.
DGR_HD inline double form_sum_up(const double* c, const double (&w)[W]) {
#pragma clang fp contract(off)
  double t = 0.0;
  for (int m = 0; m < W; ++m) t = t + (c[m] != 0.0 ? c[m] * w[m] : 0.0);
  return t;
}
DGR_HD inline double form_sum_down(const double* c, const double (&w)[W]) {
#pragma clang fp contract(off)
  double t = 0.0;
  for (int m = W - 1; m >= 0; --m) t = t + (c[m] != 0.0 ? c[m] * w[m] : 0.0);
  return t;
}
// Final scale: scales[] are ih2, ih12, isq, isq12, pre, smr1, smr2, smr3,
// ih60, isq180.
DGR_HD inline double form_scale(double sum, int sel, const double* scales) {
#pragma clang fp contract(off)
  const double s = scales[sel];
  return (sel >= S_SMR1 && sel <= S_SMR3) ? sum / s : sum * s;
}

// The deriv42 rows of DGR_FORMS_TABLE written out over the 7-point window
// m3 .. p3 (points q-3 .. q+3; every deriv42 row lies inside it), term by
// term in the row's summation order (descending for the low-face KO rows),
// then scaled as form_scale does.  The zero terms of a row are left out:
// the running sum starts at +0.0, so adding +0.0 never changes it (except
// for the sign of an all-zero result).  One function per operator family;
// f outside the family's closure rows gives the family's interior row.
DGR_HD inline double closure42_d1(int f, double m2, double m1, double z0, double p1, double p2,
                                  const double* sc) {
#pragma clang fp contract(off)
  switch (f) {
    case F_D1_L3: return ((-3.0 * z0 + 4.0 * p1) - p2) * sc[S_IH2];
    case F_D1_L4:
    case F_D1_H2: return (-m1 + p1) * sc[S_IH2];
    case F_D1_H1: return ((m2 - 4.0 * m1) + 3.0 * z0) * sc[S_IH2];
    default: return (((m2 - 8.0 * m1) + 8.0 * p1) - p2) * sc[S_IH12];
  }
}
DGR_HD inline double closure42_d2(int f, double m3, double m2, double m1, double z0, double p1, double p2,
                                  double p3, const double* sc) {
#pragma clang fp contract(off)
  switch (f) {
    case F_D2_L3: return (((2.0 * z0 - 5.0 * p1) + 4.0 * p2) - p3) * sc[S_ISQ];
    case F_D2_L4:
    case F_D2_H2: return ((m1 - 2.0 * z0) + p1) * sc[S_ISQ];
    case F_D2_H1: return (((-m3 + 4.0 * m2) - 5.0 * m1) + 2.0 * z0) * sc[S_ISQ];
    default: return ((((-m2 + 16.0 * m1) - 30.0 * z0) + 16.0 * p1) - p2) * sc[S_ISQ12];
  }
}
DGR_HD inline double closure42_adv(int f, double m3, double m2, double m1, double z0, double p1, double p2,
                                   double p3, const double* sc) {
#pragma clang fp contract(off)
  switch (f) {
    case F_ADV_L3:
    case F_ADV_L4P:
    case F_ADV_H3P: return ((-3.0 * z0 + 4.0 * p1) - p2) * sc[S_IH2];
    case F_ADV_L4N:
    case F_ADV_H2P: return (-m1 + p1) * sc[S_IH2];
    case F_ADV_L5N:
    case F_ADV_H1:
    case F_ADV_H2N: return ((m2 - 4.0 * m1) + 3.0 * z0) * sc[S_IH2];
    case F_ADV_N:
    case F_ADV_H3N: return ((((-m3 + 6.0 * m2) - 18.0 * m1) + 10.0 * z0) + 3.0 * p1) * sc[S_IH12];
    default: return ((((-3.0 * m1 - 10.0 * z0) + 18.0 * p1) - 6.0 * p2) + p3) * sc[S_IH12];  // P, L5P
  }
}
DGR_HD inline double closure42_ko(int f, double m3, double m2, double m1, double z0, double p1, double p2,
                                  double p3, const double* sc) {
#pragma clang fp contract(off)
  switch (f) {
    case F_KO_L3: return (((p3 - 3.0 * p2) + 3.0 * p1) - z0) / sc[S_SMR3];
    case F_KO_L4: return ((((p3 - 6.0 * p2) + 12.0 * p1) - 10.0 * z0) + 3.0 * m1) / sc[S_SMR2];
    case F_KO_L5: return (((((p3 - 6.0 * p2) + 15.0 * p1) - 19.0 * z0) + 12.0 * m1) - 3.0 * m2) / sc[S_SMR1];
    case F_KO_H3: return (((((m3 - 6.0 * m2) + 15.0 * m1) - 19.0 * z0) + 12.0 * p1) - 3.0 * p2) / sc[S_SMR1];
    case F_KO_H2: return ((((m3 - 6.0 * m2) + 12.0 * m1) - 10.0 * z0) + 3.0 * p1) / sc[S_SMR2];
    case F_KO_H1: return (((m3 - 3.0 * m2) + 3.0 * m1) - z0) / sc[S_SMR3];
    default:
      return ((((((-m3 + 6.0 * m2) - 15.0 * m1) + 20.0 * z0) - 15.0 * p1) + 6.0 * p2) - p3) * sc[S_PRE];
  }
}

}  // namespace dgr_forms

#endif

// Host build of the boundary-tile stencil tables (stencil_forms.h) over one
// padded block, for comparison with the oracle's operators in
// tests/test_stencil_forms.py.  Writes interior points along `axis` only
// and leaves every other point untouched.  kind: 0..3 deriv42 first,
// second, advective, KO; 4..6 deriv64 first, second, advective (= first);
// 7..10 the deriv42 operators of kinds 0..3 through the written-out rows
// (closure42_*) that the boundary kernels evaluate.
#include <cstdint>

#include "../../dendro-5.01_amd/csrc/stencil_forms.h"

using namespace dgr_forms;

static const Form kForms[F_NFORMS] = DGR_FORMS_TABLE;
static const Form kForms64[G_NFORMS] = DGR_FORMS64_TABLE;

extern "C" void forms_deriv(int kind, int axis, double* Du, const double* u, double h, const unsigned* sz,
                            unsigned bflag, const double* beta) {
  const int n0 = (int)sz[0], n1 = (int)sz[1], n2 = (int)sz[2];
  const int n[3] = {n0, n1, n2};
  const long st[3] = {1, (long)n0, (long)n0 * n1};
  const double ih = 1.0 / h, isq = 1.0 / (h * h);
  const double scales[10] = {0.5 * ih, ih / 12.0, isq, isq / 12.0, -1.0 / 64.0 / h,
                             49.0 / 48.0 * 64 * h, 43.0 / 48.0 * 64 * h, 59.0 / 48.0 * 64 * h, ih / 60.0, isq / 180.0};
  const bool lf = (bflag >> (2 * axis)) & 1u, hf = (bflag >> (2 * axis + 1)) & 1u;
  const int e = n[axis] - 3;
  for (int k = 3; k < n2 - 3; ++k)
    for (int j = 3; j < n1 - 3; ++j)
      for (int i = 3; i < n0 - 3; ++i) {
        const int idx[3] = {i, j, k};
        const int q = idx[axis];
        const long c = i + (long)n0 * (j + (long)n1 * k);
        double w[W];
        for (int m = 0; m < W; ++m) {
          const int p = q + m - 4;
          w[m] = (p >= 0 && p < n[axis]) ? u[c + (m - 4) * st[axis]] : 0.0;
        }
        if (kind >= 7) {
          const bool pos = beta[c] > 0.0;
          double r;
          if (kind == 7) r = closure42_d1(form_d1(q, e, lf, hf), w[2], w[3], w[4], w[5], w[6], scales);
          else if (kind == 8) r = closure42_d2(form_d2(q, e, lf, hf), w[1], w[2], w[3], w[4], w[5], w[6], w[7], scales);
          else if (kind == 9)
            r = closure42_adv(form_adv(q, e, lf, hf, pos), w[1], w[2], w[3], w[4], w[5], w[6], w[7], scales);
          else r = closure42_ko(form_ko(q, e, lf, hf), w[1], w[2], w[3], w[4], w[5], w[6], w[7], scales);
          Du[c] = r;
          continue;
        }
        const Form* F;
        if (kind == 0) F = &kForms[form_d1(q, e, lf, hf)];
        else if (kind == 1) F = &kForms[form_d2(q, e, lf, hf)];
        else if (kind == 2) F = &kForms[form_adv(q, e, lf, hf, beta[c] > 0.0)];
        else if (kind == 3) F = &kForms[form_ko(q, e, lf, hf)];
        else if (kind == 4 || kind == 6) F = &kForms64[form64_d1(q, e, lf, hf)];
        else F = &kForms64[form64_d2(q, e, lf, hf)];
        const double s = F->desc ? form_sum_down(F->c, w) : form_sum_up(F->c, w);
        Du[c] = form_scale(s, F->sel, scales);
      }
}

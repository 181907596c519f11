/*
 * Reference harness -- test infrastructure only.
 *
 * Compiles the reference's own generated BSSN point equations,
 * BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp, where they lie under
 * the reference tree (REF/BSSN_GR/src, on the include path), the same way
 * BSSN_GR/src/rhs.cpp:243-244 uses them: textually included inside the
 * point loop, with every input named <var>[pp] and every output <rhs>[pp].
 * The harness only supplies those arrays and the per-point eta; it copies
 * nothing from the reference.  The built library lands in oracle/_ref/ and
 * is used to write the golden vectors under tests/golden/ and to check the
 * oracle's point equations in this container.
 *
 * Only the point equations are built: BSSN_GR/src/rhs.cpp and derivs.cpp
 * include TreeNode.h -> mpi.h (include/TreeNode.h:23), which this image
 * does not have, so those files are not buildable here.
 */
#include <cmath>
#include <cstddef>
#include "bssn_point_io.h"

extern "C" void ref_point_eqs(long n, const double* in, const double* eta_pt,
                              const double* lam, const double* lamf, double* out) {
#define REF_IN(name, k) const double* name = in + (size_t)(k) * (size_t)n;
  DGR_POINT_INPUTS(REF_IN)
#undef REF_IN
#define REF_OUT(name, k) double* name = out + (size_t)(k) * (size_t)n;
  DGR_POINT_OUTPUTS(REF_OUT)
#undef REF_OUT
  // rhs.cpp:152-155 keeps lambda as unsigned int and lambda_f as double
  const unsigned int lambda[4] = {(unsigned)lam[0], (unsigned)lam[1], (unsigned)lam[2], (unsigned)lam[3]};
  const double lambda_f[2] = {lamf[0], lamf[1]};
  for (long pp = 0; pp < n; ++pp) {
    const double eta = eta_pt[pp];
#include "bssneqs_eta_const_standard_gauge.cpp"
  }
}

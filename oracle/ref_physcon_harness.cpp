/*
 * Reference harness -- test infrastructure only.
 *
 * Compiles the reference's own generated constraint equations,
 * BSSN_GR/src/physconeqs.cpp, where they lie under the reference tree
 * (REF/BSSN_GR/src on the include path), the way BSSN_GR/src/physcon.cpp:
 * 66-76 uses them: textually included inside the point loop with the point
 * coordinates x, y, z in scope, every input named <var>[pp] and every output
 * <constraint>[pp].  The harness only supplies those arrays; it copies
 * nothing from the reference.  Built into oracle/_ref/ by `make ref`; used to
 * write tests/golden/physcon_*.npz and to check the oracle's restatement.
 */
#include <cmath>
#include <cstddef>

#include "physcon_point_io.h"

extern "C" void ref_physcon_point(long n, const double* in, double* out) {
#define REF_IN(name, k) const double* name = in + (size_t)(k) * (size_t)n;
  DGR_PHYSCON_INPUTS(REF_IN)
#undef REF_IN
#define REF_OUT(name, k) double* name = out + (size_t)(k) * (size_t)n;
  DGR_PHYSCON_OUTPUTS(REF_OUT)
#undef REF_OUT
  for (long pp = 0; pp < n; ++pp) {
    const double x = px[pp], y = py[pp], z = pz[pp];
#include "physconeqs.cpp"
  }
}

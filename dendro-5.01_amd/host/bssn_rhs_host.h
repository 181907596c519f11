/*
 * Host-side mirror of the reference's BSSN RHS interface, on top of the
 * C ABI in include/dgr_bssn.h.
 *
 *   bssnRHS(rhs, u, blkList, numBlocks)      BSSN_GR/include/rhs.h:45
 *   bssnrhs(rhs, u, offset, ptmin, ptmax, sz, bflag)   BSSN_GR/include/rhs.h:47-50
 *   physical_constraints(con, u, offset, pmin, pmax, sz, bflag)
 *                                            BSSN_GR/include/physcon.h:17-20
 *
 * Same names, argument meaning and data layout: 24 unzipped arrays in VAR
 * order, blocks at `offset`, only interior points of the RHS written.  The
 * arrays may be host or device memory: device arrays run in place; host
 * arrays are staged through device buffers (u and rhs copied in, rhs copied
 * back, the PCIe round trip the reference GPU path makes on every call,
 * BSSN_GR/cuda_gr/src/rhs_cuda.cu:164,187).
 *
 * Parameters come from dgr::bssn_params(), whose field names and defaults
 * are those of the reference's bssn:: globals (BSSN_GR/src/parameters.cpp:
 * 87-97).  The reference functions return void and have no error path;
 * here a failure throws std::runtime_error with the library's message.
 */
#ifndef DGR_BSSN_RHS_HOST_H
#define DGR_BSSN_RHS_HOST_H

#include "../../include/dgr_bssn.h"

namespace dgr {

struct BSSNGlobals {
  unsigned int BSSN_LAMBDA[4] = {1, 1, 1, 1};
  double BSSN_LAMBDA_F[2] = {1.0, 0.0};
  double ETA_CONST = 2.0;
  double ETA_R0 = 50.0;
  double ETA_DAMPING_EXP = 1.0;
  double KO_DISS_SIGMA = 0.01;
  double CHI_FLOOR = 0.1;
};

// Process-wide parameter set used by bssnRHS / bssnrhs.
BSSNGlobals& bssn_params();

// Process-wide TwoPunctures parameters (the reference's TPID:: globals,
// BSSN_GR/include/parameters.h:262-310), starting from its defaults; set
// them before the first TwoPunctures call, which solves.
dgr_tp_params& tp_params();

// What bssnRHS reads from an ot::Block (BSSN_GR/src/rhs.cpp:45-62): the
// offset, the allocation sizes, the boundary flag and the padded extent
// (corner -/+ 3 dx).
struct BlockDesc {
  unsigned int offset;
  unsigned int sz[3];
  unsigned int bflag;
  double ptmin[3];
  double ptmax[3];
};

}  // namespace dgr

void bssnRHS(double** uzipVarsRHS, const double** uZipVars, const dgr::BlockDesc* blkList,
             unsigned int numBlocks);

void bssnrhs(double** uzipVarsRHS, const double** uZipVars, const unsigned int& offset, const double* ptmin,
             const double* ptmax, const unsigned int* sz, const unsigned int& bflag);

// con: 6 arrays (ham, mom0..2, psi4_real, psi4_img; VAR_CONSTRAINT order,
// BSSN_GR/include/grDef.h:34) of the same unzipped layout as u.
void physical_constraints(double** uZipConVars, const double** uZipVars, const unsigned int& offset,
                          const double* pmin, const double* pmax, const unsigned int* sz, const unsigned int& bflag);

// TwoPunctures (BSSN_GR/include/TwoPunctures.h, BSSN_GR/src/TwoPunctures.cpp:
// 24-498): the first call solves with dgr::tp_params(), every call writes the
// 24 variables of the point (x, y, z) into vars (VAR order) and the bare
// masses, puncture ADM masses, total ADM mass and angular momentum.  The
// coordinates are physical: the reference's GRIDX_TO_X is the caller's.  One
// point per call on the host, as the reference; dgr_tp_evaluate fills whole
// grids on the device.
void TwoPunctures(const double xx1, const double yy1, const double zz1, double* vars, double* mp, double* mm,
                  double* mp_adm, double* mm_adm, double* E, double* J1, double* J2, double* J3);

#endif

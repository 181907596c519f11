// Host-side mirror of bssnRHS / bssnrhs (see bssn_rhs_host.h).
#include "bssn_rhs_host.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgr_bssn.h"

namespace dgr {

BSSNGlobals& bssn_params() {
  static BSSNGlobals g;
  return g;
}

dgr_tp_params& tp_params() {
  static dgr_tp_params p = [] {
    dgr_tp_params d;
    dgr_tp_params_default(&d);
    return d;
  }();
  return p;
}

}  // namespace dgr

namespace {

dgr_bssn_params to_params(const dgr::BSSNGlobals& g) {
  dgr_bssn_params p{};
  for (int i = 0; i < 4; ++i) p.lambda[i] = g.BSSN_LAMBDA[i];
  p.lambda_f[0] = g.BSSN_LAMBDA_F[0];
  p.lambda_f[1] = g.BSSN_LAMBDA_F[1];
  p.eta_const = g.ETA_CONST;
  p.eta_r0 = g.ETA_R0;
  p.eta_damping_exp = g.ETA_DAMPING_EXP;
  p.ko_sigma = g.KO_DISS_SIGMA;
  p.chi_floor = g.CHI_FLOOR;
  p.deriv_order = 42;
  return p;
}

void check(int rc, const char* what) {
  if (rc != DGR_OK) throw std::runtime_error(std::string(what) + ": " + dgr_error_string(rc));
}

void hipcheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

bool is_device(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice;
}

// The last block list's plan and the host-path staging buffers, kept
// between calls: rkBSSN calls bssnRHS four times per step on one block list
// (BSSN_GR/src/rkBSSN.cpp:647-783), so only the first call of a list pays
// for the plan (block and tile tables, neighbour search, a stream) and the
// device buffers.  A different block list, parameter set or device
// rebuilds the plan; staging buffers only grow.  Calls are serialised by a
// mutex (the reference's bssnrhs is called from OpenMP threads,
// BSSN_GR/src/rhsTestBed.cpp:181).
struct ShimCache {
  std::mutex m;
  std::vector<dgr_block> blocks;
  dgr_bssn_params p{};
  int device = -1;
  dgr_plan_t plan = nullptr;
  double* du = nullptr;
  double* dr = nullptr;
  size_t cap = 0;  // doubles per variable in du / dr
  ~ShimCache() {
    // the HIP runtime may already be gone at process exit: no calls here
  }
};

ShimCache& cache() {
  static ShimCache c;
  return c;
}

dgr_plan_t cached_plan(ShimCache& c, const std::vector<dgr_block>& blocks, const dgr_bssn_params& p) {
  int dev = 0;
  hipcheck(hipGetDevice(&dev), "hipGetDevice");
  const bool same = c.plan && dev == c.device && blocks.size() == c.blocks.size() &&
                    std::memcmp(blocks.data(), c.blocks.data(), sizeof(dgr_block) * blocks.size()) == 0 &&
                    std::memcmp(&p, &c.p, sizeof(p)) == 0;
  if (same) return c.plan;
  if (c.plan) {
    dgr_plan_destroy(c.plan);
    c.plan = nullptr;
  }
  check(dgr_plan_create(&c.plan, blocks.data(), (uint32_t)blocks.size(), &p), "bssnRHS");
  c.blocks = blocks;
  c.p = p;
  c.device = dev;
  return c.plan;
}

void run(double** rhs, const double** u, const std::vector<dgr_block>& blocks) {
  const dgr_bssn_params p = to_params(dgr::bssn_params());
  bool dev = true;
  for (int v = 0; v < DGR_NUM_VARS; ++v) dev = dev && is_device(u[v]) && is_device(rhs[v]);
  ShimCache& c = cache();
  std::lock_guard<std::mutex> lock(c.m);
  if (dev) {
    dgr_plan_t plan = cached_plan(c, blocks, p);
    check(dgr_bssn_rhs(plan, rhs, u, nullptr), "bssnRHS");
    hipcheck(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    return;
  }
  // host arrays: stage the DOF range the blocks cover through device memory
  uint64_t lo = UINT64_MAX, hi = 0;
  for (const auto& b : blocks) {
    lo = std::min<uint64_t>(lo, b.offset);
    hi = std::max<uint64_t>(hi, b.offset + (uint64_t)b.sz[0] * b.sz[1] * b.sz[2]);
  }
  const size_t n = (size_t)(hi - lo);
  std::vector<dgr_block> shifted(blocks);
  for (auto& b : shifted) b.offset -= lo;
  dgr_plan_t plan = cached_plan(c, shifted, p);
  if (n > c.cap) {
    (void)hipFree(c.du);
    (void)hipFree(c.dr);
    c.du = c.dr = nullptr;
    c.cap = 0;
    hipcheck(hipMalloc(&c.du, sizeof(double) * n * DGR_NUM_VARS), "hipMalloc");
    hipcheck(hipMalloc(&c.dr, sizeof(double) * n * DGR_NUM_VARS), "hipMalloc");
    c.cap = n;
  }
  const double* up[DGR_NUM_VARS];
  double* rp[DGR_NUM_VARS];
  for (int v = 0; v < DGR_NUM_VARS; ++v) {
    up[v] = c.du + (size_t)v * c.cap;
    rp[v] = c.dr + (size_t)v * c.cap;
    hipcheck(hipMemcpy((void*)up[v], u[v] + lo, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    hipcheck(hipMemcpy(rp[v], rhs[v] + lo, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
  }
  check(dgr_bssn_rhs(plan, rp, up, nullptr), "bssnRHS");
  for (int v = 0; v < DGR_NUM_VARS; ++v)
    hipcheck(hipMemcpy(rhs[v] + lo, rp[v], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
}

}  // namespace

void bssnRHS(double** uzipVarsRHS, const double** uZipVars, const dgr::BlockDesc* blkList, unsigned int numBlocks) {
  if (numBlocks == 0) return;
  std::vector<dgr_block> blocks(numBlocks);
  for (unsigned b = 0; b < numBlocks; ++b) {
    blocks[b].offset = blkList[b].offset;
    blocks[b].bflag = blkList[b].bflag;
    for (int d = 0; d < 3; ++d) {
      blocks[b].sz[d] = blkList[b].sz[d];
      blocks[b].pmin[d] = blkList[b].ptmin[d];
      blocks[b].pmax[d] = blkList[b].ptmax[d];
    }
  }
  run(uzipVarsRHS, uZipVars, blocks);
}

void bssnrhs(double** uzipVarsRHS, const double** uZipVars, const unsigned int& offset, const double* ptmin,
             const double* ptmax, const unsigned int* sz, const unsigned int& bflag) {
  dgr::BlockDesc b;
  b.offset = offset;
  b.bflag = bflag;
  for (int d = 0; d < 3; ++d) {
    b.sz[d] = sz[d];
    b.ptmin[d] = ptmin[d];
    b.ptmax[d] = ptmax[d];
  }
  bssnRHS(uzipVarsRHS, uZipVars, &b, 1);
}

void physical_constraints(double** uZipConVars, const double** uZipVars, const unsigned int& offset,
                          const double* pmin, const double* pmax, const unsigned int* sz, const unsigned int& bflag) {
  const dgr_bssn_params p = to_params(dgr::bssn_params());
  dgr_block b{};
  b.offset = offset;
  b.bflag = bflag;
  for (int d = 0; d < 3; ++d) {
    b.sz[d] = sz[d];
    b.pmin[d] = pmin[d];
    b.pmax[d] = pmax[d];
  }
  const size_t n = (size_t)sz[0] * sz[1] * sz[2];
  bool dev = true;
  for (int v = 0; v < DGR_NUM_VARS; ++v) dev = dev && is_device(uZipVars[v]);
  for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q) dev = dev && is_device(uZipConVars[q]);
  // host arrays: the block's padded range goes through device memory
  double* buf = nullptr;
  const double* up[DGR_NUM_VARS];
  double* cp[DGR_NUM_CONSTRAINTS];
  if (dev) {
    for (int v = 0; v < DGR_NUM_VARS; ++v) up[v] = uZipVars[v];
    for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q) cp[q] = uZipConVars[q];
  } else {
    hipcheck(hipMalloc(&buf, sizeof(double) * n * (DGR_NUM_VARS + DGR_NUM_CONSTRAINTS)), "hipMalloc");
    b.offset = 0;
    for (int v = 0; v < DGR_NUM_VARS; ++v) up[v] = buf + (size_t)v * n;
    for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q) cp[q] = buf + (size_t)(DGR_NUM_VARS + q) * n;
  }
  dgr_plan_t plan = nullptr;
  try {
    if (!dev) {
      for (int v = 0; v < DGR_NUM_VARS; ++v)
        hipcheck(hipMemcpy((void*)up[v], uZipVars[v] + offset, sizeof(double) * n, hipMemcpyHostToDevice),
                 "hipMemcpy");
      for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q)
        hipcheck(hipMemcpy(cp[q], uZipConVars[q] + offset, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    check(dgr_plan_create(&plan, &b, 1, &p), "physical_constraints");
    check(dgr_physcon(plan, cp, up, nullptr), "physical_constraints");
    hipcheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    if (!dev)
      for (int q = 0; q < DGR_NUM_CONSTRAINTS; ++q)
        hipcheck(hipMemcpy(uZipConVars[q] + offset, cp[q], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  } catch (...) {
    if (plan) dgr_plan_destroy(plan);
    (void)hipFree(buf);
    throw;
  }
  dgr_plan_destroy(plan);
  (void)hipFree(buf);
}

void TwoPunctures(const double xx1, const double yy1, const double zz1, double* vars, double* mp, double* mm,
                  double* mp_adm, double* mm_adm, double* E, double* J1, double* J2, double* J3) {
  // the solve on the first call, kept for the process (TwoPunctures.cpp:48-200)
  static std::mutex m;
  static dgr_tp_t tp = nullptr;
  static dgr_tp_info_t info;
  {
    std::lock_guard<std::mutex> lock(m);
    if (!tp) {
      check(dgr_tp_solve(&tp, &dgr::tp_params()), "TwoPunctures (dgr_tp_solve)");
      check(dgr_tp_info(tp, &info), "dgr_tp_info");
    }
  }
  check(dgr_tp_point(tp, xx1, yy1, zz1, vars), "TwoPunctures (dgr_tp_point)");
  *mp = info.mp;
  *mm = info.mm;
  *mp_adm = info.mp_adm;
  *mm_adm = info.mm_adm;
  *E = info.E;
  *J1 = info.J[0];
  *J2 = info.J[1];
  *J3 = info.J[2];
}

// Host-side mirror of bssnRHS / bssnrhs (see bssn_rhs_host.h).
#include "bssn_rhs_host.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/dgr_bssn.h"

namespace dgr {

BSSNGlobals& bssn_params() {
  static BSSNGlobals g;
  return g;
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

void run(double** rhs, const double** u, const std::vector<dgr_block>& blocks) {
  const dgr_bssn_params p = to_params(dgr::bssn_params());
  bool dev = true;
  for (int v = 0; v < DGR_NUM_VARS; ++v) dev = dev && is_device(u[v]) && is_device(rhs[v]);
  if (dev) {
    check(dgr_bssn_rhs_blocks(rhs, u, blocks.data(), (uint32_t)blocks.size(), &p, nullptr), "bssnRHS");
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
  double* du = nullptr;
  double* dr = nullptr;
  hipcheck(hipMalloc(&du, sizeof(double) * n * DGR_NUM_VARS), "hipMalloc");
  if (hipMalloc(&dr, sizeof(double) * n * DGR_NUM_VARS) != hipSuccess) {
    (void)hipFree(du);
    throw std::runtime_error("bssnRHS: hipMalloc failed");
  }
  const double* up[DGR_NUM_VARS];
  double* rp[DGR_NUM_VARS];
  try {
    for (int v = 0; v < DGR_NUM_VARS; ++v) {
      up[v] = du + (size_t)v * n;
      rp[v] = dr + (size_t)v * n;
      hipcheck(hipMemcpy((void*)up[v], u[v] + lo, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
      hipcheck(hipMemcpy(rp[v], rhs[v] + lo, sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy");
    }
    check(dgr_bssn_rhs_blocks(rp, up, shifted.data(), (uint32_t)shifted.size(), &p, nullptr), "bssnRHS");
    for (int v = 0; v < DGR_NUM_VARS; ++v)
      hipcheck(hipMemcpy(rhs[v] + lo, rp[v], sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy");
  } catch (...) {
    (void)hipFree(du);
    (void)hipFree(dr);
    throw;
  }
  (void)hipFree(du);
  (void)hipFree(dr);
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

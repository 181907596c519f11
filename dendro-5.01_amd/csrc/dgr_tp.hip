/*
 * TwoPunctures initial data (BSSN_GR/src/TwoPunctures.cpp:24-498 with the
 * solver of TPNewton.cpp, TPFuncAndJacobian.cpp, TPCoordTransf.cpp and
 * TPEquations.cpp): two Bowen-York punctures at x = +-b (plus a centre
 * offset) with momenta P and spins S; the conformal factor is
 *   psi = 1 + m+/(2 r+) + m-/(2 r-) + u,
 * where u solves the Hamiltonian constraint
 *   Lap u + (1/8) A_ij A^ij psi^-7 = 0        (TPEquations.cpp:151-169)
 * with the Bowen-York A_ij (TPEquations.cpp:81-145).
 *
 * Discretisation (Ansorg, Bruegmann, Tichy 2004): prolate spheroidal
 * coordinates x = b cosh X cos R, rho = b sinh X sin R, phi around the x
 * axis, compactified as A = tanh(X/2) in (0, 1) and B = tan((R - pi/2)/2)
 * in (-1, 1); u = (A - 1) V with V expanded in Chebyshev polynomials of
 * a = 2A - 1 and of B and in a Fourier series of phi, collocated at the
 * Chebyshev-Gauss points and at phi_k = 2 pi k / n_phi (the reference's
 * grid, TPFuncAndJacobian.cpp:221-225).  The flat Laplacian in (X, R, phi) is
 *   [u_XX + coth X u_X + u_RR + cot R u_R] / (b^2 (sinh^2 X + sin^2 R))
 *     + u_phiphi / (b^2 sinh^2 X sin^2 R),
 * taken to (A, B) by the chain rule (TPCoordTransf.cpp:18-50 does the same
 * in its own variables).
 *
 * Newton's method on the collocation equations: each step assembles the
 * dense Jacobian (N = n_A n_B n_phi unknowns) on the GPU and solves it with
 * rocSOLVER's LU (FP64, the work of matrix multiplications); the rest of
 * the solve is O(N n) host work.  The reference solves the same linear
 * systems with a relaxation-preconditioned iterative method
 * (TPNewton.cpp).  With the solution's spectral coefficients on the GPU,
 * the initial data of every padded point of a plan's blocks is one kernel:
 * the spectral sum for u, then psi, the Bowen-York A_ij, the lapse and the
 * ADM-to-BSSN conversion (TwoPunctures.cpp:270-498).
 */
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "bssn_staged.h"
#include "dgr_internal.h"

struct dgr_tp_s {
  dgr_tp_params p{};
  int n1 = 0, n2 = 0, n3 = 0, na = 0;  // na: Chebyshev-in-a terms padded to the kernel's register count
  std::vector<double> V;                // solution at the collocation points, i + n1 (j + n2 k)
  std::vector<double> coef;             // coefficients [f][l][m], m < na (zero beyond n1)
  double mp = 0, mm = 0, mp_adm = 0, mm_adm = 0, E = 0, J[3] = {0, 0, 0};
  int newton_steps = 0;
  double last_update = 0;
  double* d_coef = nullptr;
  int device = 0;
};

namespace {

using namespace dgr_internal;

constexpr double kPi = 3.14159265358979323846;

// Differentiation matrix at distinct nodes x (barycentric form), row-major.
std::vector<double> diff_matrix(const std::vector<double>& x) {
  const int n = (int)x.size();
  std::vector<double> w(n, 1.0), D((size_t)n * n, 0.0);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k)
      if (k != j) w[j] /= x[j] - x[k];
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) {
        D[(size_t)i * n + j] = (w[j] / w[i]) / (x[i] - x[j]);
        s += D[(size_t)i * n + j];
      }
    D[(size_t)i * n + i] = -s;
  }
  return D;
}

std::vector<double> matmul(const std::vector<double>& a, const std::vector<double>& b, int n) {
  std::vector<double> c((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k)
      for (int j = 0; j < n; ++j) c[(size_t)i * n + j] += a[(size_t)i * n + k] * b[(size_t)k * n + j];
  return c;
}

// Second derivative of the trigonometric interpolant at n equispaced points.
std::vector<double> fourier_d2(int n) {
  std::vector<double> D((size_t)n * n, 0.0);
  if (n < 2) return D;
  const double h = 2.0 * kPi / n;
  for (int k = 0; k < n; ++k)
    for (int l = 0; l < n; ++l) {
      const int d = k - l;
      const double sg = (d % 2 == 0) ? 1.0 : -1.0;
      if (d == 0)
        D[(size_t)k * n + l] = n % 2 == 0 ? -kPi * kPi / (3.0 * h * h) - 1.0 / 6.0 : -(double)(n * n - 1) / 12.0;
      else if (n % 2 == 0)
        D[(size_t)k * n + l] = -sg / (2.0 * std::pow(std::sin(d * h / 2.0), 2));
      else
        D[(size_t)k * n + l] = -sg * std::cos(d * h / 2.0) / (2.0 * std::pow(std::sin(d * h / 2.0), 2));
    }
  return D;
}

// Bowen-York A_ij of both punctures at (x, y, z) relative to the centre
// (TPEquations.cpp:81-145).  With eps > 0 the squared distances are
// regularised as sqrt(r^4 + eps^4) and floored at tiny^2, as for the grid.
template <typename Real>
__host__ __device__ inline void by_aij(const dgr_tp_params& p, Real x, Real y, Real z, Real A[3][3], bool regular) {
  const Real d[2][3] = {{x - p.par_b, y, z}, {x + p.par_b, y, z}};
  const double* P[2] = {p.par_P_plus, p.par_P_minus};
  const double* S[2] = {p.par_S_plus, p.par_S_minus};
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) A[i][j] = 0.0;
  for (int q = 0; q < 2; ++q) {
    Real r2 = d[q][0] * d[q][0] + d[q][1] * d[q][1] + d[q][2] * d[q][2];
    if (regular) {
      r2 = sqrt(r2 * r2 + p.TP_epsilon * p.TP_epsilon * p.TP_epsilon * p.TP_epsilon);
      if (r2 < p.TP_Tiny * p.TP_Tiny) r2 = p.TP_Tiny * p.TP_Tiny;
    }
    const Real r = sqrt(r2), r3 = r * r2;
    const Real n[3] = {d[q][0] / r, d[q][1] / r, d[q][2] / r};
    const Real nP = n[0] * P[q][0] + n[1] * P[q][1] + n[2] * P[q][2];
    const Real nS[3] = {n[1] * S[q][2] - n[2] * S[q][1], n[2] * S[q][0] - n[0] * S[q][2],
                        n[0] * S[q][1] - n[1] * S[q][0]};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        Real a = 1.5 * (P[q][i] * n[j] + P[q][j] * n[i] + nP * n[i] * n[j]) / r2 -
                 3.0 * (nS[i] * n[j] + nS[j] * n[i]) / r3;
        if (i == j) a -= 1.5 * nP / r2;
        A[i][j] += a;
      }
  }
}

// Collocation grid, 1D operators and the pointwise coefficients of the
// linear part: Lap((A-1) V) = cAA (2 V_A + T V_AA) + cA (V + T V_A)
//   + T (cBB V_BB + cB V_B + cpp V_pp),  T = A - 1.
struct Grid {
  int n1, n2, n3, N;
  std::vector<double> a, B, ph;
  std::vector<double> DA, DAA, DB, DBB, Dpp;
  std::vector<double> T, cAA, cA, cBB, cB, cpp, psiBL, AA, xyz;
};

// Brill-Lindquist part of psi at the collocation points for bare masses mp, mm.
void set_masses(Grid& g, const dgr_tp_params& p, double mp, double mm) {
  const double b = p.par_b;
  for (int q = 0; q < g.N; ++q) {
    const double x = g.xyz[3 * (size_t)q], y = g.xyz[3 * (size_t)q + 1], z = g.xyz[3 * (size_t)q + 2];
    const double rp = std::sqrt((x - b) * (x - b) + y * y + z * z);
    const double rm = std::sqrt((x + b) * (x + b) + y * y + z * z);
    g.psiBL[q] = 1.0 + 0.5 * mp / rp + 0.5 * mm / rm;
  }
}

Grid make_grid(const dgr_tp_params& p) {
  Grid g;
  g.n1 = p.npoints_A;
  g.n2 = p.npoints_B;
  g.n3 = p.npoints_phi;
  g.N = g.n1 * g.n2 * g.n3;
  for (int i = 0; i < g.n1; ++i) g.a.push_back(-std::cos(kPi * (2 * i + 1) / (2.0 * g.n1)));
  for (int j = 0; j < g.n2; ++j) g.B.push_back(-std::cos(kPi * (2 * j + 1) / (2.0 * g.n2)));
  for (int k = 0; k < g.n3; ++k) g.ph.push_back(2.0 * kPi * k / g.n3);
  g.DA = diff_matrix(g.a);
  for (double& v : g.DA) v *= 2.0;  // d/dA = 2 d/da
  g.DAA = matmul(g.DA, g.DA, g.n1);
  g.DB = diff_matrix(g.B);
  g.DBB = matmul(g.DB, g.DB, g.n2);
  g.Dpp = fourier_d2(g.n3);
  const double b = p.par_b;
  g.T.resize(g.N); g.cAA.resize(g.N); g.cA.resize(g.N); g.cBB.resize(g.N); g.cB.resize(g.N);
  g.cpp.resize(g.N); g.psiBL.resize(g.N); g.AA.resize(g.N); g.xyz.resize(3 * (size_t)g.N);
  for (int k = 0; k < g.n3; ++k)
    for (int j = 0; j < g.n2; ++j)
      for (int i = 0; i < g.n1; ++i) {
        const int q = i + g.n1 * (j + g.n2 * k);
        const double A = 0.5 * (g.a[i] + 1.0), Bv = g.B[j];
        const double AX = 0.5 * (1.0 - A * A), AXX = -A * AX;
        const double BR = 0.5 * (1.0 + Bv * Bv), BRR = Bv * BR;
        const double coth = (1.0 + A * A) / (2.0 * A), sh = 2.0 * A / (1.0 - A * A);
        const double cot = -2.0 * Bv / (1.0 - Bv * Bv), sn = (1.0 - Bv * Bv) / (1.0 + Bv * Bv);
        const double den = b * b * (sh * sh + sn * sn);
        g.T[q] = A - 1.0;
        g.cAA[q] = AX * AX / den;
        g.cA[q] = (AXX + coth * AX) / den;
        g.cBB[q] = BR * BR / den;
        g.cB[q] = (BRR + cot * BR) / den;
        g.cpp[q] = 1.0 / (b * b * sh * sh * sn * sn);
        const double chX = (1.0 + A * A) / (1.0 - A * A), cR = -2.0 * Bv / (1.0 + Bv * Bv);
        const double x = b * chX * cR, rho = b * sh * sn;
        const double y = rho * std::cos(g.ph[k]), z = rho * std::sin(g.ph[k]);
        g.xyz[3 * (size_t)q] = x;
        g.xyz[3 * (size_t)q + 1] = y;
        g.xyz[3 * (size_t)q + 2] = z;
        double Aij[3][3];
        by_aij<double>(p, x, y, z, Aij, false);
        double s = 0.0;
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) s += Aij[r][c] * Aij[r][c];
        g.AA[q] = s;
      }
  return g;
}

// Laplacian part applied to V (host).
void apply_lap(const Grid& g, const std::vector<double>& V, std::vector<double>& out) {
  out.assign(g.N, 0.0);
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        const int q = i + n1 * (j + n2 * k);
        double vA = 0, vAA = 0, vB = 0, vBB = 0, vpp = 0;
        for (int m = 0; m < n1; ++m) {
          const double v = V[m + n1 * (j + n2 * k)];
          vA += g.DA[(size_t)i * n1 + m] * v;
          vAA += g.DAA[(size_t)i * n1 + m] * v;
        }
        for (int m = 0; m < n2; ++m) {
          const double v = V[i + n1 * (m + n2 * k)];
          vB += g.DB[(size_t)j * n2 + m] * v;
          vBB += g.DBB[(size_t)j * n2 + m] * v;
        }
        for (int m = 0; m < n3; ++m) vpp += g.Dpp[(size_t)k * n3 + m] * V[i + n1 * (j + n2 * m)];
        const double T = g.T[q];
        out[q] = g.cAA[q] * (2.0 * vA + T * vAA) + g.cA[q] * (V[q] + T * vA) +
                 T * (g.cBB[q] * vBB + g.cB[q] * vB + g.cpp[q] * vpp);
      }
}

struct DevGrid {
  const double *DA, *DAA, *DB, *DBB, *Dpp, *T, *cAA, *cA, *cBB, *cB, *cpp, *diag, *rs;
};

// Largest magnitude in each row of the Jacobian, from the operator's
// structure (row p couples its A line, B line and phi line).
void row_max(const Grid& g, const std::vector<double>& diag, std::vector<double>& out) {
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  out.resize(g.N);
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        const int p = i + n1 * (j + n2 * k);
        const double T = g.T[p];
        double m = 0.0, dsum = diag[p];
        for (int i2 = 0; i2 < n1; ++i2) {
          const double v = g.cAA[p] * (2.0 * g.DA[(size_t)i * n1 + i2] + T * g.DAA[(size_t)i * n1 + i2]) +
                           g.cA[p] * ((i == i2 ? 1.0 : 0.0) + T * g.DA[(size_t)i * n1 + i2]);
          if (i2 == i) dsum += v; else m = std::max(m, std::fabs(v));
        }
        for (int j2 = 0; j2 < n2; ++j2) {
          const double v = T * (g.cBB[p] * g.DBB[(size_t)j * n2 + j2] + g.cB[p] * g.DB[(size_t)j * n2 + j2]);
          if (j2 == j) dsum += v; else m = std::max(m, std::fabs(v));
        }
        for (int k2 = 0; k2 < n3; ++k2) {
          const double v = T * g.cpp[p] * g.Dpp[(size_t)k * n3 + k2];
          if (k2 == k) dsum += v; else m = std::max(m, std::fabs(v));
        }
        out[p] = std::max(m, std::fabs(dsum));
      }
}

// Dense Jacobian, column-major (LAPACK): J[p + q N], row p scaled by rs[p].
__global__ void __launch_bounds__(256) k_tp_jacobian(DevGrid G, int n1, int n2, int n3, double* __restrict__ J) {
  const long long N = (long long)n1 * n2 * n3, NN = N * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < NN; e += (long long)gridDim.x * blockDim.x) {
    const long long p = e % N, q = e / N;
    const int i = (int)(p % n1), j = (int)((p / n1) % n2), k = (int)(p / ((long long)n1 * n2));
    const int i2 = (int)(q % n1), j2 = (int)((q / n1) % n2), k2 = (int)(q / ((long long)n1 * n2));
    const double T = G.T[p];
    double v = 0.0;
    if (j == j2 && k == k2)
      v += G.cAA[p] * (2.0 * G.DA[i * n1 + i2] + T * G.DAA[i * n1 + i2]) +
           G.cA[p] * ((i == i2 ? 1.0 : 0.0) + T * G.DA[i * n1 + i2]);
    if (i == i2 && k == k2) v += T * (G.cBB[p] * G.DBB[j * n2 + j2] + G.cB[p] * G.DB[j * n2 + j2]);
    if (i == i2 && j == j2) v += T * G.cpp[p] * G.Dpp[k * n3 + k2];
    if (p == q) v += G.diag[p];
    J[e] = v * G.rs[p];
  }
}

struct Solver {
  const Grid& g;
  rocblas_handle h = nullptr;
  hipStream_t st = nullptr;
  double *dJ = nullptr, *dF = nullptr, *dsmall = nullptr;
  rocblas_int *dpiv = nullptr, *dinfo = nullptr;
  DevGrid G{};
  explicit Solver(const Grid& grid) : g(grid) {}
  ~Solver() {
    (void)hipFree(dJ); (void)hipFree(dF); (void)hipFree(dsmall); (void)hipFree(dpiv); (void)hipFree(dinfo);
    if (h) (void)rocblas_destroy_handle(h);
    if (st) (void)hipStreamDestroy(st);
  }
  int init() {
    const size_t N = g.N;
    DGR_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    if (rocblas_create_handle(&h) != rocblas_status_success) return DGR_E_NOMEM;
    if (rocblas_set_stream(h, st) != rocblas_status_success) return DGR_E_INVALID;
    if (hipMalloc(&dJ, sizeof(double) * N * N) != hipSuccess || hipMalloc(&dF, sizeof(double) * N) != hipSuccess ||
        hipMalloc(&dpiv, sizeof(rocblas_int) * N) != hipSuccess || hipMalloc(&dinfo, sizeof(rocblas_int)) != hipSuccess)
      return DGR_E_NOMEM;
    // small operators and pointwise coefficients in one buffer
    const size_t n1 = g.n1, n2 = g.n2, n3 = g.n3;
    const size_t sizes[13] = {n1 * n1, n1 * n1, n2 * n2, n2 * n2, n3 * n3, N, N, N, N, N, N, N, N};
    size_t tot = 0;
    for (size_t s : sizes) tot += s;
    DGR_HIP(hipMalloc(&dsmall, sizeof(double) * tot));
    const std::vector<double>* src[11] = {&g.DA, &g.DAA, &g.DB, &g.DBB, &g.Dpp, &g.T, &g.cAA, &g.cA, &g.cBB, &g.cB, &g.cpp};
    const double** dst[13] = {&G.DA,  &G.DAA, &G.DB, &G.DBB, &G.Dpp, &G.T,   &G.cAA,
                              &G.cA,  &G.cBB, &G.cB, &G.cpp, &G.diag, &G.rs};
    size_t off = 0;
    for (int q = 0; q < 13; ++q) {
      *dst[q] = dsmall + off;
      if (q < 11) DGR_HIP(hipMemcpyAsync(dsmall + off, src[q]->data(), sizeof(double) * sizes[q], hipMemcpyHostToDevice, st));
      off += sizes[q];
    }
    DGR_HIP(hipStreamSynchronize(st));
    return DGR_OK;
  }
  // One Newton step on V; returns max |update|.  Each equation (row of the
  // Jacobian and entry of the residual) is divided by the row's largest
  // magnitude: the rows next to the punctures and the axis are up to 1e17
  // times larger than others, and partial pivoting on the unscaled matrix
  // (condition ~1e19-1e23 at 16-30 points per direction, ~1e3 scaled)
  // returns updates dominated by rounding.
  int step(std::vector<double>& V, double* upd) {
    const int N = g.N;
    std::vector<double> F, diag(N), rs;
    apply_lap(g, V, F);
    for (int q = 0; q < N; ++q) {
      const double psi = g.psiBL[q] + g.T[q] * V[q];
      const double p2 = psi * psi, p4 = p2 * p2, p7 = psi * p2 * p4, p8 = p4 * p4;
      F[q] += 0.125 * g.AA[q] / p7;
      diag[q] = -0.875 * g.AA[q] / p8 * g.T[q];
    }
    row_max(g, diag, rs);
    for (int q = 0; q < N; ++q) {
      rs[q] = rs[q] > 0.0 ? 1.0 / rs[q] : 1.0;
      F[q] *= rs[q];
    }
    DGR_HIP(hipMemcpyAsync(const_cast<double*>(G.diag), diag.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    DGR_HIP(hipMemcpyAsync(const_cast<double*>(G.rs), rs.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    DGR_HIP(hipMemcpyAsync(dF, F.data(), sizeof(double) * N, hipMemcpyHostToDevice, st));
    const long long NN = (long long)N * N;
    const unsigned grid = (unsigned)std::min<long long>((NN + 255) / 256, 1 << 20);
    hipLaunchKernelGGL(k_tp_jacobian, dim3(grid), dim3(256), 0, st, G, g.n1, g.n2, g.n3, dJ);
    DGR_HIP(hipGetLastError());
    if (rocsolver_dgetrf(h, N, N, dJ, N, dpiv, dinfo) != rocblas_status_success) return DGR_E_INVALID;
    if (rocsolver_dgetrs(h, rocblas_operation_none, N, 1, dJ, N, dpiv, dF, N) != rocblas_status_success)
      return DGR_E_INVALID;
    rocblas_int info = 0;
    DGR_HIP(hipMemcpyAsync(&info, dinfo, sizeof(info), hipMemcpyDeviceToHost, st));
    DGR_HIP(hipMemcpyAsync(F.data(), dF, sizeof(double) * N, hipMemcpyDeviceToHost, st));
    DGR_HIP(hipStreamSynchronize(st));
    if (info != 0) return DGR_E_INVALID;  // singular Jacobian
    double m = 0.0;
    for (int q = 0; q < N; ++q) {
      V[q] -= F[q];
      m = std::max(m, std::fabs(F[q]));
    }
    *upd = m;
    return DGR_OK;
  }
};

// Chebyshev T_0..T_{n-1}(x).
void cheb(double x, int n, double* t) {
  if (n > 0) t[0] = 1.0;
  if (n > 1) t[1] = x;
  for (int m = 2; m < n; ++m) t[m] = 2.0 * x * t[m - 1] - t[m - 2];
}

// Real Fourier basis at phi: 1, cos phi, sin phi, cos 2phi, ..., and for
// even n the last one cos(n/2 phi).
__host__ __device__ inline void fourier(double phi, int n, double* f) {
  f[0] = 1.0;
  const int M = (n - 1) / 2;
  for (int m = 1; m <= M; ++m) {
    f[2 * m - 1] = cos(m * phi);
    f[2 * m] = sin(m * phi);
  }
  if (n % 2 == 0 && n >= 2) f[n - 1] = cos((n / 2) * phi);
}

// Spectral coefficients of V: Chebyshev in a and B (discrete orthogonality
// at Gauss points), real Fourier in phi; layout [f][l][m], m < na.
std::vector<double> spectral(const Grid& g, const std::vector<double>& V, int na) {
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  std::vector<double> c1((size_t)n1 * n2 * n3), c2((size_t)n1 * n2 * n3);
  std::vector<double> t(std::max(n1, n2));
  // along a
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int m = 0; m < n1; ++m) {
        double s = 0.0;
        for (int i = 0; i < n1; ++i) s += V[i + n1 * (j + n2 * k)] * std::cos(m * (kPi - kPi * (2 * i + 1) / (2.0 * n1)));
        c1[m + n1 * (j + n2 * k)] = s * (m == 0 ? 1.0 : 2.0) / n1;
      }
  // along B
  for (int k = 0; k < n3; ++k)
    for (int l = 0; l < n2; ++l)
      for (int m = 0; m < n1; ++m) {
        double s = 0.0;
        for (int j = 0; j < n2; ++j) s += c1[m + n1 * (j + n2 * k)] * std::cos(l * (kPi - kPi * (2 * j + 1) / (2.0 * n2)));
        c2[m + n1 * (l + n2 * k)] = s * (l == 0 ? 1.0 : 2.0) / n2;
      }
  // along phi
  std::vector<double> out((size_t)na * n2 * n3, 0.0);
  const int M = (n3 - 1) / 2;
  for (int l = 0; l < n2; ++l)
    for (int m = 0; m < n1; ++m) {
      auto at = [&](int k) { return c2[m + n1 * (l + n2 * k)]; };
      auto put = [&](int f, double v) { out[m + (size_t)na * (l + (size_t)n2 * f)] = v; };
      double s0 = 0.0;
      for (int k = 0; k < n3; ++k) s0 += at(k);
      put(0, s0 / n3);
      for (int q = 1; q <= M; ++q) {
        double sc = 0.0, ss = 0.0;
        for (int k = 0; k < n3; ++k) {
          sc += at(k) * std::cos(q * g.ph[k]);
          ss += at(k) * std::sin(q * g.ph[k]);
        }
        put(2 * q - 1, 2.0 * sc / n3);
        put(2 * q, 2.0 * ss / n3);
      }
      if (n3 % 2 == 0 && n3 >= 2) {
        double s = 0.0;
        for (int k = 0; k < n3; ++k) s += at(k) * ((k % 2 == 0) ? 1.0 : -1.0);
        put(n3 - 1, s / n3);
      }
    }
  return out;
}

// (A, B, phi) of a point relative to the centre: cosh X = (r+ + r-)/(2b),
// cos R = (r- - r+)/(2b), A = tanh(X/2), B = tan((R - pi/2)/2).
__host__ __device__ inline void to_abphi(double b, double x, double y, double z, double& A, double& B, double& phi) {
  const double rp = sqrt((x - b) * (x - b) + y * y + z * z), rm = sqrt((x + b) * (x + b) + y * y + z * z);
  double s = (rp + rm) / (2.0 * b), c = (rm - rp) / (2.0 * b);
  s = s < 1.0 ? 1.0 : s;
  c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
  A = sqrt((s - 1.0) / (s + 1.0));
  // tan(R/2 - pi/4) = (tan(R/2) - 1) / (tan(R/2) + 1), tan(R/2) = sqrt((1-c)/(1+c)),
  // in a form finite at both punctures (c = +-1)
  const double sm = sqrt(1.0 - c), sp = sqrt(1.0 + c);
  B = (sm - sp) / (sm + sp);
  phi = atan2(z, y);
}

// V at (A, B, phi) from the coefficients (host).
double eval_host(const dgr_tp_s& T, double A, double B, double phi) {
  std::vector<double> ta(T.n1), tb(T.n2), f(T.n3);
  cheb(2.0 * A - 1.0, T.n1, ta.data());
  cheb(B, T.n2, tb.data());
  fourier(phi, T.n3, f.data());
  double v = 0.0;
  for (int q = 0; q < T.n3; ++q)
    for (int l = 0; l < T.n2; ++l)
      for (int m = 0; m < T.n1; ++m) v += f[q] * tb[l] * ta[m] * T.coef[m + (size_t)T.na * (l + (size_t)T.n2 * q)];
  return v;
}

// u at a point relative to the centre (host).
double u_host(const dgr_tp_s& T, double x, double y, double z) {
  double A, B, phi;
  to_abphi(T.p.par_b, x, y, z, A, B, phi);
  return (A - 1.0) * eval_host(T, A, B, phi);
}

// M (3 r^4 / (8 R^5) - 5 r^2 / (4 R^3) + 15 / (8 R)): the smooth
// continuation of M / r inside r < R (TwoPunctures.cpp:17-21).
__host__ __device__ inline double extend(double M, double r, double R) {
  return M * (3.0 / 8.0 * r * r * r * r / (R * R * R * R * R) - 5.0 / 4.0 * r * r / (R * R * R) + 15.0 / 8.0 / R);
}

// The 24 BSSN variables at a point (coordinates relative to the centre)
// from u there (TwoPunctures.cpp:285-498): psi with the regularised
// distances, the lapse of initial_lapse, gd = psi^4 delta, K_ij =
// A_ij / psi^2, then chi = det(gd)^(-1/3), gt = chi gd, K = tr K and
// At = chi (K_ij - gd_ij K / 3); Gt = beta = B = 0.
__host__ __device__ inline void tp_point(const dgr_tp_params& P, double mp, double mm, double xx, double yy,
                                         double zz, double U, double out[24]) {
  double rp = sqrt((xx - P.par_b) * (xx - P.par_b) + yy * yy + zz * zz);
  double rm = sqrt((xx + P.par_b) * (xx + P.par_b) + yy * yy + zz * zz);
  const double e4 = P.TP_epsilon * P.TP_epsilon * P.TP_epsilon * P.TP_epsilon;
  rp = pow(rp * rp * rp * rp + e4, 0.25);
  rm = pow(rm * rm * rm * rm + e4, 0.25);
  if (rp < P.TP_Tiny) rp = P.TP_Tiny;
  if (rm < P.TP_Tiny) rm = P.TP_Tiny;
  const double R = P.TP_Extend_Radius;
  const double ip = rp < R ? extend(1.0, rp, R) : 1.0 / rp;  // 1/r, continued inside R
  const double im = rm < R ? extend(1.0, rm, R) : 1.0 / rm;
  const double psi1 = 1.0 + 0.5 * mp * ip + 0.5 * mm * im + U;
  double alpha = 1.0;
  if (P.initial_lapse == DGR_TP_LAPSE_ANTISYMMETRIC || P.initial_lapse == DGR_TP_LAPSE_AVERAGED) {
    alpha = (1.0 - 0.5 * mp * ip - 0.5 * mm * im) / (1.0 + 0.5 * mp * ip + 0.5 * mm * im);
    if (P.initial_lapse == DGR_TP_LAPSE_AVERAGED) alpha = 0.5 * (1.0 + alpha);
  } else if (P.initial_lapse == DGR_TP_LAPSE_PSIN || P.initial_lapse == DGR_TP_LAPSE_BROWNSVILLE) {
    const double ps = 1.0 + 0.5 * mp * ip + 0.5 * mm * im;  // psi without u
    alpha = pow(ps, P.initial_lapse_psi_exponent);
    if (P.initial_lapse == DGR_TP_LAPSE_BROWNSVILLE) alpha = 2.0 / (1.0 + alpha);
  }
  double Aij[3][3];
  by_aij<double>(P, xx, yy, zz, Aij, true);
  const double p2 = psi1 * psi1, p4 = p2 * p2;
  const double det = p4 * p4 * p4, chi = pow(1.0 / det, 1.0 / 3.0);
  const double gu = 1.0 / p4;
  const double trK = gu * (Aij[0][0] + Aij[1][1] + Aij[2][2]) / p2;
  out[dgr::V_ALPHA] = alpha;
  out[dgr::V_CHI] = chi;
  out[dgr::V_K] = trK;
  for (int c = 0; c < 3; ++c) {
    out[dgr::V_GT0 + c] = 0.0;
    out[dgr::V_BETA0 + c] = 0.0;
    out[dgr::V_B0 + c] = 0.0;
  }
  for (int q = 0; q < 6; ++q) {
    const int r = dgr::spi(q), c = dgr::spj(q);
    const double gd = r == c ? p4 : 0.0;
    out[dgr::V_SG0 + q] = chi * gd;
    out[dgr::V_SA0 + q] = chi * (Aij[r][c] / p2 - gd * trK / 3.0);
  }
}

constexpr int kLdsCoef = 16384;  // coefficients staged in LDS (128 KB) when they fit

template <int NA, bool LDS>
__global__ void __launch_bounds__(512) k_tp_eval(const DevBlock* __restrict__ blocks, uint32_t nblocks,
                                                 dgr_tp_params P, double mp, double mm, int n2, int n3,
                                                 const double* __restrict__ coef, OutPtrs u) {
  __shared__ double sc[LDS ? kLdsCoef : 1];
  const int ncoef = NA * n2 * n3;
  if constexpr (LDS) {
    for (int e = threadIdx.x; e < ncoef; e += blockDim.x) sc[e] = coef[e];
    __syncthreads();
  }
  const double* C = LDS ? sc : coef;
  for (uint32_t b = blockIdx.x; b < nblocks; b += gridDim.x) {
    const DevBlock& Bk = blocks[b];
    const int s0 = Bk.sz[0], s1 = Bk.sz[1], n = s0 * s1 * (int)Bk.sz[2];
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const int i = e % s0, j = (e / s0) % s1, k = e / (s0 * s1);
      double xx, yy, zz;
      {
#pragma clang fp contract(off)
        xx = Bk.pmin[0] + i * Bk.h[0] - P.center_offset[0];
        yy = Bk.pmin[1] + j * Bk.h[1] - P.center_offset[1];
        zz = Bk.pmin[2] + k * Bk.h[2] - P.center_offset[2];
      }
      // u = (A - 1) V, V the spectral sum
      double A, Bc, phi;
      to_abphi(P.par_b, xx, yy, zz, A, Bc, phi);
      double ta[NA];
      const double a = 2.0 * A - 1.0;
      ta[0] = 1.0;
      ta[1] = a;
#pragma unroll
      for (int m = 2; m < NA; ++m) ta[m] = 2.0 * a * ta[m - 1] - ta[m - 2];
      double V = 0.0;
      for (int f = 0; f < n3; ++f) {
        const int fm = (f + 1) / 2;
        const double ff = f == 0 ? 1.0 : ((f % 2 == 1 || (n3 % 2 == 0 && f == n3 - 1)) ? cos(fm * phi) : sin(fm * phi));
        double tb0 = 1.0, tb1 = Bc, sf = 0.0;
        for (int l = 0; l < n2; ++l) {
          const double tb = l == 0 ? tb0 : (l == 1 ? tb1 : 2.0 * Bc * tb1 - tb0);
          if (l >= 2) { tb0 = tb1; tb1 = tb; }
          const double* c = C + (size_t)NA * (l + (size_t)n2 * f);
          double s = 0.0;
#pragma unroll
          for (int m = 0; m < NA; ++m) s += ta[m] * c[m];
          sf += tb * s;
        }
        V += ff * sf;
      }
      const double U = (A - 1.0) * V;
      double out[24];
      tp_point(P, mp, mm, xx, yy, zz, U, out);
      const uint64_t o = Bk.offset + (uint64_t)e;
#pragma unroll
      for (int v = 0; v < 24; ++v) u.v[v][o] = out[v];
    }
  }
}

int check_params(const dgr_tp_params* p) {
  if (!p || !(p->par_b > 0.0) || p->npoints_A < 4 || p->npoints_B < 4 || p->npoints_phi < 1 ||
      p->npoints_A > 64 || p->npoints_B > 256 || p->npoints_phi > 256)
    return DGR_E_INVALID;
  const long long N = (long long)p->npoints_A * p->npoints_B * p->npoints_phi;
  if (N > DGR_TP_MAX_UNKNOWNS) return DGR_E_INVALID;
  if (p->give_bare_mass ? !(p->par_m_plus > 0.0 && p->par_m_minus > 0.0)
                        : !(p->target_M_plus > 0.0 && p->target_M_minus > 0.0 && p->par_m_plus > 0.0 &&
                            p->par_m_minus > 0.0))
    return DGR_E_INVALID;
  if (p->Newton_maxit < 1 || p->initial_lapse < 0 || p->initial_lapse > DGR_TP_LAPSE_UNIT) return DGR_E_INVALID;
  return DGR_OK;
}

// Newton steps until the update is below tol or maxit steps were taken.
int newton(Solver& S, std::vector<double>& V, double tol, int maxit, dgr_tp_s* tp) {
  for (int it = 0; it < maxit; ++it) {
    double upd = 0.0;
    const int rc = S.step(V, &upd);
    if (rc) return rc;
    tp->newton_steps++;
    tp->last_update = upd;
    if (upd < tol) break;
  }
  return DGR_OK;
}

}  // namespace

extern "C" void dgr_tp_params_default(dgr_tp_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  // BSSN_GR/src/parameters.cpp:142-151 and include/parameters.h:271-281
  p->par_b = 1.0;
  p->par_m_plus = 1.0;
  p->par_m_minus = 1.0;
  p->target_M_plus = 1.0;
  p->target_M_minus = 1.0;
  p->give_bare_mass = 0;
  p->npoints_A = 30;
  p->npoints_B = 30;
  p->npoints_phi = 16;
  p->Newton_tol = 1.0e-10;
  p->Newton_maxit = 5;
  p->adm_tol = 1.0e-10;
  p->initial_lapse = DGR_TP_LAPSE_PSIN;
  p->initial_lapse_psi_exponent = -2.0;
  p->TP_epsilon = 1.0e-6;
  p->TP_Tiny = 1.0e-15;
  p->TP_Extend_Radius = 0.0;
}

extern "C" int dgr_tp_solve(dgr_tp_t* out, const dgr_tp_params* params) {
  if (!out) return DGR_E_INVALID;
  *out = nullptr;
  int rc = check_params(params);
  if (rc) return rc;
  dgr_tp_s* tp = new dgr_tp_s();
  tp->p = *params;
  const dgr_tp_params& p = tp->p;
  tp->n1 = p.npoints_A;
  tp->n2 = p.npoints_B;
  tp->n3 = p.npoints_phi;
  tp->na = tp->n1 <= 16 ? 16 : (tp->n1 <= 32 ? 32 : 64);
  (void)hipGetDevice(&tp->device);
  tp->mp = p.par_m_plus;
  tp->mm = p.par_m_minus;
  const double b = p.par_b;
  auto punct = [&](double sx) {  // u at the punctures x = +-b
    return u_host(*tp, sx * b, 0.0, 0.0);
  };
  Grid g = make_grid(p);
  std::vector<double> V(g.N, 0.0);
  double up = 0.0, um = 0.0;
  {
    Solver S(g);
    rc = S.init();
    if (rc) { delete tp; return rc; }
    if (!p.give_bare_mass) {
      // bare masses from the target ADM masses (TwoPunctures.cpp:95-150):
      // one Newton step per bare-mass update, then solve P mp + mp mm/(4b)
      // = M+ and Q mm + mp mm/(4b) = M- (P = 1 + u+, Q = 1 + u-) for the
      // bare masses (the positive root of the quadratic)
      const double Mp = p.target_M_plus, Mm = p.target_M_minus;
      for (int outer = 0; outer < 100; ++outer) {
        set_masses(g, p, tp->mp, tp->mm);
        rc = newton(S, V, p.Newton_tol, 1, tp);
        if (rc) { delete tp; return rc; }
        tp->coef = spectral(g, V, tp->na);
        up = punct(1.0);
        um = punct(-1.0);
        tp->mp_adm = (1.0 + up) * tp->mp + tp->mp * tp->mm / (4.0 * b);
        tp->mm_adm = (1.0 + um) * tp->mm + tp->mp * tp->mm / (4.0 * b);
        if (std::fabs(Mp - tp->mp_adm) <= p.adm_tol && std::fabs(Mm - tp->mm_adm) <= p.adm_tol) break;
        const double PQ = (1.0 + up) * (1.0 + um);
        const double tmp = -4.0 * b * PQ + std::sqrt(16.0 * b * Mm * PQ + (Mp - Mm + 4.0 * b * PQ) * (Mp - Mm + 4.0 * b * PQ));
        tp->mp = (tmp + Mp - Mm) / (2.0 * (1.0 + up));
        tp->mm = (tmp - Mp + Mm) / (2.0 * (1.0 + um));
      }
    }
    set_masses(g, p, tp->mp, tp->mm);
    rc = newton(S, V, p.Newton_tol, p.Newton_maxit, tp);
    if (rc) { delete tp; return rc; }
    tp->coef = spectral(g, V, tp->na);
  }
  tp->V = V;
  up = punct(1.0);
  um = punct(-1.0);
  tp->mp_adm = (1.0 + up) * tp->mp + tp->mp * tp->mm / (4.0 * b);
  tp->mm_adm = (1.0 + um) * tp->mm + tp->mp * tp->mm / (4.0 * b);
  // total ADM mass: u ~ (A - 1) V and 1 - A ~ b / r at infinity, so
  // M = m+ + m- - 2 b V(A = 1, B = 0, phi = 0)
  tp->E = tp->mp + tp->mm - 2.0 * b * eval_host(*tp, 1.0, 0.0, 0.0);
  // angular momentum: sum of r x P over the punctures plus the spins
  const double cp[3] = {p.center_offset[0] + b, p.center_offset[1], p.center_offset[2]};
  const double cm[3] = {p.center_offset[0] - b, p.center_offset[1], p.center_offset[2]};
  for (int a = 0; a < 3; ++a) {
    const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
    tp->J[a] = cp[a1] * p.par_P_plus[a2] - cp[a2] * p.par_P_plus[a1] + cm[a1] * p.par_P_minus[a2] -
               cm[a2] * p.par_P_minus[a1] + p.par_S_plus[a] + p.par_S_minus[a];
  }
  if (hipMalloc(&tp->d_coef, sizeof(double) * tp->coef.size()) != hipSuccess ||
      hipMemcpy(tp->d_coef, tp->coef.data(), sizeof(double) * tp->coef.size(), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(tp->d_coef);
    delete tp;
    return DGR_E_NOMEM;
  }
  *out = tp;
  return DGR_OK;
}

extern "C" int dgr_tp_info(dgr_tp_t tp, dgr_tp_info_t* info) {
  if (!tp || !info) return DGR_E_INVALID;
  info->mp = tp->mp;
  info->mm = tp->mm;
  info->mp_adm = tp->mp_adm;
  info->mm_adm = tp->mm_adm;
  info->E = tp->E;
  for (int a = 0; a < 3; ++a) info->J[a] = tp->J[a];
  info->newton_steps = tp->newton_steps;
  info->last_update = tp->last_update;
  return DGR_OK;
}

extern "C" int dgr_tp_collocation(dgr_tp_t tp, double* values, uint64_t n) {
  if (!tp || !values || n < tp->V.size()) return DGR_E_INVALID;
  std::memcpy(values, tp->V.data(), sizeof(double) * tp->V.size());
  return DGR_OK;
}

extern "C" int dgr_tp_u(dgr_tp_t tp, const double* xyz, double* out, uint64_t n) {
  if (!tp || (n && (!xyz || !out))) return DGR_E_INVALID;
  for (uint64_t q = 0; q < n; ++q)
    out[q] = u_host(*tp, xyz[3 * q] - tp->p.center_offset[0], xyz[3 * q + 1] - tp->p.center_offset[1],
                    xyz[3 * q + 2] - tp->p.center_offset[2]);
  return DGR_OK;
}

extern "C" int dgr_tp_point(dgr_tp_t tp, double x, double y, double z, double* vars) {
  if (!tp || !vars) return DGR_E_INVALID;
  const double xx = x - tp->p.center_offset[0], yy = y - tp->p.center_offset[1], zz = z - tp->p.center_offset[2];
  tp_point(tp->p, tp->mp, tp->mm, xx, yy, zz, u_host(*tp, xx, yy, zz), vars);
  return DGR_OK;
}

extern "C" int dgr_tp_evaluate(dgr_tp_t tp, dgr_plan_t plan, double* const* u, void* stream) {
  if (!tp || !plan || !u) return DGR_E_INVALID;
  OutPtrs pu;
  for (int v = 0; v < 24; ++v) {
    if (!u[v]) return DGR_E_INVALID;
    pu.v[v] = u[v];
  }
  if (plan->nblocks == 0) return DGR_OK;
  const uint32_t g = std::min<uint32_t>(plan->nblocks, 2048);
  const hipStream_t st = (hipStream_t)stream;
  const bool lds = (size_t)tp->na * tp->n2 * tp->n3 <= (size_t)kLdsCoef;
  const double mp = tp->mp, mm = tp->mm;
#define DGR_TP_LAUNCH(NA, L)                                                                                      \
  hipLaunchKernelGGL((k_tp_eval<NA, L>), dim3(g), dim3(512), 0, st, plan->d_blocks, plan->nblocks, tp->p, mp, mm, \
                     tp->n2, tp->n3, (const double*)tp->d_coef, pu)
  if (tp->na == 16) {
    if (lds) DGR_TP_LAUNCH(16, true); else DGR_TP_LAUNCH(16, false);
  } else if (tp->na == 32) {
    if (lds) DGR_TP_LAUNCH(32, true); else DGR_TP_LAUNCH(32, false);
  } else {
    if (lds) DGR_TP_LAUNCH(64, true); else DGR_TP_LAUNCH(64, false);
  }
#undef DGR_TP_LAUNCH
  DGR_HIP(hipGetLastError());
  return DGR_OK;
}

extern "C" int dgr_tp_destroy(dgr_tp_t tp) {
  if (!tp) return DGR_E_INVALID;
  (void)hipFree(tp->d_coef);
  delete tp;
  return DGR_OK;
}

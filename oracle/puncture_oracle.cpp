/*
 * CPU ORACLE -- test infrastructure only (see bssn_oracle.h).
 *
 * Restatement of punctureData (BSSN_GR/src/grUtils.cpp:398-702), the
 * reference's puncture initial data, written from its formulas, not copied:
 * Brill-Lindquist conformal factor, the first-order spin / boost /
 * spin-boost corrections of the conformal factor (arXiv:0711.1165 as the
 * reference implements them, grUtils.cpp:478-606), alpha = psi^-2 and
 * chi = psi^-4 floored at CHI_FLOOR, the Bowen-York extrinsic curvature
 * (grUtils.cpp:644-700), gt = delta, everything else zero.
 *
 * grUtils.cpp is not buildable here (it includes the mesh and MPI headers),
 * so this restatement is pinned by closed forms (tests/test_oracle_puncture.py):
 * Brill-Lindquist values for non-spinning, non-boosted holes; At trace-free
 * and divergence-free in the flat conformal metric (the momentum constraint
 * Bowen-York data satisfy); the sign and normalisation of At against the
 * textbook Bowen-York momentum and spin terms; the floors.  The conformal-
 * factor corrections for spin and boost are checked only for their limits
 * (zero spin / momentum, decay far from the hole): parity partially pinned.
 * Compiled with -ffp-contract=off.
 */
#include <cmath>

extern "C" {

/* hole: mass, pos[3], p[3], spin, spin_theta, spin_phi (10 doubles) */
void orc_puncture_point(const double* xyz, const double* holes, double chi_floor, double* out) {
  const double x = xyz[0], y = xyz[1], z = xyz[2];
  double r[2], n[2][3], S[2][3], P[2][3], m[2], ptot[2], spin[2];
  double psi = 1.0;
  for (int q = 0; q < 2; ++q) {
    const double* h = holes + 10 * q;
    m[q] = h[0];
    const double dx = x - h[1], dy = y - h[2], dz = z - h[3];
    r[q] = std::sqrt(dx * dx + dy * dy + dz * dz);
    n[q][0] = dx / r[q];
    n[q][1] = dy / r[q];
    n[q][2] = dz / r[q];
    for (int d = 0; d < 3; ++d) P[q][d] = h[4 + d];
    ptot[q] = std::sqrt(P[q][0] * P[q][0] + P[q][1] * P[q][1] + P[q][2] * P[q][2]);
    spin[q] = h[7];
    S[q][0] = spin[q] * std::sin(h[8]) * std::cos(h[9]);
    S[q][1] = spin[q] * std::sin(h[8]) * std::sin(h[9]);
    S[q][2] = spin[q] * std::cos(h[8]);
    psi += m[q] / (2.0 * r[q]);
  }
  double corr = 0.0;
  for (int q = 0; q < 2; ++q) {
    const double R = 2.0 * r[q] / m[q];
    const double l = 1.0 / (1.0 + R);
    if (std::fabs(spin[q]) > 1.e-6) {
      const double J = 4.0 * spin[q] / (m[q] * m[q]);
      const double u0 = (l + l * l + std::pow(l, 3) - 4.0 * std::pow(l, 4) + 2.0 * std::pow(l, 5)) / 40.0;
      const double u2 = -std::pow(l, 5) / 20.0;
      const double mu = (n[q][0] * S[q][0] + n[q][1] * S[q][1] + n[q][2] * S[q][2]) / std::fabs(spin[q]);
      corr += J * J * (u0 + u2 * R * R * (3.0 * mu * mu - 1.0) / 2.0);
    }
    if (ptot[q] > 1.e-6) {
      const double Pa = 2.0 * ptot[q] / m[q];
      const double u0 = (l - 2.0 * l * l + 2.0 * std::pow(l, 3) - std::pow(l, 4) + 0.2 * std::pow(l, 5)) * 5.0 / 32.0;
      const double u2 = (15.0 * l + 132.0 * l * l + 53.0 * std::pow(l, 3) + 96.0 * std::pow(l, 4) +
                         82.0 * std::pow(l, 5) + 84.0 / R * (std::pow(l, 5) + std::log(l) / R)) /
                        (80.0 * R);
      const double mu = (n[q][0] * P[q][0] + n[q][1] * P[q][1] + n[q][2] * P[q][2]) / ptot[q];
      corr += Pa * Pa * (u0 + u2 * (3.0 * mu * mu - 1.0) / 2.0);
    }
    if (ptot[q] > 1.e-6 && std::fabs(spin[q]) > 1.e-6) {
      double c[3] = {P[q][1] * S[q][2] - P[q][2] * S[q][1], P[q][2] * S[q][0] - P[q][0] * S[q][2],
                     P[q][0] * S[q][1] - P[q][1] * S[q][0]};
      const double v1 = (c[0] * n[q][0] + c[1] * n[q][1] + c[2] * n[q][2]) * 16.0 / std::pow(m[q], 4) * r[q];
      corr += v1 * (1.0 + 5.0 * R + 10.0 * R * R) * std::pow(l, 5) / 80.0;
    }
  }
  const double pu = psi + corr;
  for (int v = 0; v < 24; ++v) out[v] = 0.0;
  out[0] = std::fmax(std::pow(pu, -2.0), chi_floor);
  out[1] = std::fmax(std::pow(pu, -4.0), chi_floor);
  out[12] = out[15] = out[17] = 1.0;
  static const int I[6] = {0, 0, 0, 1, 1, 2}, Jx[6] = {0, 1, 2, 1, 2, 2};
  static const double eps[3][3][3] = {{{0, 0, 0}, {0, 0, 1}, {0, -1, 0}},
                                      {{0, 0, -1}, {0, 0, 0}, {1, 0, 0}},
                                      {{0, 1, 0}, {-1, 0, 0}, {0, 0, 0}}};
  for (int p = 0; p < 6; ++p) {
    const int a = I[p], b = Jx[p];
    double A = 0.0;
    for (int q = 0; q < 2; ++q) {
      double sp = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l)
          sp += eps[a][k][l] * S[q][k] * n[q][l] * n[q][b] + eps[b][k][l] * S[q][k] * n[q][l] * n[q][a];
      double pn = 0.0;
      for (int k = 0; k < 3; ++k) pn += P[q][k] * n[q][k];
      const double mp = P[q][a] * n[q][b] + P[q][b] * n[q][a] + pn * (n[q][a] * n[q][b] - (a == b ? 1.0 : 0.0));
      A += 3.0 / (std::pow(pu, 6) * std::pow(r[q], 3)) * (sp + 0.5 * r[q] * mp);
    }
    out[18 + p] = A;
  }
}

/* n points: xyz (3, n) coordinate-major, out (24, n) variable-major */
void orc_puncture_points(long n, const double* xyz, const double* holes, double chi_floor, double* out) {
  for (long i = 0; i < n; ++i) {
    const double p[3] = {xyz[i], xyz[n + i], xyz[2 * n + i]};
    double o[24];
    orc_puncture_point(p, holes, chi_floor, o);
    for (int v = 0; v < 24; ++v) out[(size_t)v * n + i] = o[v];
  }
}

}  // extern "C"

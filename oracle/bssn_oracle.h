/*
 * CPU ORACLE -- test infrastructure only.
 *
 * A plain C++ restatement of Dendro-5's BSSN right-hand side on unzipped,
 * ghost-padded blocks (BSSN_GR/src/rhs.cpp:86-392, BSSN_GR/src/derivs.cpp,
 * BSSN_GR/src/bssn_constraints.cpp:12-172, BSSN_GR/src/rkBSSN.cpp:895-1011),
 * of its constraint monitor (BSSN_GR/src/physcon.cpp) and of its puncture
 * initial data (BSSN_GR/src/grUtils.cpp:398-702).
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline; the product
 * path (libdgr_bssn.so) never calls into it.
 *
 * How each part is pinned (see the header comment of bssn_oracle.cpp):
 *   - point equations: golden vectors written by the reference's own
 *     generated equations (oracle/_ref, built from
 *     BSSN_GR/src/bssneqs_eta_const_standard_gauge.cpp);
 *   - stencils, closures, KO, Sommerfeld terms, constraint projection and
 *     RK4 combine: closed forms (polynomial exactness, known differences of
 *     monomials, analytic boundary data) and invariants in tests/.
 */
#ifndef DGR_BSSN_ORACLE_H
#define DGR_BSSN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Face bits, same numbering as OCT_DIR_* (include/dendro.h:84-89). */
enum { ORC_LEFT = 0, ORC_RIGHT = 1, ORC_DOWN = 2, ORC_UP = 3, ORC_BACK = 4, ORC_FRONT = 5 };

typedef struct {
  double lambda[4];   /* BSSN_LAMBDA   (parameters.cpp:87) */
  double lambda_f[2]; /* BSSN_LAMBDA_F (parameters.cpp:88) */
  double eta_const;   /* ETA_CONST */
  double eta_r0;      /* ETA_R0 */
  double eta_damping_exp; /* ETA_DAMPING_EXP */
  double ko_sigma;    /* KO_DISS_SIGMA */
  int deriv_order;    /* 42: deriv42 (BSSN_GR/include/rhs.h:23-37); 64: deriv64 first/second/
                         advective with ko_deriv42 (QuadGrav/include/derivs.h:80-96) */
  int pad;
} orc_params;

/* Single-operator restatements (derivs.cpp).  axis: 0=x, 1=y, 2=z.
 * kind: 0 first (deriv42_*), 1 second (deriv42_xx/yy/zz),
 *       2 advective (deriv42adv_*, needs beta), 3 Kreiss-Oliger (ko_deriv42_*),
 *       4 first (deriv64_*), 5 second (deriv64_xx/...), 6 advective-6 (deriv64adv_*).
 * Writes only the points the reference loop writes. */
void orc_deriv(int kind, int axis, double* Du, const double* u, double h,
               const unsigned* sz, unsigned bflag, const double* beta);

/* One block: restates bssnrhs(rhs, u, offset, pmin, pmax, sz, bflag). */
void orc_bssnrhs(double* const* rhs, const double* const* u, uint64_t offset,
                 const double* pmin, const double* pmax, const unsigned* sz,
                 unsigned bflag, const orc_params* p);

/* Many blocks with an OpenMP parallel for schedule(dynamic) loop
 * (rhsTestBed.cpp:181-215).  nthreads<=0 keeps the OpenMP default. */
void orc_bssnrhs_blocks(double* const* rhs, const double* const* u, uint32_t nblocks,
                        const uint64_t* offsets, const unsigned* sz3,
                        const unsigned* bflags, const double* pmin3, const double* pmax3,
                        const orc_params* p, int nthreads);

/* Point equations only: in[NIN*n] (input-major, bssn_point_io.h order),
 * eta[n], out[24*n]. */
void orc_point_eqs(long n, const double* in, const double* eta, const orc_params* p, double* out);

/* enforce_bssn_constraints on n nodes of 24 arrays (bssn_constraints.cpp:12-172).
 * Returns 0, or 1 if a node had det<0 / |det-1|>1e-6 / |tr A|>1e-6 (the
 * reference exits there). */
int orc_enforce_constraints(double* const* u, long n, double chi_floor);

/* u_out = u_prev + c*dt*k (rkBSSN.cpp:895-905), elementwise over n. */
void orc_rk4_stage(double* u_out, const double* u_prev, const double* k, double cdt, long n);
/* u = u_prev + sum_s C[s]*dt*k_s, s = 0..3 in order (rkBSSN.cpp:996-1011). */
void orc_rk4_final(double* u, const double* u_prev, const double* const* k4, double dt, long n);

int orc_num_point_inputs(void);

/* physical_constraints on one block (physcon_oracle.cpp): con[6] = ham,
 * mom0..2, psi4_real, psi4_img on interior points. */
void orc_physcon(double* const* con, const double* const* u, uint64_t offset, const double* pmin,
                 const double* pmax, const unsigned* sz, unsigned bflag);
/* Constraint point equations: in[113*n] (physcon_point_io.h order), out[6*n]. */
void orc_physcon_point(long n, const double* in, double* out);
int orc_num_physcon_inputs(void);

/* punctureData at one point / n points (puncture_oracle.cpp); holes: 2 x
 * (mass, pos[3], p[3], spin, spin_theta, spin_phi). */
void orc_puncture_point(const double* xyz, const double* holes, double chi_floor, double* out);
void orc_puncture_points(long n, const double* xyz, const double* holes, double chi_floor, double* out);

#ifdef __cplusplus
}
#endif
#endif

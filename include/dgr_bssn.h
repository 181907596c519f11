/*
 * dgr_bssn.h -- C ABI of the MI355X BSSN right-hand-side library
 * (libdgr_bssn.so).
 *
 * Drop-in for Dendro-5's BSSN RHS on unzipped (ghost-padded) octree blocks.
 * Every entry point takes plain pointers and sizes; device arrays are HIP
 * device pointers and `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  All calls return 0 on success or a negative DGR_E_*
 * code; dgr_error_string() names it.  Nothing here calls exit().
 *
 * Data layout (same as the reference): 24 variables, each a flat double
 * array of the unzipped DOF; block b occupies sz[0]*sz[1]*sz[2] consecutive
 * doubles at `offset`, x fastest, IDX(i,j,k) = i + sz0*(j + sz1*k)
 * (BSSN_GR/include/rhs.h:21).  Variables are in VAR order
 * (BSSN_GR/include/grDef.h:31): alpha, chi, K, Gt0-2, beta0-2, B0-2,
 * gt0-5, At0-5.  Only interior points [3, sz-3)^3 of the right-hand side
 * are written; padding is left untouched (BSSN_GR/src/rhs.cpp:210-216).
 */
#ifndef DGR_BSSN_H
#define DGR_BSSN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGR_NUM_VARS 24

enum {
  DGR_OK = 0,
  DGR_E_INVALID = -1,   /* bad argument (null pointer, size out of range) */
  DGR_E_HIP = -2,       /* a HIP runtime call failed */
  DGR_E_NOMEM = -3,     /* device allocation failed */
  DGR_E_UNSUPPORTED = -4,
  DGR_E_CONSTRAINT = -5 /* constraint projection met det<0, |det-1|>1e-6 or |tr A|>1e-6 */
};

/* Block descriptor: the arguments bssnrhs takes per block
 * (BSSN_GR/include/rhs.h:47-50): offset, sz, bflag, ptmin/ptmax of the
 * padded block (ptmin = corner - 3 dx, ptmax = corner + 3 dx,
 * BSSN_GR/src/rhs.cpp:51-62).  bflag uses the OCT_DIR_* bit numbering
 * (include/dendro.h:84-89): bit0 -x, bit1 +x, bit2 -y, bit3 +y, bit4 -z,
 * bit5 +z. */
typedef struct dgr_block {
  uint64_t offset;
  uint32_t sz[3];
  uint32_t bflag;
  double pmin[3];
  double pmax[3];
  uint32_t group;    /* independent block sets (e.g. one per served job): ghost
                        refill only pairs blocks of the same group, and RK4
                        time steps are set per group; 0 for a single mesh */
  uint32_t reserved;
} dgr_block;

/* BSSN parameters, passed by value instead of the reference's bssn::
 * globals (BSSN_GR/include/parameters.h:191-222) or its GPU struct
 * cuda::BSSNComputeParams (BSSN_GR/cuda_gr/include/params_cu.h:93-108). */
typedef struct dgr_bssn_params {
  double lambda[4];       /* BSSN_LAMBDA */
  double lambda_f[2];     /* BSSN_LAMBDA_F */
  double eta_const;       /* ETA_CONST */
  double eta_r0;          /* ETA_R0 */
  double eta_damping_exp; /* ETA_DAMPING_EXP */
  double ko_sigma;        /* KO_DISS_SIGMA */
  double chi_floor;       /* CHI_FLOOR (constraint projection) */
  uint32_t deriv_order;   /* 42: deriv42 (the reference default); 64: deriv64 first / second /
                             advective with ko_deriv42 (QuadGrav/include/derivs.h:80-96) */
  uint32_t reserved;
} dgr_bssn_params;

/* Opaque plan: the block table and the tile work table on the device. */
typedef struct dgr_plan_s* dgr_plan_t;

/* How a plan cuts blocks into tiles and orders them for the persistent
 * kernel; chosen per workload by the autotuner (pydgr.autotune).  No
 * counterpart in the reference, whose GPU path launches a fixed thread
 * block per unzipped block (BSSN_GR/cuda_gr/src/rhs_cuda.cu:35-143).
 * Size classes by the largest padded extent of a block:
 *   0: <= 13, 1: <= 19, 2: <= 25, 3: larger. */
#define DGR_SIZE_CLASSES 4
enum {
  DGR_ORDER_BLOCK = 0,   /* tiles in block order */
  DGR_ORDER_XCD = 1,     /* a block's tiles on one XCD (shared L2 for the halos) */
  DGR_ORDER_LARGEST = 2, /* tiles of the largest blocks first */
  DGR_ORDER_SFC = 3      /* blocks in Morton order cut into 8 contiguous runs, one per XCD:
                            neighbouring blocks share an L2 (gathered ghosts) */
};
typedef struct dgr_schedule {
  uint32_t tile_max[DGR_SIZE_CLASSES]; /* largest tile extent per axis, 3..7 (default 7) */
  uint32_t order;                      /* DGR_ORDER_* (default DGR_ORDER_BLOCK) */
  uint32_t grid_mult;                  /* persistent grid = grid_mult x CUs; 0 = one workgroup per tile (default 1) */
} dgr_schedule;
void dgr_schedule_default(dgr_schedule* s);

const char* dgr_error_string(int code);

/* Select the device for this thread and load the kernels. */
int dgr_init(int device);

/* Build a plan for a block list.  Replaces the per-call setup of
 * cuda::computeRHS (BSSN_GR/cuda_gr/src/rhs_cuda.cu:35-143): the block table
 * is uploaded once and split into tiles of at most 7x7x7 interior points.
 * Blocks need sz >= 7 in every direction, sz >= 12 along every direction
 * if any face is flagged (one-sided closures, as in bssnrhs). */
int dgr_plan_create(dgr_plan_t* plan, const dgr_block* blocks, uint32_t nblocks,
                    const dgr_bssn_params* params);
/* Same with an explicit schedule (NULL = default); results are identical
 * for every schedule, only the speed differs. */
int dgr_plan_create_ex(dgr_plan_t* plan, const dgr_block* blocks, uint32_t nblocks,
                       const dgr_bssn_params* params, const dgr_schedule* sched);
int dgr_plan_destroy(dgr_plan_t plan);
/* Interior points covered by the plan. */
uint64_t dgr_plan_points(dgr_plan_t plan);
uint32_t dgr_plan_tiles(dgr_plan_t plan);

/* RHS of all blocks in one batched launch on `stream` (asynchronous).
 * rhs and u are arrays of 24 device pointers.  Replaces
 * bssnRHS(double**, const double**, const ot::Block*, unsigned)
 * (BSSN_GR/include/rhs.h:45) and cuda::computeRHS
 * (BSSN_GR/cuda_gr/include/rhs_cuda.cuh:36), with device-resident data. */
int dgr_bssn_rhs(dgr_plan_t plan, double* const* rhs, const double* const* u, void* stream);

/* Same result, one kernel launch per block: the per-block loop of
 * bssnRHS (BSSN_GR/src/rhs.cpp:43-70) calling bssnrhs for each block
 * (BSSN_GR/include/rhs.h:47-50).  Kept as the baseline the batched path is
 * measured against. */
int dgr_bssn_rhs_per_block(dgr_plan_t plan, double* const* rhs, const double* const* u, void* stream);

/* Block-wise RHS of the time-stepper context (BSSNCtx::rhs_blkwise,
 * BSSN_GR/include/bssnCtx.h:79-91, BSSN_GR/src/bssnCtx.cpp:186-245): the
 * blocks blk_ids[0..n) of the plan, same arrays and layout as dgr_bssn_rhs,
 * one batched launch.  Asynchronous. */
int dgr_bssn_rhs_blkwise(dgr_plan_t plan, double* const* rhs, const double* const* u, const uint32_t* blk_ids,
                         uint32_t n, void* stream);
/* Single-block RHS (BSSNCtx::rhs_blk, BSSN_GR/include/bssnCtx.h:93,
 * BSSN_GR/src/bssnCtx.cpp:247-305): block blk_id of the plan with its own
 * block-local arrays, variable v of the padded block at in + v*NN and
 * out + v*NN (NN = sz0*sz1*sz2), dof = 24.  Asynchronous. */
int dgr_bssn_rhs_blk(dgr_plan_t plan, double* out, const double* in, uint32_t dof, uint32_t blk_id,
                     void* stream);

/* One-shot form of bssnRHS: builds a temporary plan, runs, synchronises. */
int dgr_bssn_rhs_blocks(double* const* rhs, const double* const* u, const dgr_block* blocks,
                        uint32_t nblocks, const dgr_bssn_params* params, void* stream);

/* One-shot bssnRHS that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bitwise the same result as
 * dgr_bssn_rhs_blocks.  The block, tile and pointer tables go to a
 * file-scope device array of DGR_NOALLOC_SCRATCH_BYTES (about 4.1 MB) that
 * each GPU the library runs on holds until the process ends; at most
 * DGR_NOALLOC_MAX_BLOCKS blocks cut into at most DGR_NOALLOC_MAX_TILES
 * tiles of up to 7^3 interior points (a 13^3 block is one tile, a 37^3
 * block 125), else DGR_E_INVALID before any device work.  Asynchronous on
 * `stream` (the tables are copied from host memory before the call
 * returns).  Every call on a GPU shares that GPU's array: calls on one
 * stream are ordered, calls that overlap on different streams race. */
#define DGR_NOALLOC_MAX_BLOCKS 8192u
#define DGR_NOALLOC_MAX_TILES 65536u
#define DGR_NOALLOC_SCRATCH_BYTES 4063744u
int dgr_bssn_rhs_blocks_noalloc(double* const* rhs, const double* const* u, const dgr_block* blocks,
                                uint32_t nblocks, const dgr_bssn_params* params, void* stream);

/* "Unzip" of the padding between RK stages (ghost exchange + unzip,
 * BSSN_GR/src/rkBSSN.cpp:909-914, Mesh::unzip include/mesh.tcc:7436):
 * every ghost point of every block (outside [3, sz-3)^3) with a same-level
 * neighbour (equal size and spacing, corners sz-7 points apart) takes that
 * neighbour's interior value; on 2:1-balanced layouts of 13^3 blocks, ghosts
 * next to a coarser block are interpolated from its element (order-6
 * Lagrange, x then y then z) and ghosts next to finer blocks copy their
 * nodes.  Ghosts outside the domain are left untouched.  Synchronous. */
int dgr_ghost_refill(dgr_plan_t plan, double* const* u, void* stream);
/* Number of level-jump interpolation boxes and node copies the plan found
 * (0 on uniform layouts).  No counterpart in the reference (its unzip
 * recomputes these from the mesh on every call). */
int dgr_plan_level_fills(dgr_plan_t plan, uint32_t* boxes, uint64_t* copies);

/* Same-level neighbour table the ghost refill and the gathering RK4 stages
 * use, computed on the host (no device call): out[b*27 + (dx+1) + 3(dy+1) +
 * 9(dz+1)] = index of the block whose corner is sz-7 points away along each
 * axis with d != 0 (equal size, spacing and group), or -1.  Corners are
 * matched on the lattice of their class; a block off that lattice by more
 * than 1e-6 of a spacing has no neighbours and is no one's neighbour.  No
 * counterpart in the reference, whose unzip finds neighbours through the
 * mesh's element-to-element map (Mesh::unzip, include/mesh.tcc:7436). */
int dgr_block_neighbours(const dgr_block* blocks, uint32_t nblocks, int32_t* out);

/* enforce_bssn_constraints (BSSN_GR/src/bssn_constraints.cpp:12-172) on
 * every interior point: det(gt) = 1, tr(At) = 0, chi and alpha floored at
 * params.chi_floor.  Returns DGR_E_CONSTRAINT where the reference exits
 * (det < 0, |det-1| > 1e-6 after rescaling, |tr A| > 1e-6); points with
 * det < 0 are left unchanged.  Synchronous. */
int dgr_enforce_constraints(dgr_plan_t plan, double* const* u, void* stream);

/* Constraint monitor.  Outputs in VAR_CONSTRAINT order
 * (BSSN_GR/include/grDef.h:34): ham, mom0, mom1, mom2, psi4_real, psi4_img. */
#define DGR_NUM_CONSTRAINTS 6
/* physical_constraints (BSSN_GR/src/physcon.cpp:12-96, declared in
 * BSSN_GR/include/physcon.h:17-20) over every block of the plan: con[6]
 * device arrays of the same unzipped layout as u[24], written on interior
 * points only.  deriv42 stencils with face closures whatever the plan's
 * deriv_order (the reference's physcon.h:8-14).  Asynchronous on stream. */
int dgr_physcon(dgr_plan_t plan, double* const* con, const double* const* u, void* stream);
/* Masked constraint norms (computeConstraintL2Norm,
 * BSSN_GR/include/grUtils.tcc:40-88, called per constraint by
 * extractConstraints, grUtils.tcc:91-130): over the unique nodes of the
 * zipped vector (dgr_zip: each node once, hanging nodes at level jumps
 * excluded) whose mask value (the reference's EXTRACTION_VAR_ID, alpha by
 * default, parameters.cpp:157-158) is >= threshold.  out[13] (host):
 * sqrt(mean of squares) of the 6 outputs, their max |.|, and the node
 * count.  Deterministic (fixed reduction order).  Synchronous. */
/* Mesh::zip (include/mesh.tcc:3775-3822): the unique nodal values of one
 * unzipped array (device, the plan's layout) into zipped[0..n) (device),
 * n from dgr_zip_size.  A node shared by same-level blocks is taken from the
 * block that has it on its low side; a fine-block node on a side whose
 * neighbour is coarser is either a hanging node (no zipped entry) or a
 * coarse node, taken from the coarse block.  Order: blocks in plan order,
 * each block's nodes x fastest.  The node tables are built on the first call
 * (synchronous then); asynchronous on `stream` afterwards. */
int dgr_zip_size(dgr_plan_t plan, uint64_t* n);
int dgr_zip(dgr_plan_t plan, double* zipped, const double* unzipped, void* stream);
int dgr_physcon_norms(dgr_plan_t plan, const double* const* con, const double* mask, double threshold,
                      double* out, void* stream);

/* One puncture of the initial data: the fields of the reference's BH
 * (BSSN_GR/include/bh.h:58-80) that punctureData reads. */
typedef struct {
  double mass;                  /* getBHMass */
  double pos[3];                /* getBHCoordX/Y/Z */
  double p[3];                  /* getVx/Vy/Vz: linear momentum (boost) */
  double spin;                  /* getBHSpin */
  double spin_theta, spin_phi;  /* getBHSpinTheta/Phi */
} dgr_puncture;
/* punctureData (BSSN_GR/src/grUtils.cpp:398-702) at every padded point of
 * every block of the plan (x = pmin + i h): Brill-Lindquist conformal
 * factor with the reference's spin / boost corrections, alpha and chi
 * floored at params.chi_floor (the reference's CHI_FLOOR), Bowen-York At,
 * gt = delta, all else zero.  holes[2]; masses must be positive.
 * Asynchronous on stream. */
int dgr_init_puncture(dgr_plan_t plan, double* const* u, const dgr_puncture* holes, void* stream);
/* No counterpart in the reference: synthetic perturbed flat space for
 * benchmarks and served jobs.  At every padded point, alpha = chi = 1,
 * gt = delta, all else 0, plus on variable v the sum over the bumps b of
 * amp exp(-|x - c_vb|^2 / width^2); centers: device array of 24 x nbumps x 3
 * doubles (c_vb at 3 (v nbumps + b)), nbumps <= 16.  Asynchronous on stream. */
int dgr_init_bumps(dgr_plan_t plan, double* const* u, const double* centers, int nbumps, double amp, double width,
                   void* stream);

/* TwoPunctures initial data (BSSN_GR/src/TwoPunctures.cpp:24-498, solver in
 * TPNewton.cpp / TPFuncAndJacobian.cpp / TPCoordTransf.cpp / TPEquations.cpp).
 * The reference's TwoPunctures(x, y, z, vars, &mp, &mm, &mp_adm, &mm_adm, &E,
 * &J1, &J2, &J3) solves on its first call and then fills one point per call;
 * here the solve is dgr_tp_solve and the fill of every padded point of a
 * plan's blocks is dgr_tp_evaluate.  Field names are the reference's TPID
 * parameters (BSSN_GR/include/parameters.h:262-310); dgr_tp_params_default
 * gives its defaults.  Punctures sit at center_offset +- (par_b, 0, 0).
 * Always evaluated from the spectral coefficients (the reference's
 * grid_setup_method EVALUATION); no multiply_old_lapse. */
enum {
  DGR_TP_LAPSE_ANTISYMMETRIC = 0, /* (1 - m+/2r+ - m-/2r-) / (1 + m+/2r+ + m-/2r-) */
  DGR_TP_LAPSE_AVERAGED = 1,      /* (1 + antisymmetric) / 2 */
  DGR_TP_LAPSE_PSIN = 2,          /* psi_BL^initial_lapse_psi_exponent (the reference's default) */
  DGR_TP_LAPSE_BROWNSVILLE = 3,   /* 2 / (1 + psi_BL^exponent) */
  DGR_TP_LAPSE_UNIT = 4           /* 1 */
};
#define DGR_TP_MAX_UNKNOWNS 40000 /* npoints_A x npoints_B x npoints_phi (a dense N x N Jacobian in HBM) */
typedef struct {
  double par_b;                           /* punctures at x = +-par_b */
  double par_m_plus, par_m_minus;         /* bare masses (initial guesses unless give_bare_mass) */
  double target_M_plus, target_M_minus;   /* ADM masses sought when give_bare_mass == 0 */
  double par_P_plus[3], par_P_minus[3];   /* Bowen-York momenta */
  double par_S_plus[3], par_S_minus[3];   /* Bowen-York spins */
  double center_offset[3];
  int give_bare_mass;
  int npoints_A, npoints_B, npoints_phi;  /* spectral resolution: A, B <= 64 / 256, phi <= 256 */
  int Newton_maxit;
  int initial_lapse;                      /* DGR_TP_LAPSE_* */
  double Newton_tol;                      /* stop when the largest Newton update is below it */
  double adm_tol;
  double initial_lapse_psi_exponent;
  double TP_epsilon, TP_Tiny, TP_Extend_Radius;
} dgr_tp_params;
typedef struct {
  double mp, mm;            /* bare masses used */
  double mp_adm, mm_adm;    /* puncture ADM masses */
  double E;                 /* total ADM mass */
  double J[3];              /* ADM angular momentum */
  int newton_steps;         /* Newton steps taken (all bare-mass iterations) */
  double last_update;       /* largest change of the last step */
} dgr_tp_info_t;
typedef struct dgr_tp_s* dgr_tp_t;
void dgr_tp_params_default(dgr_tp_params* p);
/* The spectral solve (synchronous; uses the current device and about
 * 8 N^2 bytes of device memory while it runs). */
int dgr_tp_solve(dgr_tp_t* tp, const dgr_tp_params* p);
int dgr_tp_info(dgr_tp_t tp, dgr_tp_info_t* info);
/* The solution V (u = (A - 1) V) at the collocation points, index
 * i + npoints_A (j + npoints_B k); n >= npoints_A npoints_B npoints_phi. */
int dgr_tp_collocation(dgr_tp_t tp, double* values, uint64_t n);
/* u at n points (xyz: 3 n host doubles, absolute coordinates), on the host. */
int dgr_tp_u(dgr_tp_t tp, const double* xyz, double* out, uint64_t n);
/* The reference's per-point TwoPunctures call on the host: the 24 BSSN
 * variables at one point (absolute coordinates) into vars[24], the same
 * formulas as dgr_tp_evaluate (host libm instead of the device's). */
int dgr_tp_point(dgr_tp_t tp, double x, double y, double z, double* vars);
/* The 24 BSSN variables at every padded point of every block (x = pmin +
 * i h): chi, gt, At, K from psi^4 delta and the Bowen-York curvature, the
 * lapse of initial_lapse, beta = Gt = B = 0.  Asynchronous on stream. */
int dgr_tp_evaluate(dgr_tp_t tp, dgr_plan_t plan, double* const* u, void* stream);
int dgr_tp_destroy(dgr_tp_t tp);

/* Device-resident RK4 evolution of the 24 state arrays u (caller-owned
 * device memory, ghost layers filled).  One step replaces the RK4 branch of
 * RK_BSSN::performSingleIteration (BSSN_GR/src/rkBSSN.cpp:803-1011):
 * four RHS evaluations, the stage combines u_prev + RK4_U[s+1] dt k_s with
 * enforce_bssn_constraints, the final u_prev + sum_s RK4_C[s] dt k_s
 * (rkBSSN.h:55-58) and a same-level ghost refill after every stage.  The
 * step is captured into a hipGraph on first use (use_graph != 0) and
 * replayed.  The RHS of stages 1-4 reads ghost points that have a same-level
 * neighbour from that neighbour's interior, which gives the same values as a
 * refill after each stage.  Within one dgr_rk4_run / _step / _steps call
 * the state lives in the work memory; the call ends by writing it back into
 * u and refilling u's ghost layers (same-level, then level jumps).  Work
 * memory: dgr_rk4_work_doubles(plan) doubles. */
typedef struct dgr_rk4_s* dgr_rk4_t;
int dgr_rk4_create(dgr_rk4_t* rk, dgr_plan_t plan, double* const* u, int use_graph);
/* Same, with caller-owned work memory of at least dgr_rk4_work_doubles(plan)
 * doubles = 48 x (dof + interior points) (the state during a call and the
 * stage state: per variable the padded layout, then every block's interior
 * points compact) + 4 x 24 x 343 x tiles (k and the running sum in the
 * compact tile-major layout). */
uint64_t dgr_rk4_work_doubles(dgr_plan_t plan);
int dgr_rk4_create_with_work(dgr_rk4_t* rk, dgr_plan_t plan, double* const* u, double* work, int use_graph);
/* Copy the state into the stage buffer, stream-ordered on `stream`.  Ghost
 * points that no neighbouring block fills and that lie beyond no flagged
 * face (a concave corner of a block set that is not a full octree) are read
 * by stages 2-4 from the stage buffer, where they keep the state's values;
 * creation does this copy once, and a caller that writes a new state into an
 * existing evolver (a served job reusing an evolver) calls this after it.
 * No counterpart in the reference, whose meshes have no such ghosts. */
int dgr_rk4_reset_stage(dgr_rk4_t rk, void* stream);
/* Per-group time steps dt[g], g < ngroups (every group in the plan must be
 * covered); dgr_rk4_step(s) set the same dt for all groups. */
int dgr_rk4_set_group_dt(dgr_rk4_t rk, const double* dt, uint32_t ngroups);
int dgr_rk4_run(dgr_rk4_t rk, int nsteps, void* stream);
int dgr_rk4_step(dgr_rk4_t rk, double dt, void* stream);
int dgr_rk4_steps(dgr_rk4_t rk, double dt, int nsteps, void* stream);
/* Waits for the evolver and returns DGR_E_CONSTRAINT if any projection
 * since the last call met a condition the reference exits on. */
int dgr_rk4_status(dgr_rk4_t rk);
int dgr_rk4_destroy(dgr_rk4_t rk);

/* Explicit Runge-Kutta time stepper with a Butcher tableau: the reference's
 * ETS (ODE/include/ets.h:337-445 set_ets_coefficients, :469-545 evolve)
 * driving the BSSN context (BSSN_GR/src/bssnCtx.cpp:151-181 rhs).  Per stage
 * s: tmp = u + sum_{p<s} (a_sp dt) k_p on interior points, constraint
 * projection of tmp (as BSSNCtx::rhs does on its input), ghost fill of tmp,
 * k_s = RHS(tmp); then u += (b_k dt) k_k for k in order, and u's ghost
 * layers are refilled.  Stage buffers are allocated at creation (24 x
 * (1 + stages) padded arrays).  Types follow ETSType (ODE/include/ts.h:26);
 * RK5 is disabled in the reference and here. */
typedef struct dgr_ets_s* dgr_ets_t;
enum { DGR_ETS_RK3 = 0, DGR_ETS_RK4 = 1 };
int dgr_ets_create(dgr_ets_t* ets, dgr_plan_t plan, double* const* u, int type);
/* set_ets_coefficients(aij, bi, ci, num_stages): aij row-major stages x
 * stages, strictly lower triangular (explicit), at most 8 stages. */
int dgr_ets_create_tableau(dgr_ets_t* ets, dgr_plan_t plan, double* const* u, const double* aij,
                           const double* bi, const double* ci, uint32_t stages);
/* One step of size dt (ETS::evolve), asynchronous on stream. */
int dgr_ets_evolve(dgr_ets_t ets, double dt, void* stream);
/* Time advanced so far (sum of the dt of all evolve calls). */
double dgr_ets_time(dgr_ets_t ets);
/* DGR_E_CONSTRAINT if a stage projection met a point where the reference
 * exits since the last call; synchronises stream. */
int dgr_ets_status(dgr_ets_t ets, void* stream);
int dgr_ets_destroy(dgr_ets_t ets);

#ifdef __cplusplus
}
#endif
#endif

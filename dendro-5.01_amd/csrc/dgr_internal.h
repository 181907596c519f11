/*
 * Internal declarations shared by the BSSN RHS kernels (dgr_bssn.hip) and
 * the RK4 / ghost-refill / constraint kernels (dgr_rk4.hip).
 */
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "../../include/dgr_bssn.h"

namespace dgr_internal {

// Per-axis stencil constants, computed on the host exactly as derivs.cpp
// computes them (IEEE double, no contraction).
struct AxisC {
  double ih2, ih12, isq, isq12, pre, smr1, smr2, smr3, ih60, isq180;
  int e;
  int lf, hf;
  int pad;
};

struct DevBlock {
  uint64_t offset;
  uint32_t sz[3];
  uint32_t bflag;
  double pmin[3];
  double h[3];
  AxisC ax[3];
  uint32_t group;
  uint32_t cbase;  // first slot of the block's interior points in a compact region
};

struct DevTile {
  uint32_t block;
  uint16_t i0, j0, k0;  // first interior point of the tile (block coords)
  uint8_t nx, ny, nz;
  uint8_t pad;  // boundary tiles: thread map (0 default, else 1 + face map index)
};

struct VarPtrs {
  const double* v[24];
};
struct OutPtrs {
  double* v[24];
};
// The 48 array pointers live in device memory and are loaded where used,
// which keeps them out of the scalar register file.
struct IOPtrs {
  VarPtrs u;
  OutPtrs r;
};

struct DevParams {
  double lambda[4];
  double lambda_f[2];
  double eta_const, eta_r0, eta_damping_exp, ko_sigma;
};

// Ghost fill across a level jump, fine side: the ghost box of a block (one
// of its 26 directions) lies inside one element of a block twice as coarse
// and is interpolated from that element's 7^3 nodes.
struct FillItem {
  uint64_t dst;        // padded index of the box's first point in the fine block
  uint64_t src;        // padded index of the coarse block's node (3,3,3)
  uint32_t dsz[3];     // fine block sizes (strides of dst)
  uint32_t ssz[3];     // coarse block sizes (strides of src)
  uint8_t n[3];        // box extent per axis (3 or sz-6)
  uint8_t s0[3];       // position of the box's first point in the coarse element, in half node spacings
  uint8_t pad[2];
  uint64_t csrc;       // src in a stage buffer: its slot in the compact interior region
};

// Ghost fill across a level jump, coarse side: every ghost point coincides
// with a node of a finer block and copies it.
struct FillCopy {
  uint64_t dst, src;
  uint64_t csrc;  // src in a stage buffer: its slot in the compact interior region
};


// Same-level neighbour of a block in direction (dx,dy,dz) in {-1,0,1}^3,
// slot (dx+1) + 3 (dy+1) + 9 (dz+1); -1 when there is none (domain
// boundary, or a neighbour of another size or spacing).
constexpr int NNBR = 27;

int hip_fail(hipError_t e, const char* what);

#define DGR_HIP(call)                                 \
  do {                                                \
    hipError_t _e = (call);                           \
    if (_e != hipSuccess) return dgr_internal::hip_fail(_e, #call); \
  } while (0)

}  // namespace dgr_internal

struct dgr_plan_s {
  int device = 0;
  uint32_t nblocks = 0;
  uint32_t ntiles = 0;
  uint64_t points = 0;
  dgr_internal::DevBlock* d_blocks = nullptr;
  dgr_internal::DevTile* d_tiles = nullptr;      // scheduled order (batched launches)
  dgr_internal::DevTile* d_tiles_blk = nullptr;  // block order (per-block launches)
  uint32_t* d_tile_slot = nullptr;        // block-order index of each scheduled tile
  uint32_t ntiles_interior = 0;           // tiles [0, n) belong to blocks with bflag 0
  std::vector<uint32_t> block_tile_begin;  // first tile of block b in d_tiles_blk
  std::vector<uint32_t> block_tile_count;
  uint32_t grid_mult = 1;                 // persistent grid = grid_mult x CUs; 0: one workgroup per tile
  dgr_internal::DevParams params{};
  int grid = 0;                          // persistent grid size
  dgr_internal::IOPtrs* d_io = nullptr;  // device copy of the 48 array pointers
  dgr_internal::IOPtrs h_io{};           // last uploaded pointers
  bool io_valid = false;
  int32_t* d_nbr = nullptr;              // NNBR neighbour slots per block
  long long* d_delta = nullptr;          // NNBR gather distances per block (gathering RHS)
  // Stage buffers of the RK4 evolver: per variable the padded layout
  // (dof_end points; only its ghosts are read) followed by a compact region
  // of every block's interior points, block after block (block b's from
  // DevBlock::cbase, x fastest).  d_cdelta: per block and direction, the
  // distance from the compact index a point would have in its own block
  // to its source in such a buffer (the neighbour's compact slot, the
  // block's own for direction 13), or 0 where the source is the block's
  // own padding.
  uint64_t ncomp = 0;
  long long* d_cdelta = nullptr;
  std::vector<int32_t> nbr;              // host copy of the neighbour table
  uint64_t dof_end = 0;                  // one past the last padded point of any block
  double chi_floor = 0.1;                // CHI_FLOOR of the constraint projection
  uint32_t ngroups = 1;                  // 1 + largest block group
  int order = 42;                        // stencil set: 42 (deriv42) or 64 (deriv64 + ko_deriv42)
  // level jumps (adaptive layouts): interpolation boxes and copies
  dgr_internal::FillItem* d_fill_items = nullptr;  // sorted by coarse element
  uint32_t nfill_items = 0;
  uint32_t* d_fill_groups = nullptr;  // first item of each group (one coarse element), + nfill_items
  uint32_t nfill_groups = 0;
  dgr_internal::FillCopy* d_fill_copies = nullptr;
  uint64_t nfill_copies = 0;
  // constraint monitor: pointer table, norm partials (allocated on first use)
  dgr_internal::IOPtrs* d_io_con = nullptr;
  dgr_internal::IOPtrs h_io_con{};
  bool io_con_valid = false;
  double* d_norm = nullptr;  // 13 x ntiles partials + 13 results
  // block-wise RHS: host copy of the block-order tiles, a device list of a
  // block subset's tiles and a pointer table (allocated on first use)
  std::vector<dgr_internal::DevTile> h_tiles_blk;
  dgr_internal::DevTile* d_tiles_sub = nullptr;
  dgr_internal::IOPtrs* d_io_blk = nullptr;
  // batched RHS: the boundary-tile kernel runs on `aux`, forked from and
  // joined back to the caller's stream, so its workgroups fill the CUs that
  // the interior kernel's tail leaves idle (default priority)
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // flag word of dgr_enforce_constraints (allocated with the plan)
  unsigned* d_flag = nullptr;
  // zip to unique nodal values (built on first use): host block table, per
  // block the bit set of the 27 directions whose neighbour is coarser (from
  // the level-jump fills), the owner flag of every padded point and the
  // padded index of every unique node in zipped order
  std::vector<dgr_internal::DevBlock> h_blocks;
  std::vector<uint32_t> coarse_dirs;
  bool zip_built = false;
  uint64_t nzip = 0;
  uint8_t* d_own = nullptr;
  uint64_t* d_zidx = nullptr;
};

namespace dgr_internal {
// Launch the RHS kernels of `plan` with a device pointer table.  With
// gather, ghost points that have a same-level neighbour are read from that
// neighbour's interior (the RK4 stages), so no refill is needed before, and
// the right-hand side is written in the compact tile-major layout of the
// stage kernels: variable v of interior point (a, b, c) of block-order tile
// w at rhs[v][w * 343 + a + 7 (b + 7 c)].
// With stage, the inputs are stage buffers (interior points read from the
// compact region; gather only).
int launch_rhs(dgr_plan_t plan, const IOPtrs* d_io, hipStream_t st, bool gather, bool stage = false);

// Find the level-jump ghost fills of a plan's blocks and upload them
// (plan->d_fill_*); blocks without a same-level neighbour in a direction are
// looked up one level coarser and one level finer.
int build_level_fill(dgr_plan_t plan, const dgr_block* blocks, const std::vector<DevBlock>& hb);
// Fill the level-jump ghosts of the 24 arrays u (asynchronous on st); with
// stage, u are stage buffers whose interior points are read from the
// compact region.
int launch_level_fill(dgr_plan_t plan, const OutPtrs& u, hipStream_t st, bool stage = false);
}  // namespace dgr_internal

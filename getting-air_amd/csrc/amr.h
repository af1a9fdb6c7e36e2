// Launch interface of the block-pool (AMR) kernels, amr_kernels.hip.
//
// Level 0 is a dense solve's lattice (its block layout, dense.h lat_elem);
// each level m >= 1 is a pool of tiles.  Level m is cut into blocks of RB^3
// cells of its PARENT level m - 1; a refined block owns a slot holding FB^3
// level-m cells, FB = 2 RB + 2 GHOST (interior [GHOST, GHOST + 2 RB) plus a
// two-cell ring).  A pool is plane-major:
// g[i * pool_plane + (slot * FB + fz) * FB * FB + fy * FB + fx], with
// pool_plane = max_slots * FB^3, holding the shifted populations g = f - w.
// A level's PoolGeom describes its parent grid: nx, ny, nz are the parent
// level's global cell counts (L0's times 2^(m-1)), so its blocks, cell
// classes and surface lists are those of level m over level m - 1, and only
// how a parent cell's data is addressed differs between the dense parent
// (m = 1) and a pool parent (m >= 2, ParentRef).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "collide.h"
#include "dense.h"

namespace lbm {

constexpr int kGhost = 2;

// A pinned sphere solved on the fine level alone, in L0 cell units; the
// refined region covers every L0 cell within 1 of it.
struct PoolBody {
  int32_t on;
  float cx, cy, cz, r;
};

struct PoolGeom {
  int32_t nx, ny, nz;        // parent-level cells (L0's at level 1)
  int32_t rb, fb;            // block edge (parent cells), tile edge (this level's cells)
  int32_t nbx, nby, nbz;     // blocks per axis
  uint32_t ncells;           // parent-level cells
  uint32_t pool_plane;       // max_slots * fb^3
  int32_t max_slots;
  const int32_t* block_slot;     // per block: slot or -1
  const int32_t* slot_to_block;  // per slot: block or -1
  const uint8_t* ccls;           // per parent cell: kCovered | kNear | kSeam
  const int32_t* tile_table;     // kTileWords per active tile (k_tile_table)
  PoolBody body;
  // Position of this level's cells in L0 units: p = u * l0_scale + l0_offset
  // with u the parent-unit position origin - 1/4 + (j - GHOST)/2;
  // l0_scale = 2^-(m-1), l0_offset = -(1 - 2^-(m-1))/2 (cell-centred
  // refinement, common_d3_amr_step1.wgsl:64-84).  1 and 0 at level 1.
  float l0_scale, l0_offset;
  float dx2;                     // the body force's weight dx^2 = 4^-m
  // The next finer level's block map (null: none): an interior cell under
  // one of its tiles is solved there, and this level's step leaves it alone
  // (coveredByChild, common_d3_amr_step1.wgsl:107-143).
  const int32_t* child_block_slot;
  int32_t cnbx, cnby;            // the child level's blocks per axis (2 nx / rb ...)
};

// Where a parent cell's populations and moments live.  The dense parent (L0)
// holds them in the lattice's block layout and float4 moments; a pool parent
// (level m - 1 >= 1) in its tiles, planar, found through its block map: the
// tile that owns the cell, at interior index cell - 2 RB block + GHOST
// (parentIndex, common_d3_parent_pool.wgsl:37-59).
struct ParentRef {
  float* g;                   // the parent's populations at the coupling's time level
  float* mac;                 // its moments
  const int32_t* block_slot;  // pool parent: its block map; null: dense parent
  int32_t nbx, nby, nbz;      // pool parent: its blocks per axis
  uint32_t plane;             // pool parent: its pool plane
};

// Per active tile (index into the active list), rebuilt with the cell
// classes: words 0-26 the slots of the 27 blocks around it, word (ox + 1) +
// 3 (oy + 1) + 9 (oz + 1) for the block at offset (ox, oy, oz) (-1 when
// unrefined; word 13 is the tile's own slot), word 27 a mask of the refined
// ones, words 28-30 the tile's block coordinates.
constexpr int kTileWords = 32;

// L0 cell classes, rebuilt by launch_cell_class after every topology change.
constexpr uint8_t kCovered = 1;  // the cell's block holds a tile
constexpr uint8_t kNear = 2;     // unrefined, with a covered lattice neighbour
constexpr uint8_t kSeam = 4;     // covered, with an unrefined lattice neighbour

// Block edges the pool kernels are compiled for.
constexpr bool supported_rb(int rb) { return rb == 4 || rb == 8 || rb == 16; }

// Manager inputs and outputs (one workgroup; see k_manage).
struct ManageArgs {
  int32_t nblocks, max_slots;
  const float* crit;         // per block criterion value
  float refine_at;           // an unrefined block refines when crit >= refine_at
  float keep_at;             // a refined block stays while crit >= keep_at
  const uint8_t* geom;       // per block, or null: bit 0 the body's geometry wants it (always
                             // refined), bit 1 the field criterion may not refine it
  int32_t* block_slot;
  int32_t* slot_to_block;
  uint8_t* want;             // out: per block decision
  int32_t* dying;            // out: [count, (block, slot) pairs ...] ascending block id
  int32_t* fresh;            // out: [count, (block, slot) pairs ...] ascending block id
  int32_t* active;           // out: [count, slot ...] ascending
  int32_t* free_list;        // scratch, max_slots
  int32_t* status;           // [0] wanted blocks left without a slot (sticky sum); [1] with
                             // `decided`, wanted blocks whose geom bit 1 is set (sticky sum)
  int32_t decided;           // 1: want[] already holds the decision (the manager at depth);
                             // crit, refine_at, keep_at unused, geom only counted into status[1]
};

// The manager at depth decides every level's want before any level
// allocates (common_d3_manage.wgsl decide, main-3d.js:3086-3107).
struct DecideArgs {
  int32_t nbx, nby, nbz;       // this level's blocks per axis
  int32_t shift;               // m - 1: a level-m block's level-1 block is (b >> shift)
  int32_t cnbx, cnby;          // the criterion's grid (level-1 blocks per axis)
  const float* crit;           // per level-1 block, or null: no field term
  float refine_at, keep_at;    // this level's rung of the ladder
  const uint8_t* geom;         // per block, or null: bit 0 the body wants it, bit 1 the field may not
  const int32_t* block_slot;   // held blocks are judged against keep_at
  uint8_t* want;               // out
};

// Work lists of the surface passes, rebuilt with the cell classes after every
// topology change (entries in no particular order: each entry's pass writes
// only its own cell, so the order cannot change a result).
struct SurfacePlan {
  uint32_t* seam;   // covered L0 cells with an unrefined lattice neighbour
  uint32_t* near;   // unrefined L0 cells with a covered lattice neighbour
  uint4* ring;      // every ring cell of each active tile (ring_per_tile entries per tile, in
                    // class-major order): pool cell (slot * FB^3 + local) or kAbsent over a
                    // refined L0 cell, L0 cell, explode decisions, L0 coordinates (k_ring_plan)
  uint32_t* flux;   // unrefined L0 cells with a covered cell among their 26 neighbours
                    // (the seam flux correction's cells); null without it
  uint32_t* count;  // [seam, near, (unused), flux]
  uint32_t cap_seam, cap_near, cap_ring, cap_flux;
};
// Ring cells per tile, (2 rb + 4)^dim - (2 rb)^dim: the ring list holds this
// many entries for every active tile.
uint32_t ring_per_tile(int rb, int dim);
// Capacity of the lists for a pool of max_slots tiles.
void plan_capacity(int rb, int max_slots, uint32_t ncells, int dim, uint32_t* seam, uint32_t* near, uint32_t* ring);

// Rebuilds ccls and the plan lists (count must be zeroed first).
// parent: which parent cells exist (a pool parent's cells outside its tiles
// are never listed); the classes cover the whole parent grid.
hipError_t launch_cell_class(int q, uint8_t* ccls, const PoolGeom& g, const SurfacePlan& plan, const int32_t* active,
                             int nactive, const ParentRef& parent, hipStream_t s);
// mac_pool null: the substep's macroscopic field is not written.
// collide_ring: ring cells collide as well (the interp interface).
// sub: 0 or 1, the first or second substep of the explode interface, whose
// cells nothing reads afterwards load nothing; -1 steps every cell.
hipError_t launch_pool_step(int q, const float* gin, float* gout, float* mac_pool, const PoolGeom& g,
                            const CollideParams& fine, int nactive, int collide_ring, int sub, hipStream_t s);
// Momentum exchange on the fine-level body from the pool state at the start
// of the macro-step: one FP64 [fx, fy, fz, tx, ty, tz] partial per thread
// block (pool_force_blocks of them), weighted dx^2 = 1/4.
uint32_t pool_force_blocks(int rb, int nactive);
hipError_t launch_pool_force(int q, const float* gpool, const PoolGeom& g, const int32_t* active, int nactive,
                             double* partials, hipStream_t s);
// Seam flux of the fine solver over one substep, from the pool buffer that
// substep streams: [dm, dpx, dpy, dpz] / 8 per flux cell, written
// (accumulate 0) or added (1).
hipError_t launch_flux(int q, const float* gpool, float* acc, const PoolGeom& g, const SurfacePlan& plan,
                       uint32_t nflux, int accumulate, hipStream_t s);
// Flux correction of the unrefined flux cells of the coarse state at t + 1
// (gc_t1, mac) from the coarse state at t and the accumulated fine flux.
hipError_t launch_reflux(int q, const float* gc_t, float* gc_t1, float* mac, const PoolGeom& g, const SurfacePlan& plan,
                         uint32_t nflux, float* acc, const float* force, hipStream_t s);
// dirs: directions per thread (0: about 2^20 threads for the ring).
hipError_t launch_explode(int q, const ParentRef& parent, float* gpool, const PoolGeom& g, const SurfacePlan& plan,
                          uint32_t nring, int linear, int dirs, hipStream_t s);
hipError_t launch_coalesce(int q, const float* gpool, const ParentRef& parent, const float* mac_pool,
                           const PoolGeom& g, const int32_t* active, int nactive, const SurfacePlan& plan,
                           uint32_t nseam, uint32_t nnear, int orphans, hipStream_t s);
// ring_only: the ring cells alone (the interp interface's per-step refresh).
// A pool parent is sampled in the parent tile that holds the child block
// (its ring included), at the parent's own time level.
hipError_t launch_interp(int q, const ParentRef& parent, float* gpool, const PoolGeom& g, const int32_t* pairs,
                         int npairs, float rescale_c2f, int ring_only, hipStream_t s);
hipError_t launch_average(int q, const float* gpool, const ParentRef& parent, const PoolGeom& g,
                          const int32_t* pairs, int npairs, float rescale_f2c, hipStream_t s);
// kind 0: |curl u|; 1: Q = 1/2 (|Omega|^2 - |S|^2)
hipError_t launch_criterion(const float* mac, float* crit, const PoolGeom& g, float lead, int kind, hipStream_t s);
hipError_t launch_manage(const ManageArgs& a, hipStream_t s);
hipError_t launch_decide(const DecideArgs& a, hipStream_t s);
// Octet completion of a level >= 2 (completeOctets): a block is wanted when
// any block of its octet is (nbx, nby, nbz even).
hipError_t launch_complete_octets(uint8_t* want, int nbx, int nby, int nbz, hipStream_t s);
// The parent closure (balance): parent block p of the level with pnb blocks
// per axis becomes wanted when any child block in [2p - 1, 2p + 2]^3
// (periodic, on the child grid of 2 pnb per axis) is; ORs, never clears.
hipError_t launch_balance(uint8_t* want_parent, int pnbx, int pnby, int pnbz, const uint8_t* want_child,
                          hipStream_t s);
// active = [count, slots with a block ...] ascending
hipError_t launch_active_list(const int32_t* slot_to_block, int max_slots, int32_t* active, hipStream_t s);
// pairs = [count, (block, slot) ...] of an active list (2 max_slots + 1 ints)
hipError_t launch_active_pairs(const int32_t* active, const int32_t* slot_to_block, int max_slots, int32_t* pairs,
                               hipStream_t s);
hipError_t launch_pool_totals(int q, const float* gcoarse, const float* gpool, const PoolGeom& g, const int32_t* active,
                              double* partials, unsigned* counter, double* out, hipStream_t s);
// FP64 [mass, px, py, pz] of the interior cells of one pool level's active
// tiles that no finer tile covers, each weighted `weight` (8^-m).
hipError_t launch_level_totals(int q, const float* gpool, const PoolGeom& g, const int32_t* active, double weight,
                               double* partials, unsigned* counter, double* out, hipStream_t s);

}  // namespace lbm

// Block-pool (AMR) kernels for gfx950, two levels (dense L0 + one pool of
// fine tiles).  Semantics follow the reference's pool shaders:
//   pool step      shaders/common_d3_amr_step1.wgsl:160-405 (DIRECT_GHOST
//                  gather, ring cells advect only, interior collides at
//                  omega_fine)
//   explode        shaders/common_d3_amr_explode.wgsl:98-214
//   coalesce       shaders/common_d3_amr_coalesce.wgsl:126-285 (+ orphans)
//   interp fill    shaders/common_d3_amr_interp.wgsl:109-185 with
//                  interpCoarseToFine3, common_d3_pool.wgsl:192-220
//   average/drain  shaders/common_d3_amr_average.wgsl:75-154
//   flux / reflux  shaders/common_d3_amr_flux.wgsl:80-194 and
//                  common_d3_amr_reflux.wgsl:60-165 (the interp interface's
//                  optional seam correction)
//   criterion      shaders/common_d3_criterion.wgsl:122-158, with the
//                  vorticity magnitude |curl u| in place of Q
//   manager        shaders/common_d3_manage.wgsl:430-546 and
//                  amr_active_list.wgsl:41-70, re-designed to be
//                  deterministic: ordered compactions in one workgroup
//                  instead of an atomic free list.
// Populations are stored shifted (g = f - w); the surface passes convert to
// f where their arithmetic is not linear in f with unit total weight.
//
// The block edge RB is a template parameter (4, 8 or 16), so every tile
// index is shift-and-mask arithmetic.  Which L0 cells are covered, next to a
// covered cell, or on the covered side of a seam is read from a class byte
// per L0 cell (k_cell_class, rebuilt after each topology change) instead of
// being re-derived from the block map in every pass.  Neighbour coordinates
// move by at most one period, so they wrap with a compare, never a modulo.
#include <hip/hip_runtime.h>
#include <type_traits>

#include <algorithm>

#include "amr.h"
#include "collide.h"
#include "device_common.h"
#include "lattice_pairs.h"
#include "lattice_tables.h"

namespace lbm {

// A tile of a D-dimensional pool (D = 3: D3Q19/D3Q27; D = 2: the D2Q9
// quadtree).  A 2D tile is one layer of FB x FB cells: its z coordinate is
// pinned at LO (an interior layer, so every z test passes) and addresses
// ignore it, and its global z is 0 on a grid one cell deep.
template <int RB, int D = 3>
struct Tile {
  static constexpr int RB2 = 2 * RB;
  static constexpr int FB = RB2 + 2 * kGhost;
  static constexpr int CELLS = D == 3 ? FB * FB * FB : FB * FB;
  static constexpr int LO = kGhost, HI = kGhost + RB2;
  static constexpr int LOG_RB = RB == 4 ? 2 : (RB == 8 ? 3 : 4);
  static constexpr int NCH = 1 << D;               // children of a parent cell
  static constexpr float INV_NCH = D == 3 ? 0.125f : 0.25f;
  // 3D: FB = 4 (RB/2 + 1), so a tile is a whole number of 64-lane waves and
  // every wave of a tile-major launch lies in one tile
  static_assert(D == 2 || CELLS % 64 == 0, "tile cells must fill whole waves");
  static constexpr int RING = CELLS - (D == 3 ? RB2 * RB2 * RB2 : RB2 * RB2);  // ring cells per tile
  // tile-local coordinates of cell r
  __device__ static void xyz(uint32_t r, int& fx, int& fy, int& fz) {
    fx = (int)(r % FB);
    fy = (int)((r / FB) % FB);
    fz = D == 3 ? (int)(r / (FB * FB)) : LO;
  }
  // Ring cell j (0 <= j < RING) in class-major order.  A ring cell's class is
  // the side of the interior each axis lies on (low f < LO, mid, high
  // f >= HI); faces (two mid axes; one in 2D) come first, then edges, then
  // corners, each class x-fastest.  In 3D with RB >= 8 a face class
  // (2 x 2RB x 2RB cells) and an edge class (2 x 2 x 2RB) are whole 64-lane
  // waves and RING is a multiple of 64, so in a launch over tiles x RING the
  // waves of faces and edges hold one class each; the eight corner classes
  // share the last wave of a tile.
  __device__ static void ring_xyz(uint32_t j, int& fx, int& fy, int& fz) {
    int rem = (int)j;
    fx = fy = fz = LO;
    for (int nmid = D - 1; nmid >= 0; nmid--) {
      for (int c = 0; c < (D == 3 ? 27 : 9); c++) {
        const int kx = c % 3, ky = (c / 3) % 3, kz = D == 3 ? c / 9 : 1;
        if ((kx == 1) + (ky == 1) + (D == 3 && kz == 1) != nmid) continue;
        const int sx = kx == 1 ? RB2 : kGhost, sy = ky == 1 ? RB2 : kGhost;
        const int sz = D == 2 ? 1 : (kz == 1 ? RB2 : kGhost);
        if (rem < sx * sy * sz) {
          const int px = rem % sx, py = (rem / sx) % sy, pz = rem / (sx * sy);
          fx = kx == 0 ? px : (kx == 1 ? LO + px : HI + px);
          fy = ky == 0 ? py : (ky == 1 ? LO + py : HI + py);
          if (D == 3) fz = kz == 0 ? pz : (kz == 1 ? LO + pz : HI + pz);
          return;
        }
        rem -= sx * sy * sz;
      }
    }
  }
};
template <int Q, int RB>
using TileQ = Tile<RB, Lattice<Q>::D>;

// Block b of a launch runs on XCD b % 8, and each XCD has its own L2.
// Renumbering so that each XCD walks one contiguous eighth of the blocks
// keeps a tile's blocks, and the neighbouring rows and planes their pulls
// share, in one L2 (without it the pool step fetched 2.5x its algorithmic
// bytes from HBM).
__device__ __forceinline__ uint32_t xcd_block() {
  const uint32_t b = blockIdx.x, per = gridDim.x >> 3;
  return b < per * 8u ? (b & 7u) * per + (b >> 3) : b;
}

// Tile index of a tile-major launch, made wave-uniform (scalar registers)
// where a wave lies in one tile (3D); per lane for a 2D tile, which is not a
// whole number of waves.
template <int RB, int D = 3>
__device__ __forceinline__ uint32_t wave_tile(uint32_t t) {
  if constexpr (D == 3) return __builtin_amdgcn_readfirstlane(t / (uint32_t)Tile<RB>::CELLS);
  else return t / (uint32_t)Tile<RB, 2>::CELLS;
}

// v in [-n, 2n) -> [0, n)
__device__ __forceinline__ int wrap1(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

__device__ __forceinline__ uint32_t coarse_cell(const PoolGeom& G, int x, int y, int z) {
  return (uint32_t)((z * G.ny + y) * G.nx + x);
}
template <int RB, int D = 3>
__device__ __forceinline__ uint32_t pool_cell(int slot, int fx, int fy, int fz) {
  constexpr int FB = Tile<RB, D>::FB;
  if constexpr (D == 3) return (uint32_t)(slot * Tile<RB>::CELLS + (fz * FB + fy) * FB + fx);
  else return (uint32_t)(slot * Tile<RB, 2>::CELLS + fy * FB + fx);
}
__device__ __forceinline__ int block_id(const PoolGeom& G, int bx, int by, int bz) {
  return (bz * G.nby + by) * G.nbx + bx;
}
// Is L0 cell (x, y, z), each coordinate in [-n, 2n), under a tile.
__device__ __forceinline__ bool covered_at(const PoolGeom& G, int x, int y, int z) {
  return G.ccls[coarse_cell(G, wrap1(x, G.nx), wrap1(y, G.ny), wrap1(z, G.nz))] & kCovered;
}
// Tile-local index of global fine coordinate g relative to block coordinate
// b on one axis, nearest periodic image (localInTile, common_d3_pool.wgsl:
// 81-93; valid with three or more blocks per axis, which the engine checks).
template <int RB>
__device__ __forceinline__ int local_in_tile(int g, int b, int nfine) {
  int d = g - b * Tile<RB>::RB2;
  if (d < -nfine / 2) d += nfine;
  if (d > nfine / 2) d -= nfine;
  return d + kGhost;
}

// Parent cells: index in the parent's storage (the dense parent's cell, or
// the pool index in the parent tile that owns the cell, kAbsent when that
// block holds no tile), and the population / moment accessors.  (x, y, z) is
// wrapped into the parent grid.
constexpr uint32_t kAbsent = 0xffffffffu;
template <bool PP, int RB>
__device__ __forceinline__ uint32_t parent_index(const PoolGeom& G, const ParentRef& P, int x, int y, int z) {
  if constexpr (!PP) {
    return coarse_cell(G, x, y, z);
  } else {
    constexpr int SH = Tile<RB>::LOG_RB + 1;
    const int bx = x >> SH, by = y >> SH, bz = z >> SH;
    const int slot = P.block_slot[(bz * P.nby + by) * P.nbx + bx];
    if (slot < 0) return kAbsent;
    return pool_cell<RB>(slot, x - (bx << SH) + kGhost, y - (by << SH) + kGhost, z - (bz << SH) + kGhost);
  }
}
template <bool PP, int Q>
__device__ __forceinline__ float* ppop(const ParentRef& P, uint32_t idx, int i) {
  if constexpr (PP) return P.g + (size_t)i * P.plane + idx;
  else return P.g + lat_elem(Q, i, idx);
}
template <bool PP>
__device__ __forceinline__ void pmac_put(const ParentRef& P, uint32_t idx, float4 v) {
  if constexpr (PP) {
    P.mac[idx] = v.x;
    P.mac[(size_t)P.plane + idx] = v.y;
    P.mac[2 * (size_t)P.plane + idx] = v.z;
    P.mac[3 * (size_t)P.plane + idx] = v.w;
  } else {
    reinterpret_cast<float4*>(P.mac)[idx] = v;
  }
}

// L0-unit position of this level's cell j (tile-local, ring included) of a
// block whose first parent cell is `origin`: the parent-unit position
// origin - 1/4 + (j - GHOST)/2 (fineToCoarseUnit3), mapped to L0 by the
// level's affine scale (1 and 0 at level 1).
__device__ __forceinline__ float level_to_l0(const PoolGeom& G, int origin, int j) {
  return ((float)origin - 0.25f + 0.5f * (float)(j - kGhost)) * G.l0_scale + G.l0_offset;
}

// Is this level's interior cell, at its global coordinate (gx, gy, gz) of
// this level's grid (2 nx wide), under a tile of the next finer level.
template <int RB>
__device__ __forceinline__ bool covered_by_child(const PoolGeom& G, int gx, int gy, int gz) {
  if (!G.child_block_slot) return false;
  gx = wrap1(gx, 2 * G.nx);
  gy = wrap1(gy, 2 * G.ny);
  gz = wrap1(gz, 2 * G.nz);
  constexpr int SH = Tile<RB>::LOG_RB;
  return G.child_block_slot[((gz >> SH) * G.cnby + (gy >> SH)) * G.cnbx + (gx >> SH)] >= 0;
}

// Velocity and weight of direction i when i is uniform over the block (a
// surface pass launched with one grid row per direction): compares and
// selects on the scalar unit instead of a table indexed at run time.
template <int Q>
__device__ __forceinline__ void direction(int i, int& ex, int& ey, int& ez, float& w) {
  using L = Lattice<Q>;
  ex = 0, ey = 0, ez = 0, w = L::w[0];
#pragma unroll
  for (int k = 1; k < Q; k++)
    if (i == k) ex = L::ex[k], ey = L::ey[k], ez = L::ez[k], w = L::w[k];
}

// Signed distance to the pinned sphere in binary32, evaluated as the dense
// classification evaluates it (common_d3_geometry.wgsl:119, :174-187), so a
// fine cell and the dense cell at the same point agree on which side they are.
__device__ __forceinline__ float sphere_phi(const PoolBody& b, float x, float y, float z) {
  const float dx = sub_rn(x, b.cx), dy = sub_rn(y, b.cy), dz = sub_rn(z, b.cz);
  return sub_rn(sqrt_rn(add_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy)), mul_rn(dz, dz))), b.r);
}

// A pull source lies at most sqrt(3)/2 L0 cells from the cell, so a cell at
// phi >= 1 has no solid source.
constexpr float kBodyReach = 1.f;

// ---------------------------------------------------------------------------
// L0 cell classes
// ---------------------------------------------------------------------------
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256) k_cell_class(uint8_t* __restrict__ ccls, PoolGeom G, SurfacePlan plan,
                                                    ParentRef P) {
  using L = Lattice<Q>;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.ncells) return;
  const int x = c % G.nx, y = (c / G.nx) % G.ny, z = c / (G.nx * G.ny);
  auto cov = [&](int cx, int cy, int cz) {
    cx = wrap1(cx, G.nx);
    cy = wrap1(cy, G.ny);
    cz = wrap1(cz, G.nz);
    return G.block_slot[block_id(G, cx >> Tile<RB>::LOG_RB, cy >> Tile<RB>::LOG_RB, cz >> Tile<RB>::LOG_RB)] >= 0;
  };
  const bool self = cov(x, y, z);
  bool other = false;  // a lattice neighbour with the opposite state
#pragma unroll
  for (int i = 1; i < Q; i++) other = other || (cov(x + L::ex[i], y + L::ey[i], z + L::ez[i]) != self);
  ccls[c] = self ? (uint8_t)(kCovered | (other ? kSeam : 0)) : (uint8_t)(other ? kNear : 0);
  // a pool parent's cells outside its tiles hold no data: never listed
  if (PP && parent_index<PP, RB>(G, P, x, y, z) == kAbsent) return;
  if (plan.flux && !self) {
    // a refined block among the blocks the 26 neighbours span
    // (common_d3_amr_flux.wgsl:101-122): any covered cell in the 3x3x3 box
    bool any = false;
#pragma unroll
    for (int k = 0; k < 27; k++) any = any || cov(x + k % 3 - 1, y + (k / 3) % 3 - 1, z + k / 9 - 1);
    if (any) {
      const uint32_t j = atomicAdd(&plan.count[3], 1u);
      if (j < plan.cap_flux) plan.flux[j] = c;
    }
  }
  if (other) {
    if (self) {
      const uint32_t k = atomicAdd(&plan.count[0], 1u);
      if (k < plan.cap_seam) plan.seam[k] = c;
    } else {
      const uint32_t k = atomicAdd(&plan.count[1], 1u);
      if (k < plan.cap_near) plan.near[k] = c;
    }
  }
}

// What explode needs to know about each ring cell, all fixed by the topology:
// one entry per ring cell of each active tile, at k * RING + j for active
// tile k and ring cell j in class-major order (Tile::ring_xyz):
//   .x  pool cell (slot * FB^3 + tile-local index), or kAbsent where the L0
//       cell under the ring cell is refined (explode writes nothing there)
//   .y  the L0 cell v under it
//   .z  bit i (i < Q): the L0 cell v + e_i is refined; bits 27-29: the two
//       L0 neighbours of v along x / y / z are both unrefined (the linear
//       reconstruction's gradient along that axis is taken)
//   .w  v's coordinates, x | y << 10 | z << 20
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256) k_ring_plan(PoolGeom G, SurfacePlan plan, const int32_t* __restrict__ active,
                                                   int nactive, ParentRef P) {
  using L = Lattice<Q>;
  using T = TileQ<Q, RB>;
  constexpr int RB2 = T::RB2, FB = T::FB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / (uint32_t)T::RING;
  if (k >= (uint32_t)nactive) return;
  int fx, fy, fz;
  T::ring_xyz(t - k * (uint32_t)T::RING, fx, fy, fz);
  const uint32_t r = L::D == 3 ? (uint32_t)((fz * FB + fy) * FB + fx) : (uint32_t)(fy * FB + fx);
  plan.ring[t] = make_uint4(kAbsent, 0u, 0u, 0u);
  const int slot = active[1 + k];
  const int id = G.slot_to_block[slot];
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  const int gx = wrap1(bx * RB2 + fx - kGhost, 2 * G.nx);
  const int gy = wrap1(by * RB2 + fy - kGhost, 2 * G.ny);
  const int gz = wrap1(bz * RB2 + fz - kGhost, 2 * G.nz);
  const int vx = gx >> 1, vy = gy >> 1, vz = gz >> 1;
  if (G.ccls[coarse_cell(G, vx, vy, vz)] & kCovered) return;
  // the parent cell always lies in a parent tile (the ring-parent coverage
  // the topology checks guarantee; checkRingParentCoverage, d3-amr.mjs:348)
  const uint32_t vc = parent_index<PP, RB>(G, P, vx, vy, vz);
  if (vc == kAbsent) return;
  // a gradient axis also needs both parent neighbours present; where one is
  // missing the explosion is uniform along it (parentPresent,
  // common_d3_parent_pool.wgsl:31-35)
  auto open = [&](int x, int y, int z) {
    return !covered_at(G, x, y, z) &&
           (!PP || parent_index<PP, RB>(G, P, wrap1(x, G.nx), wrap1(y, G.ny), wrap1(z, G.nz)) != kAbsent);
  };
  uint32_t bits = 0;
#pragma unroll
  for (int i = 0; i < Q; i++)
    if (covered_at(G, vx + L::ex[i], vy + L::ey[i], vz + L::ez[i])) bits |= 1u << i;
  if (open(vx - 1, vy, vz) && open(vx + 1, vy, vz)) bits |= 1u << 27;
  if (open(vx, vy - 1, vz) && open(vx, vy + 1, vz)) bits |= 1u << 28;
  if (open(vx, vy, vz - 1) && open(vx, vy, vz + 1)) bits |= 1u << 29;
  plan.ring[t] = make_uint4((uint32_t)slot * T::CELLS + r, vc, bits,
                            (uint32_t)vx | (uint32_t)vy << 10 | (uint32_t)vz << 20);
}

uint32_t ring_per_tile(int rb, int dim) {
  auto pw = [&](uint32_t v) { return dim == 3 ? v * v * v : v * v; };
  return pw(2 * rb + 2 * kGhost) - pw(2 * rb);
}

void plan_capacity(int rb, int max_slots, uint32_t ncells, int dim, uint32_t* seam, uint32_t* near, uint32_t* ring) {
  auto pw = [&](uint64_t v) { return dim == 3 ? v * v * v : v * v; };
  const uint64_t shell = pw((uint64_t)rb) - pw((uint64_t)(rb - 2));
  *seam = (uint32_t)std::min<uint64_t>((uint64_t)max_slots * shell, ncells);
  *near = ncells;
  *ring = (uint32_t)((uint64_t)max_slots * ring_per_tile(rb, dim));
}

// The per-tile neighbour table (kTileWords words per active tile).  The pool
// step reads it with wave-uniform loads in place of per-lane block-map
// lookups (slot -> block -> neighbour block -> slot), which sat on every
// lane's path to its population loads.
__global__ void __launch_bounds__(256) k_tile_table(PoolGeom G, const int32_t* __restrict__ active, int nactive,
                                                    int32_t* __restrict__ table) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nactive) return;
  const int slot = active[1 + k];
  const int id = G.slot_to_block[slot];
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  int32_t* t = table + (size_t)k * kTileWords;
  uint32_t mask = 0;
  for (int n = 0; n < 27; n++) {
    const int ox = n % 3 - 1, oy = (n / 3) % 3 - 1, oz = n / 9 - 1;
    const int s = G.block_slot[block_id(G, wrap1(bx + ox, G.nbx), wrap1(by + oy, G.nby), wrap1(bz + oz, G.nbz))];
    t[n] = s;
    if (s >= 0) mask |= 1u << n;
  }
  t[27] = (int32_t)mask;
  t[28] = bx;
  t[29] = by;
  t[30] = bz;
  t[31] = 0;
}

// ---------------------------------------------------------------------------
// pool step
// ---------------------------------------------------------------------------
// One pool cell's substep: pull (the DIRECT_GHOST gather), moments, and,
// when `collide_cell`, the fine collision; `mac_pool` (or null) receives
// the moments.  A lane whose result nothing reads (`load` false) loads
// nothing and stores zeros, so that every wave still writes whole cache
// lines; `put_mac` false keeps the cell's moments as they are.
// BODY: the fine-level pinned sphere (its own instantiation, so that the
// body-free kernel keeps its register budget).  A link whose geometric source
// is solid reflects this cell's own opposite population (halfway bounce-back,
// common_d3_amr_step1.wgsl:258-276; the body is at rest, so Ladd's term is
// zero), and a solid cell is held at feq(1, 0) with [1, 0, 0, 0] as its
// moments (:387-403).
template <int Q, int RB, bool BODY, bool FRC>
__device__ __forceinline__ void pool_cell_step(const float* __restrict__ gin, float* __restrict__ gout,
                                               float* __restrict__ mac_pool, const PoolGeom& G, const CollideParams& P,
                                               const int32_t* __restrict__ tt, int fx, int fy, int fz,
                                               bool collide_cell, bool load = true, bool put_mac = true) {
  const int slot = tt[13], bx = tt[28], by = tt[29], bz = tt[30];
  using L = Lattice<Q>;
  using T = TileQ<Q, RB>;
  constexpr int D = L::D, FB = T::FB, RB2 = T::RB2, LO = T::LO, HI = T::HI;
  const uint32_t cell = pool_cell<RB, D>(slot, fx, fy, fz);
  float g[Q];
  const int offx = fx <= LO ? -1 : (fx + 1 >= HI ? 1 : 0);
  const int offy = fy <= LO ? -1 : (fy + 1 >= HI ? 1 : 0);
  const int offz = fz <= LO ? -1 : (fz + 1 >= HI ? 1 : 0);
  // the slots of the neighbour blocks this cell's pulls can reach, selected
  // from the tile's table (word (ox + 1) + 3 (oy + 1) + 9 (oz + 1))
  const int nX = offx < 0 ? tt[12] : (offx > 0 ? tt[14] : -1);
  const int nY = offy < 0 ? tt[10] : (offy > 0 ? tt[16] : -1);
  const int nZ = offz < 0 ? tt[4] : (offz > 0 ? tt[22] : -1);
  const int nXY = !(offx && offy) ? -1 : (offy < 0 ? (offx < 0 ? tt[9] : tt[11]) : (offx < 0 ? tt[15] : tt[17]));
  const int nXZ = !(offx && offz) ? -1 : (offz < 0 ? (offx < 0 ? tt[3] : tt[5]) : (offx < 0 ? tt[21] : tt[23]));
  const int nYZ = !(offy && offz) ? -1 : (offz < 0 ? (offy < 0 ? tt[1] : tt[7]) : (offy < 0 ? tt[19] : tt[25]));
  const int nXYZ = !(offx && offy && offz)
                       ? -1
                       : (offz < 0 ? (offy < 0 ? (offx < 0 ? tt[0] : tt[2]) : (offx < 0 ? tt[6] : tt[8]))
                                   : (offy < 0 ? (offx < 0 ? tt[18] : tt[20]) : (offx < 0 ? tt[24] : tt[26])));
#pragma unroll
  for (int i = 0; i < Q; i++) {
    int sx = fx - L::ex[i], sy = fy - L::ey[i], sz = fz - L::ez[i];
    const int ox = (sx < LO || sx >= HI) ? offx : 0;
    const int oy = (sy < LO || sy >= HI) ? offy : 0;
    const int oz = (sz < LO || sz >= HI) ? offz : 0;
    int ns = -1;
    if (ox && oy && oz) ns = nXYZ;
    else if (ox && oy) ns = nXY;
    else if (ox && oz) ns = nXZ;
    else if (oy && oz) ns = nYZ;
    else if (ox) ns = nX;
    else if (oy) ns = nY;
    else if (oz) ns = nZ;
    int src_slot = slot;
    if (ns >= 0) {
      // the source lies in a refined neighbour: its interior, same time level
      src_slot = ns;
      sx -= ox * RB2;
      sy -= oy * RB2;
      sz -= oz * RB2;
    }
    sx = min(max(sx, 0), FB - 1);
    sy = min(max(sy, 0), FB - 1);
    sz = min(max(sz, 0), FB - 1);
    g[i] = load ? *at(gin + (size_t)i * G.pool_plane, pool_cell<RB, D>(src_slot, sx, sy, sz)) : 0.f;
  }
  float phi = 1e30f;
  if (BODY) {
    // positions in L0 units (exact: every term is a small dyadic rational)
    const float px = level_to_l0(G, bx * RB, fx), py = level_to_l0(G, by * RB, fy), pz = level_to_l0(G, bz * RB, fz);
    phi = sphere_phi(G.body, px, py, pz);
    if (load && phi < kBodyReach) {
#pragma unroll
      for (int i = 1; i < Q; i++) {
        // the geometric source, one cell of this level back
        const float sx = level_to_l0(G, bx * RB, fx - L::ex[i]), sy = level_to_l0(G, by * RB, fy - L::ey[i]),
                    sz = level_to_l0(G, bz * RB, fz - L::ez[i]);
        if (sphere_phi(G.body, sx, sy, sz) < 0.f) g[i] = *at(gin + (size_t)L::opp[i] * G.pool_plane, cell);
      }
    }
  }

  if (mac_pool || collide_cell) {
    float drho, rho, ux, uy, uz;
    float S[Pairs<Q>::NP], D[Pairs<Q>::NP];
    pair_sums<Q>(g, S, D);
    // D2Q9's f32 weights do not sum to 1: the density adds their excess, as
    // the dense step does
    pair_moments<Q, FRC>(Q == 9 ? g[0] + P.wsum_m1 : g[0], S, D, P.F, drho, rho, ux, uy, uz);
    if (mac_pool && put_mac) {
      *at(mac_pool, cell) = rho;
      *at(mac_pool + G.pool_plane, cell) = ux;
      *at(mac_pool + 2 * (size_t)G.pool_plane, cell) = uy;
      *at(mac_pool + 3 * (size_t)G.pool_plane, cell) = uz;
    }
    if (collide_cell) collide_pairs<Q, FRC>(g, S, D, P, drho, rho, ux, uy, uz);
  }
  if (BODY && phi < 0.f) {
#pragma unroll
    for (int i = 0; i < Q; i++) g[i] = 0.f;  // feq(1, 0) - w
    if (mac_pool && put_mac) {
      *at(mac_pool, cell) = 1.f;
      *at(mac_pool + G.pool_plane, cell) = 0.f;
      *at(mac_pool + 2 * (size_t)G.pool_plane, cell) = 0.f;
      *at(mac_pool + 3 * (size_t)G.pool_plane, cell) = 0.f;
    }
  }
#pragma unroll
  for (int i = 0; i < Q; i++) *at(gout + (size_t)i * G.pool_plane, cell) = g[i];
}

// One thread per pool cell of every active tile, in memory order (a wave is
// 64 consecutive cells of one tile), and every cell stores: a 20-cell row is
// not a whole number of cache lines, so a lane that stored nothing would
// leave every line of its wave partly written.  On the explode interface
// (sub 0 or 1, the first or second substep) the cells whose result nothing
// reads load nothing and store zeros:
// - ring cells over a refined neighbour block, in both substeps: pulls into
//   that region take the neighbour tile's interior (the gather above), and
//   explode, coalesce and the orphans touch only ring cells over unrefined
//   blocks;
// - the outer ring layer in the first substep: what it holds can only reach,
//   in the second, the inward and tangential populations of the ring, and
//   coalesce and the orphans read populations leaving the interior alone;
// - interior cells a finer tile covers (their moments are the child's
//   coalesced ones, kept; their populations are read by nothing before the
//   next coalesce or restriction rewrites the ones that are).
// sub -1 (the interp interface) steps every cell.
template <int Q, int RB, bool BODY, bool FRC>
__global__ void __launch_bounds__(256)
k_pool_step(const float* __restrict__ gin, float* __restrict__ gout, float* __restrict__ mac_pool, PoolGeom G,
            CollideParams P, int nactive, int collide_ring, int sub) {
  using T = TileQ<Q, RB>;
  constexpr int FB = T::FB, LO = T::LO, HI = T::HI;
  const uint32_t t = xcd_block() * blockDim.x + threadIdx.x;
  const uint32_t k = wave_tile<RB, Lattice<Q>::D>(t);
  if (k >= (uint32_t)nactive) return;
  const uint32_t r = t - k * (uint32_t)T::CELLS;
  const int32_t* tt = G.tile_table + (size_t)k * kTileWords;
  int fx, fy, fz;
  T::xyz(r, fx, fy, fz);
  const bool interior = fx >= LO && fx < HI && fy >= LO && fy < HI && fz >= LO && fz < HI;
  bool load = true, put_mac = true;
  if (interior) {
    // An interior cell a finer level solves keeps the moments the child's
    // coalesce wrote (coveredByChild, common_d3_amr_step1.wgsl:107-143).
    if (G.child_block_slot && covered_by_child<RB>(G, tt[28] * T::RB2 + fx - kGhost, tt[29] * T::RB2 + fy - kGhost,
                                                   tt[30] * T::RB2 + fz - kGhost))
      load = put_mac = false;
  } else if (sub >= 0) {
    const int ox = fx < LO ? -1 : (fx >= HI ? 1 : 0);
    const int oy = fy < LO ? -1 : (fy >= HI ? 1 : 0);
    const int oz = fz < LO ? -1 : (fz >= HI ? 1 : 0);
    const bool over_refined = ((uint32_t)tt[27] >> ((ox + 1) + 3 * (oy + 1) + 9 * (oz + 1))) & 1u;
    const bool outer = fx == 0 || fx == FB - 1 || fy == 0 || fy == FB - 1 || fz == 0 || fz == FB - 1;
    load = !over_refined && !(sub == 0 && outer);
  }
  // A wave none of whose cells' results is read stores nothing either: in 3D
  // a wave is 64 consecutive cells of one tile from a line-aligned start, so
  // its stores would cover whole cache lines that no other wave writes.
  if (!__any(load)) return;
  // Ring cells advect and store without colliding (common_d3_amr_step1.wgsl
  // COLLIDE_RING = 0): coalesce averages uncollided interface states.  The
  // interp interface collides them too (COLLIDE_RING = 1, main-3d.js:1395).
  pool_cell_step<Q, RB, BODY, FRC>(gin, gout, mac_pool, G, P, tt, fx, fy, fz, load && (interior || collide_ring),
                                   load, put_mac);
}

// Momentum exchange on the fine-level body (common_d3_force_pool.wgsl, the
// bounce-back branch of a body at rest): every interior fluid cell with a
// solid geometric source of direction i adds -e_i 2 f_opp(i), read from the
// pool state the macro-step starts from; the sum is weighted by the fine
// cell's face area dx^2 = 1/4 and its torque taken about the centre in L0
// units.  One thread per interior cell of each active tile; each block leaves
// its FP64 partial in `partials`, summed in block order afterwards.
template <int Q, int RB>
__global__ void __launch_bounds__(256)
k_pool_force(const float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ active, int nactive,
             double* __restrict__ partials) {
  using L = Lattice<Q>;
  constexpr int RB2 = 2 * RB;
  constexpr uint32_t per = RB2 * RB2 * RB2;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / per;
  double fv[6] = {0, 0, 0, 0, 0, 0};
  if (k < (uint32_t)nactive) {
    const uint32_t r = t - k * per;
    const int fx = kGhost + (int)(r % RB2), fy = kGhost + (int)((r / RB2) % RB2), fz = kGhost + (int)(r / (RB2 * RB2));
    const int slot = active[1 + k];
    const int id = G.slot_to_block[slot];
    const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
    const float px = level_to_l0(G, bx * RB, fx), py = level_to_l0(G, by * RB, fy), pz = level_to_l0(G, bz * RB, fz);
    const float phi = sphere_phi(G.body, px, py, pz);
    if (phi >= 0.f && phi < kBodyReach) {
      const uint32_t cell = pool_cell<RB>(slot, fx, fy, fz);
      double f0 = 0.0, f1 = 0.0, f2 = 0.0;
#pragma unroll
      for (int i = 1; i < Q; i++) {
        const float sx = level_to_l0(G, bx * RB, fx - L::ex[i]), sy = level_to_l0(G, by * RB, fy - L::ey[i]),
                    sz = level_to_l0(G, bz * RB, fz - L::ez[i]);
        if (sphere_phi(G.body, sx, sy, sz) < 0.f) {
          const double m = 2.0 * ((double)*at(gp + (size_t)L::opp[i] * G.pool_plane, cell) + (double)L::w[i]);
          f0 -= L::ex[i] * m;
          f1 -= L::ey[i] * m;
          f2 -= L::ez[i] * m;
        }
      }
      // weighted dx^2 = 4^-m (common_d3_force_pool.wgsl DX_WEIGHT)
      f0 *= (double)G.dx2;
      f1 *= (double)G.dx2;
      f2 *= (double)G.dx2;
      const double rx = (double)px - G.body.cx, ry = (double)py - G.body.cy, rz = (double)pz - G.body.cz;
      fv[0] = f0;
      fv[1] = f1;
      fv[2] = f2;
      fv[3] = ry * f2 - rz * f1;
      fv[4] = rz * f0 - rx * f2;
      fv[5] = rx * f1 - ry * f0;
    }
  }
  __shared__ double lds[4 * 6];
  block_sum<6>(fv, lds);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 6; c++) partials[(size_t)blockIdx.x * 6 + c] = fv[c];
  }
}

// ---------------------------------------------------------------------------
// explode: coarse state at t -> ring of the pool buffer the first substep reads
// ---------------------------------------------------------------------------
// Bit i set for the directions whose component on `axis` is v.
template <int Q>
__device__ constexpr uint32_t dirs_with(int axis, int v) {
  uint32_t m = 0;
  for (int i = 0; i < Q; i++) {
    const int e = axis == 0 ? Lattice<Q>::ex[i] : (axis == 1 ? Lattice<Q>::ey[i] : Lattice<Q>::ez[i]);
    if (e == v) m |= 1u << i;
  }
  return m;
}
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_explode(ParentRef P, float* __restrict__ gp, PoolGeom G, const uint4* __restrict__ ring, uint32_t nring, int linear,
          int per) {
  using L = Lattice<Q>;
  using T = TileQ<Q, RB>;
  constexpr int LO = T::LO, HI = T::HI;
  const uint32_t t = xcd_block() * blockDim.x + threadIdx.x;
  if (t >= nring) return;
  const uint4 d = ring[t];
  if (d.x == kAbsent) return;  // the L0 cell under this ring cell is refined
  // Grid row g takes directions [g per, (g + 1) per): a ring cell's Q
  // populations are independent, so a small ring spreads them over threads
  // (one thread per ring cell leaves too few waves to hide the loads) and a
  // large one keeps them in one thread (which shares the index work among
  // them).
  const int i0 = (int)blockIdx.y * per, i1 = min(Q, i0 + per);
  // a ring cell over an unrefined L0 cell, and its explode decisions
  // (k_ring_plan)
  const uint32_t cell = d.x, vc = d.y, bits = d.z;
  int fx, fy, fz;
  T::xyz(cell % (uint32_t)T::CELLS, fx, fy, fz);
  const int vx = d.w & 1023, vy = (d.w >> 10) & 1023, vz = d.w >> 20;
  const bool okx = (bits >> 27) & 1, oky = (bits >> 28) & 1, okz = (bits >> 29) & 1;
  // the gradient's neighbours (present wherever their axis is ok)
  auto pidx = [&](int x, int y, int z) { return parent_index<PP, RB>(G, P, x, y, z); };
  const uint32_t cmx = okx ? pidx(wrap1(vx - 1, G.nx), vy, vz) : vc, cpx = okx ? pidx(wrap1(vx + 1, G.nx), vy, vz) : vc;
  const uint32_t cmy = oky ? pidx(vx, wrap1(vy - 1, G.ny), vz) : vc, cpy = oky ? pidx(vx, wrap1(vy + 1, G.ny), vz) : vc;
  const uint32_t cmz = okz ? pidx(vx, vy, wrap1(vz - 1, G.nz)) : vc, cpz = okz ? pidx(vx, vy, wrap1(vz + 1, G.nz)) : vc;
  // position of the fine cell inside its coarse cell, in coarse units: the
  // global fine coordinate is 2 * block + local - 2, so its parity is the
  // local one's
  const float dx = ((float)(fx & 1) - 0.5f) * 0.5f;
  const float dy = ((float)(fy & 1) - 0.5f) * 0.5f;
  const float dz = ((float)(fz & 1) - 0.5f) * 0.5f;
  // Ring populations only advect (the ring does not collide), so the one in
  // direction e at ring cell p reaches the interior within the macro-step's
  // two substeps only if p + e or p + 2e is an interior cell; nothing else
  // reads the others (coalesce and the orphans take populations that left
  // the interior at t).  Without a body on this level explode writes the
  // directions that point from the cell's side of the ring towards the
  // interior on every axis (e = +1 on an axis below the interior, -1 above
  // it, any value on an axis within it): a superset of those, fixed by the
  // cell's ring class and so the same across a wave of one class.  With a
  // body, the ring's bounce-back reflects a cell's own opposite population,
  // so every direction is written.
  constexpr uint32_t ALL = (1u << Q) - 1u;
  auto side = [&](int f, uint32_t plus, uint32_t minus) { return f < LO ? plus : (f >= HI ? minus : ALL); };
  const uint32_t dirs = G.body.on ? ALL
                                  : side(fx, dirs_with<Q>(0, 1), dirs_with<Q>(0, -1)) &
                                        side(fy, dirs_with<Q>(1, 1), dirs_with<Q>(1, -1)) &
                                        side(fz, dirs_with<Q>(2, 1), dirs_with<Q>(2, -1));
  auto inside = [&](int x, int y, int z) { return x >= LO && x < HI && y >= LO && y < HI && z >= LO && z < HI; };
#pragma unroll
  for (int i = 0; i < Q; i++) {
    if (i < i0 || i >= i1) continue;  // wave-uniform
    if (!((dirs >> i) & 1u)) continue;
    const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
    const float wi = L::w[i];
    const bool into = inside(fx + ex, fy + ey, fz + ez);
    auto l0 = [&](uint32_t c) { return *ppop<PP, Q>(P, c, i); };  // population i of parent cell c
    float out;
    if (!((bits >> i) & 1)) {
      // The coarse target is unrefined: only a population that streams from
      // this ring cell into this tile's interior is kept (in-orphan, uniform
      // injection); everything else is zeroed (f = 0, i.e. g = -w).
      out = into ? l0(vc) : -wi;
    } else if (!linear || i == 0) {
      out = l0(vc);
    } else {
      // linear reconstruction d . G with the gradient projected
      // perpendicular to e_i (common_d3_amr_explode.wgsl:197-212)
      float Gx = okx ? 0.5f * (l0(cpx) - l0(cmx)) : 0.f;
      float Gy = oky ? 0.5f * (l0(cpy) - l0(cmy)) : 0.f;
      float Gz = okz ? 0.5f * (l0(cpz) - l0(cmz)) : 0.f;
      const float ee = (float)(ex * ex + ey * ey + ez * ez);
      const float proj = (Gx * ex + Gy * ey + Gz * ez) / ee;
      Gx -= ex * proj;
      Gy -= ey * proj;
      Gz -= ez * proj;
      out = l0(vc) + (dx * Gx + dy * Gy + dz * Gz);
    }
    *at(gp + (size_t)i * G.pool_plane, cell) = out;
  }
}

// ---------------------------------------------------------------------------
// coalesce: fine ring outflux -> time-t coarse cells under refined blocks,
// plus the orphan correction on unrefined cells next to them
// ---------------------------------------------------------------------------
// mac projection of every covered L0 cell from its 8 children (4 in 2D;
// rho mean, mass-weighted u; common_d3_amr_coalesce.wgsl:245-261)
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_coalesce_mac(ParentRef P, const float* __restrict__ mac_pool, PoolGeom G, const int32_t* __restrict__ active,
               int nactive) {
  using T = TileQ<Q, RB>;
  constexpr int D = Lattice<Q>::D;
  constexpr uint32_t per = D == 3 ? RB * RB * RB : RB * RB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / per;
  if (k >= (uint32_t)nactive) return;
  const uint32_t r = t - k * per;
  const int slot = active[1 + k];
  const int id = G.slot_to_block[slot];
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  const int lx = r % RB, ly = (r / RB) % RB, lz = r / (RB * RB);
  const int f0x = kGhost + 2 * lx, f0y = kGhost + 2 * ly, f0z = kGhost + 2 * lz;
  float rs = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
#pragma unroll
  for (int c = 0; c < T::NCH; c++) {
    const uint32_t pc = pool_cell<RB, D>(slot, f0x + (c & 1), f0y + ((c >> 1) & 1), f0z + ((c >> 2) & 1));
    const float rho = *at(mac_pool, pc);
    rs += rho;
    mx += rho * *at(mac_pool + G.pool_plane, pc);
    my += rho * *at(mac_pool + 2 * (size_t)G.pool_plane, pc);
    mz += rho * *at(mac_pool + 3 * (size_t)G.pool_plane, pc);
  }
  const float inv = 1.f / fmaxf(rs, 1e-6f);
  const uint32_t cc = parent_index<PP, RB>(G, P, bx * RB + lx, by * RB + ly, bz * RB + lz);
  if (PP && cc == kAbsent) return;  // a tile's parent cells lie in parent tiles
  pmac_put<PP>(P, cc, make_float4(rs * T::INV_NCH, mx * inv, my * inv, mz * inv));
}

// Covered seam cells: for every direction whose target is unrefined, the
// mean of the 8 ring cells covering the target (common_d3_amr_coalesce.wgsl:
// 264-284).  One thread per (seam cell, direction i = 1 + grid row).
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_coalesce_seam(const float* __restrict__ gp, ParentRef P, PoolGeom G, const uint32_t* __restrict__ seam,
                uint32_t nseam) {
  using T = TileQ<Q, RB>;
  constexpr int FB = T::FB, D = Lattice<Q>::D;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseam) return;
  const int i = 1 + blockIdx.y;
  int ex, ey, ez;
  float wi;
  direction<Q>(i, ex, ey, ez, wi);
  const uint32_t c = seam[t];
  const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
  const int tx = cx + ex, ty = cy + ey, tz = cz + ez;
  if (covered_at(G, tx, ty, tz)) return;
  const int bx = cx >> T::LOG_RB, by = cy >> T::LOG_RB, bz = cz >> T::LOG_RB;
  const int slot = G.block_slot[block_id(G, bx, by, bz)];
  // the target is one coarse cell outside the block: its 8 children are a
  // 2x2x2 cube of this tile's ring
  const int lx = local_in_tile<RB>(2 * tx, bx, 2 * G.nx);
  const int ly = local_in_tile<RB>(2 * ty, by, 2 * G.ny);
  const int lz = local_in_tile<RB>(2 * tz, bz, 2 * G.nz);
  const float* p = gp + (size_t)i * G.pool_plane;
  const uint32_t b0 = pool_cell<RB, D>(slot, lx, ly, lz);
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < T::NCH; k++) s += *at(p, b0 + (k & 1) + ((k >> 1) & 1) * FB + ((k >> 2) & 1) * FB * FB);
  // an average of g is the average of f minus w
  *ppop<PP, Q>(P, PP ? parent_index<PP, RB>(G, P, cx, cy, cz) : c, i) = s * T::INV_NCH;
}

// Orphans on unrefined L0 cells next to the refined region
// (coalesceOrphans, common_d3_amr_coalesce.wgsl:126-211).  One thread per
// (near cell, direction i = 1 + grid row); the 8 child lookups of a
// direction are independent loads.
template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_orphans(const float* __restrict__ gp, ParentRef P, PoolGeom G, const uint32_t* __restrict__ near, uint32_t nnear) {
  using T = TileQ<Q, RB>;
  constexpr int D = Lattice<Q>::D;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nnear) return;
  const int i = 1 + blockIdx.y;
  int ex, ey, ez;
  float wi;
  direction<Q>(i, ex, ey, ez, wi);
  const uint32_t c = near[t];
  const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
  const int NFx = 2 * G.nx, NFy = 2 * G.ny, NFz = 2 * G.nz;
  const int tx = wrap1(cx + ex, G.nx), ty = wrap1(cy + ey, G.ny), tz = wrap1(cz + ez, G.nz);
  if (G.ccls[coarse_cell(G, tx, ty, tz)] & kCovered) return;
  // Child o of the target sends population i from the fine cell
  // q = 2 (c + e) + o - e = 2 c + o + e, which is also the child whose coarse
  // cell the in-count tests (coarse cell (2 c + o + e) / 2 covered); a coarse
  // cell is covered exactly when its block holds a tile, so the in-count is
  // the number of children found.
  float s = 0.f;   // sum of f over the children found in refined tiles
  float nIn = 0.f;
#pragma unroll
  for (int k = 0; k < T::NCH; k++) {
    const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
    const int px = 2 * tx + ox, py = 2 * ty + oy, pz = 2 * tz + oz;
    const int qx = wrap1(px - ex, NFx), qy = wrap1(py - ey, NFy), qz = wrap1(pz - ez, NFz);
    const int qbx = qx >> (T::LOG_RB + 1), qby = qy >> (T::LOG_RB + 1), qbz = qz >> (T::LOG_RB + 1);
    const int qs = G.block_slot[block_id(G, qbx, qby, qbz)];
    if (qs >= 0) {
      s += *at(gp + (size_t)i * G.pool_plane,
               pool_cell<RB, D>(qs, local_in_tile<RB>(px, qbx, NFx), local_in_tile<RB>(py, qby, NFy),
                                local_in_tile<RB>(pz, qbz, NFz))) + wi;
      nIn += 1.f;
    }
  }
  float& gci = *ppop<PP, Q>(P, PP ? parent_index<PP, RB>(G, P, cx, cy, cz) : c, i);
  const float fc = gci + wi;
  const float net = s - nIn * fc;
  if (net != 0.f) gci = (fc + net * T::INV_NCH) - wi;
}

// ---------------------------------------------------------------------------
// interp fill of newly refined tiles, and average/drain of dying ones
// ---------------------------------------------------------------------------
template <int Q, bool PP>
__device__ __forceinline__ void coarse_sample(const ParentRef& P, uint32_t cell, float& rho, float& ux, float& uy,
                                              float& uz, float (&fneq)[Q]) {
  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) g[i] = *ppop<PP, Q>(P, cell, i);
  const float F0[3] = {0.f, 0.f, 0.f};
  float drho;
  moments<Q>(g, F0, drho, rho, ux, uy, uz);
#pragma unroll
  for (int i = 0; i < Q; i++) fneq[i] = g[i] - geq<Q>(i, rho, ux, uy, uz);
}

template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_interp(ParentRef P, float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ pairs, int npairs,
         float rescale, int ring_only) {
  using T = TileQ<Q, RB>;
  constexpr int LO = T::LO, HI = T::HI, D = Lattice<Q>::D;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / (uint32_t)T::CELLS;
  if (k >= (uint32_t)npairs) return;
  const uint32_t r = t - k * (uint32_t)T::CELLS;
  const int id = pairs[1 + 2 * k], slot = pairs[2 + 2 * k];
  int fx, fy, fz;
  T::xyz(r, fx, fy, fz);
  // GHOST_ONLY (common_d3_amr_interp.wgsl:4-9): interior cells keep their
  // evolved state
  if (ring_only && fx >= LO && fx < HI && fy >= LO && fy < HI && fz >= LO && fz < HI) return;
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  // coarse-unit position origin - 1/4 + (j - GHOST)/2 (fineToCoarseUnit3)
  const float px = (float)(bx * RB) - 0.25f + 0.5f * (float)(fx - kGhost);
  const float py = (float)(by * RB) - 0.25f + 0.5f * (float)(fy - kGhost);
  const float pz = (float)(bz * RB) - 0.25f + 0.5f * (float)(fz - kGhost);
  const int c0x = (int)floorf(px), c0y = (int)floorf(py), c0z = D == 3 ? (int)floorf(pz) : 0;
  // bilinear in 2D: one parent layer, weight 1
  const float tx = px - c0x, ty = py - c0y, tz = D == 3 ? pz - c0z : 0.f;
  // A pool parent cell is read in the parent tile that OWNS it, at its
  // interior index (parentIndex, common_d3_parent_pool.wgsl:37-59), so a
  // tile filled mid-run sees solved parent cells rather than a parent ring
  // that only advects.  Where no parent tile owns a stencil cell (the outer
  // ring's stencil reaches one parent cell beyond the ring-parent cover),
  // the cell is read in the ring of the parent tile that holds this block,
  // which the stencil of a whole tile never leaves (parent-local [0, FB)).
  int ps = 0, pox = 0, poy = 0, poz = 0;
  if constexpr (PP) {
    constexpr int SH = Tile<RB>::LOG_RB + 1;
    const int pbx = (bx * RB) >> SH, pby = (by * RB) >> SH, pbz = (bz * RB) >> SH;
    ps = P.block_slot[(pbz * P.nby + pby) * P.nbx + pbx];
    if (ps < 0) return;  // a tile's parent block holds a tile
    pox = kGhost - (pbx << SH);
    poy = kGhost - (pby << SH);
    poz = kGhost - (pbz << SH);
  }
  float rho = 0.f, ux = 0.f, uy = 0.f, uz = 0.f, fneq[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) fneq[i] = 0.f;
  for (int s = 0; s < T::NCH; s++) {
    const int i = s & 1, j = (s >> 1) & 1, kk = (s >> 2) & 1;
    const float w = (i ? tx : 1.f - tx) * (j ? ty : 1.f - ty) * (kk ? tz : 1.f - tz);
    const int vx = wrap1(c0x + i, G.nx), vy = wrap1(c0y + j, G.ny), vz = wrap1(c0z + kk, G.nz);
    uint32_t cell = parent_index<PP, RB>(G, P, vx, vy, vz);
    if (PP && cell == kAbsent) cell = pool_cell<RB, D>(ps, c0x + i + pox, c0y + j + poy, c0z + kk + poz);
    float sr, sx, sy, sz, sf[Q];
    coarse_sample<Q, PP>(P, cell, sr, sx, sy, sz, sf);
    rho += w * sr;
    ux += w * sx;
    uy += w * sy;
    uz += w * sz;
#pragma unroll
    for (int q = 0; q < Q; q++) fneq[q] += w * sf[q];
  }
  const uint32_t cell = pool_cell<RB, D>(slot, fx, fy, fz);
#pragma unroll
  for (int q = 0; q < Q; q++) gp[(size_t)q * G.pool_plane + cell] = geq<Q>(q, rho, ux, uy, uz) + rescale * fneq[q];
}

template <int Q, int RB, bool PP>
__global__ void __launch_bounds__(256)
k_average(const float* __restrict__ gp, ParentRef P, PoolGeom G, const int32_t* __restrict__ pairs, int npairs,
          float rescale) {
  using T = TileQ<Q, RB>;
  constexpr int D = Lattice<Q>::D;
  constexpr uint32_t per = D == 3 ? RB * RB * RB : RB * RB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / per;
  if (k >= (uint32_t)npairs) return;
  const uint32_t r = t - k * per;
  const int id = pairs[1 + 2 * k], slot = pairs[2 + 2 * k];
  const int lx = r % RB, ly = (r / RB) % RB, lz = r / (RB * RB);
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  float rs = 0.f, mx = 0.f, my = 0.f, mz = 0.f, sneq[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) sneq[q] = 0.f;
  for (int c = 0; c < T::NCH; c++) {
    const uint32_t pc = pool_cell<RB, D>(slot, kGhost + 2 * lx + (c & 1), kGhost + 2 * ly + ((c >> 1) & 1),
                                         kGhost + 2 * lz + ((c >> 2) & 1));
    float g[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) g[q] = gp[(size_t)q * G.pool_plane + pc];
    const float F0[3] = {0.f, 0.f, 0.f};
    float drho, rho, ux, uy, uz;
    moments<Q>(g, F0, drho, rho, ux, uy, uz);
    rs += rho;
    mx += rho * ux;
    my += rho * uy;
    mz += rho * uz;
#pragma unroll
    for (int q = 0; q < Q; q++) sneq[q] += g[q] - geq<Q>(q, rho, ux, uy, uz);
  }
  const float rho = rs * T::INV_NCH;
  const float inv = 1.f / fmaxf(rs, 1e-6f);
  const float ux = mx * inv, uy = my * inv, uz = mz * inv;
  const uint32_t cc = parent_index<PP, RB>(G, P, bx * RB + lx, by * RB + ly, bz * RB + lz);
  if (PP && cc == kAbsent) return;  // a tile's parent cells lie in parent tiles
#pragma unroll
  for (int q = 0; q < Q; q++) *ppop<PP, Q>(P, cc, q) = geq<Q>(q, rho, ux, uy, uz) + rescale * (sneq[q] * T::INV_NCH);
  pmac_put<PP>(P, cc, make_float4(rho, ux, uy, uz));
}

// ---------------------------------------------------------------------------
// seam flux and its correction (the interp interface, opt-in)
// ---------------------------------------------------------------------------
// What the fine solver moved across the seam in one substep, read from the
// pool buffer that substep streams (common_d3_amr_flux.wgsl:131-193).  For
// each of the flux cell's 8 fine cells p and each direction: OUT, the
// population that streams to p from the interior of a refined tile; IN, the
// population that streams from p (in the receiving tile's own ring copy)
// into a refined tile.  [dm, dp] / 8 (V_fine / V_coarse) per list entry.
template <int Q, int RB>
__global__ void __launch_bounds__(256)
k_flux(const float* __restrict__ gp, float4* __restrict__ acc, PoolGeom G, const uint32_t* __restrict__ cells,
       uint32_t n, int accumulate) {
  using L = Lattice<Q>;
  using T = Tile<RB>;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t c = cells[t];
  const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
  const int NFx = 2 * G.nx, NFy = 2 * G.ny, NFz = 2 * G.nz;
  constexpr int SH = T::LOG_RB + 1;
  float dm = 0.f, px = 0.f, py = 0.f, pz = 0.f;
  for (int oc = 0; oc < 8; oc++) {
    const int p0 = 2 * cx + (oc & 1), p1 = 2 * cy + ((oc >> 1) & 1), p2 = 2 * cz + ((oc >> 2) & 1);
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
      const float* plane = gp + (size_t)i * G.pool_plane;
      const int qx = wrap1(p0 - ex, NFx), qy = wrap1(p1 - ey, NFy), qz = wrap1(p2 - ez, NFz);
      const int so = G.block_slot[block_id(G, qx >> SH, qy >> SH, qz >> SH)];
      if (so >= 0) {
        const float f = *at(plane, pool_cell<RB>(so, local_in_tile<RB>(qx, qx >> SH, NFx),
                                                 local_in_tile<RB>(qy, qy >> SH, NFy),
                                                 local_in_tile<RB>(qz, qz >> SH, NFz))) + L::w[i];
        dm += f;
        px += f * ex;
        py += f * ey;
        pz += f * ez;
      }
      const int tx = wrap1(p0 + ex, NFx), ty = wrap1(p1 + ey, NFy), tz = wrap1(p2 + ez, NFz);
      const int si = G.block_slot[block_id(G, tx >> SH, ty >> SH, tz >> SH)];
      if (si >= 0) {
        const float f = *at(plane, pool_cell<RB>(si, local_in_tile<RB>(p0, tx >> SH, NFx),
                                                 local_in_tile<RB>(p1, ty >> SH, NFy),
                                                 local_in_tile<RB>(p2, tz >> SH, NFz))) + L::w[i];
        dm -= f;
        px -= f * ex;
        py -= f * ey;
        pz -= f * ez;
      }
    }
  }
  float4 v = make_float4(dm * 0.125f, px * 0.125f, py * 0.125f, pz * 0.125f);
  if (accumulate) {
    const float4 o = acc[t];
    v = make_float4(o.x + v.x, o.y + v.y, o.z + v.z, o.w + v.w);
  }
  acc[t] = v;
}

// The correction onto the unrefined flux cells of the coarse state at t + 1
// (common_d3_amr_reflux.wgsl:105-164): delta = fine flux + (coarse_in -
// coarse_out), both coarse terms read from the state at t, applied as
// f_i += w_i (d_rho + 3 e_i . d_mom); then the cell's [rho, u] republished
// with the Guo half-force.  The applied delta is left in acc.
template <int Q, int RB>
__global__ void __launch_bounds__(256)
k_reflux(const float* __restrict__ gct, float* __restrict__ gct1, float4* __restrict__ mac, PoolGeom G,
         const uint32_t* __restrict__ cells, uint32_t n, float4* __restrict__ acc, float Fx, float Fy, float Fz) {
  using L = Lattice<Q>;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t c = cells[t];
  const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
  const float4 a = acc[t];
  float dm = a.x, px = a.y, py = a.z, pz = a.w;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
    // coarse_out: what this cell pulled out of the refined region
    const uint32_t s = coarse_cell(G, wrap1(cx - ex, G.nx), wrap1(cy - ey, G.ny), wrap1(cz - ez, G.nz));
    if (G.ccls[s] & kCovered) {
      const float v = gct[lat_elem(Q, i, s)] + L::w[i];
      dm -= v;
      px -= v * ex;
      py -= v * ey;
      pz -= v * ez;
    }
    // coarse_in: what the refined region pulled out of this cell
    if (covered_at(G, cx + ex, cy + ey, cz + ez)) {
      const float v = gct[lat_elem(Q, i, c)] + L::w[i];
      dm += v;
      px += v * ex;
      py += v * ey;
      pz += v * ez;
    }
  }
  acc[t] = make_float4(dm, px, py, pz);
  float rho = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
    float* gi = gct1 + lat_elem(Q, i, c);
    const float g = *gi + L::w[i] * (dm + 3.f * (ex * px + ey * py + ez * pz));
    *gi = g;
    const float f = g + L::w[i];
    rho += f;
    mx += f * ex;
    my += f * ey;
    mz += f * ez;
  }
  const float den = fmaxf(rho, 1e-6f);
  mac[c] = make_float4(rho, mx / den + Fx / (2.f * den), my / den + Fy / (2.f * den), mz / den + Fz / (2.f * den));
}

// ---------------------------------------------------------------------------
// criterion: max |curl u| over each block's L0 box dilated by `lead`, from
// the macroscopic field, central differences, each operation correctly
// rounded so the host restatement reproduces it bit for bit
// ---------------------------------------------------------------------------
__device__ __forceinline__ float3 mac_u(const float4* mac, const PoolGeom& G, int x, int y, int z) {
  const float4 m = mac[coarse_cell(G, ((x % G.nx) + G.nx) % G.nx, ((y % G.ny) + G.ny) % G.ny,
                                   ((z % G.nz) + G.nz) % G.nz)];
  return make_float3(m.y, m.z, m.w);
}

__device__ __forceinline__ float vort_mag(const float4* mac, const PoolGeom& G, int x, int y, int z) {
  const float3 xp = mac_u(mac, G, x + 1, y, z), xm = mac_u(mac, G, x - 1, y, z);
  const float3 yp = mac_u(mac, G, x, y + 1, z), ym = mac_u(mac, G, x, y - 1, z);
  const float3 zp = mac_u(mac, G, x, y, z + 1), zm = mac_u(mac, G, x, y, z - 1);
  // omega = (dUz/dy - dUy/dz, dUx/dz - dUz/dx, dUy/dx - dUx/dy)
  const float wx = sub_rn(mul_rn(sub_rn(yp.z, ym.z), 0.5f), mul_rn(sub_rn(zp.y, zm.y), 0.5f));
  const float wy = sub_rn(mul_rn(sub_rn(zp.x, zm.x), 0.5f), mul_rn(sub_rn(xp.z, xm.z), 0.5f));
  const float wz = sub_rn(mul_rn(sub_rn(xp.y, xm.y), 0.5f), mul_rn(sub_rn(yp.x, ym.x), 0.5f));
  return sqrt_rn(add_rn(add_rn(mul_rn(wx, wx), mul_rn(wy, wy)), mul_rn(wz, wz)));
}

// Q = 1/2 (|Omega|^2 - |S|^2) at an L0 cell from the same centred gradient,
// J[j][i] = du_i/dx_j, summed over (i, j) in the order of qAt
// (common_d3_criterion.wgsl:102-120; host twin qOfGrad, d3-criterion.mjs).
// Pure shear gives exactly 0, solid rotation at rate W gives W^2.
__device__ __forceinline__ float q_crit(const float4* mac, const PoolGeom& G, int x, int y, int z) {
  const float3 xp = mac_u(mac, G, x + 1, y, z), xm = mac_u(mac, G, x - 1, y, z);
  const float3 yp = mac_u(mac, G, x, y + 1, z), ym = mac_u(mac, G, x, y - 1, z);
  const float3 zp = mac_u(mac, G, x, y, z + 1), zm = mac_u(mac, G, x, y, z - 1);
  const float J[3][3] = {
      {mul_rn(sub_rn(xp.x, xm.x), 0.5f), mul_rn(sub_rn(xp.y, xm.y), 0.5f), mul_rn(sub_rn(xp.z, xm.z), 0.5f)},
      {mul_rn(sub_rn(yp.x, ym.x), 0.5f), mul_rn(sub_rn(yp.y, ym.y), 0.5f), mul_rn(sub_rn(yp.z, ym.z), 0.5f)},
      {mul_rn(sub_rn(zp.x, zm.x), 0.5f), mul_rn(sub_rn(zp.y, zm.y), 0.5f), mul_rn(sub_rn(zp.z, zm.z), 0.5f)}};
  float s2 = 0.f, om2 = 0.f;
#pragma unroll
  for (int i = 0; i < 3; i++) {
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const float a = J[j][i], b = J[i][j];
      const float S = mul_rn(0.5f, add_rn(a, b)), O = mul_rn(0.5f, sub_rn(a, b));
      s2 = add_rn(s2, mul_rn(S, S));
      om2 = add_rn(om2, mul_rn(O, O));
    }
  }
  return mul_rn(0.5f, sub_rn(om2, s2));
}

// One block per AMR block: the maximum of |curl u| (kind 0) or Q (kind 1)
// over the block's L0 box dilated by `lead` cells (reduceQ,
// common_d3_criterion.wgsl:122-158).
__global__ void __launch_bounds__(64) k_criterion(const float4* __restrict__ mac, float* __restrict__ crit, PoolGeom G,
                                                  int lead, int kind) {
  const int b = blockIdx.x;
  const int bx = b % G.nbx, by = (b / G.nbx) % G.nby, bz = b / (G.nbx * G.nby);
  // a grid one cell deep (the 2D pool) takes its single layer
  const int n = G.rb + 2 * lead, nzb = G.nz == 1 ? 1 : n;
  const int x0 = bx * G.rb - lead, y0 = by * G.rb - lead, z0 = G.nz == 1 ? 0 : bz * G.rb - lead;
  const int total = n * n * nzb;
  float best = -1e30f;
  for (int k = threadIdx.x; k < total; k += 64) {
    const int kx = k % n, ky = (k / n) % n, kz = k / (n * n);
    best = fmaxf(best, kind ? q_crit(mac, G, x0 + kx, y0 + ky, z0 + kz) : vort_mag(mac, G, x0 + kx, y0 + ky, z0 + kz));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) best = fmaxf(best, __shfl_xor(best, off, 64));
  if (threadIdx.x == 0) crit[b] = best;
}

// ---------------------------------------------------------------------------
// manager: one workgroup of 1024 threads, ordered compactions (deterministic)
// ---------------------------------------------------------------------------
// Exclusive positions of the flagged indices in [0, n), ascending; `emit(i,
// pos)` runs for every flagged i.  Returns the flagged count.  All threads
// of the block must call it.
template <class Pred, class Emit>
__device__ int ordered_compact(int n, Pred pred, Emit emit, int* scan) {
  const int tid = threadIdx.x;
  int base = 0;
  for (int start = 0; start < n; start += 1024) {
    const int i = start + tid;
    const int flag = (i < n && pred(i)) ? 1 : 0;
    scan[tid] = flag;
    __syncthreads();
    // Hillis-Steele inclusive scan, ping-pong halves of `scan`
    int src = 0;
    for (int off = 1; off < 1024; off <<= 1) {
      const int v = scan[src * 1024 + tid] + (tid >= off ? scan[src * 1024 + tid - off] : 0);
      scan[(1 - src) * 1024 + tid] = v;
      src = 1 - src;
      __syncthreads();
    }
    const int incl = scan[src * 1024 + tid];
    if (flag) emit(i, base + incl - 1);
    const int chunk = scan[src * 1024 + 1023];
    __syncthreads();
    base += chunk;
  }
  return base;
}

__global__ void __launch_bounds__(1024) k_manage(ManageArgs a) {
  __shared__ int scan[2048];
  __shared__ int nfree_s;
  const int tid = threadIdx.x;
  // 1. decide, with hysteresis (common_d3_manage.wgsl decide: a held block
  //    is judged against the lower coarsen threshold), the body's geometry
  //    OR the field (:443-462).  Decided beforehand at depth; there the
  //    geometry mask only counts wanted blocks in the sponge band.
  if (!a.decided) {
    for (int b = tid; b < a.nblocks; b += 1024) {
      const bool held = a.block_slot[b] >= 0;
      const float c = a.crit[b];
      const uint8_t gm = a.geom ? a.geom[b] : 0;
      const bool field = !(gm & 2) && (held ? c >= a.keep_at : c >= a.refine_at);
      a.want[b] = ((gm & 1) || field) ? 1 : 0;
    }
  } else if (a.geom) {
    int n = 0;
    for (int b = tid; b < a.nblocks; b += 1024) n += (a.want[b] && (a.geom[b] & 2)) ? 1 : 0;
    if (n) atomicAdd(&a.status[1], n);
  }
  __syncthreads();
  // 2. dying blocks (held, no longer wanted), ascending
  const int ndying = ordered_compact(
      a.nblocks, [&](int b) { return a.block_slot[b] >= 0 && !a.want[b]; },
      [&](int b, int pos) { a.dying[1 + 2 * pos] = b; a.dying[2 + 2 * pos] = a.block_slot[b]; }, scan);
  __syncthreads();
  // 3. coarsen: release their slots
  for (int b = tid; b < a.nblocks; b += 1024) {
    const int s = a.block_slot[b];
    if (s >= 0 && !a.want[b]) {
      a.slot_to_block[s] = -1;
      a.block_slot[b] = -1;
    }
  }
  __syncthreads();
  // 4. free slots, ascending
  const int nfree = ordered_compact(
      a.max_slots, [&](int s) { return a.slot_to_block[s] < 0; }, [&](int s, int pos) { a.free_list[pos] = s; },
      scan);
  if (tid == 0) nfree_s = nfree;
  __syncthreads();
  // 5. refine: the k-th newly wanted block (ascending id) takes the k-th free slot
  const int nwant = ordered_compact(
      a.nblocks, [&](int b) { return a.want[b] && a.block_slot[b] < 0; },
      [&](int b, int pos) {
        if (pos < nfree_s) {
          const int s = a.free_list[pos];
          a.block_slot[b] = s;
          a.slot_to_block[s] = b;
          a.fresh[1 + 2 * pos] = b;
          a.fresh[2 + 2 * pos] = s;
        }
      },
      scan);
  __syncthreads();
  // 6. slots in use, ascending
  const int nactive = ordered_compact(
      a.max_slots, [&](int s) { return a.slot_to_block[s] >= 0; }, [&](int s, int pos) { a.active[1 + pos] = s; },
      scan);
  if (tid == 0) {
    a.dying[0] = ndying;
    a.fresh[0] = min(nwant, nfree);
    a.active[0] = nactive;
    a.status[0] += max(0, nwant - nfree);
  }
}

// ---------------------------------------------------------------------------
// the manager at depth: every level's want, then the 2:1 closure finest-first
// (common_d3_manage.wgsl decide / completeOctets / balance), each pass one
// thread per output block reading what it does not write
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_decide(DecideArgs a) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= a.nbx * a.nby * a.nbz) return;
  const uint8_t gm = a.geom ? a.geom[b] : 0;
  // the geometry first: a block the body needs is never put to the field
  bool w = (gm & 1) != 0;
  if (!w && a.crit && !(gm & 2)) {
    const int bx = b % a.nbx, by = (b / a.nbx) % a.nby, bz = b / (a.nbx * a.nby);
    const float c = a.crit[((bz >> a.shift) * a.cnby + (by >> a.shift)) * a.cnbx + (bx >> a.shift)];
    w = a.block_slot[b] >= 0 ? c >= a.keep_at : c >= a.refine_at;
  }
  a.want[b] = w ? 1 : 0;
}

__global__ void __launch_bounds__(256) k_complete_octets(uint8_t* want, int nbx, int nby, int nbz) {
  const int o = blockIdx.x * blockDim.x + threadIdx.x;
  const int ox = nbx >> 1, oy = nby >> 1, oz = nbz >> 1;
  if (o >= ox * oy * oz) return;
  const int x = 2 * (o % ox), y = 2 * ((o / ox) % oy), z = 2 * (o / (ox * oy));
  uint8_t any = 0;
  for (int k = 0; k < 8; k++) any |= want[((z + (k >> 2)) * nby + y + ((k >> 1) & 1)) * nbx + x + (k & 1)];
  for (int k = 0; k < 8; k++) want[((z + (k >> 2)) * nby + y + ((k >> 1) & 1)) * nbx + x + (k & 1)] = any;
}

__global__ void __launch_bounds__(256) k_balance(uint8_t* wp, int pnbx, int pnby, int pnbz, const uint8_t* wc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pnbx * pnby * pnbz) return;
  const int px = p % pnbx, py = (p / pnbx) % pnby, pz = p / (pnbx * pnby);
  const int cnx = 2 * pnbx, cny = 2 * pnby, cnz = 2 * pnbz;
  uint8_t any = 0;
  for (int k = 0; k < 64 && !any; k++) {
    const int cx = wrap1(2 * px + (k & 3) - 1, cnx);
    const int cy = wrap1(2 * py + ((k >> 2) & 3) - 1, cny);
    const int cz = wrap1(2 * pz + (k >> 4) - 1, cnz);
    any = wc[(cz * cny + cy) * cnx + cx];
  }
  if (any) wp[p] = 1;
}

// Ordered active-slot list only (after a topology upload).
__global__ void __launch_bounds__(1024) k_active_list(const int32_t* slot_to_block, int max_slots, int32_t* active) {
  __shared__ int scan[2048];
  const int n = ordered_compact(
      max_slots, [&](int s) { return slot_to_block[s] >= 0; }, [&](int s, int pos) { active[1 + pos] = s; }, scan);
  if (threadIdx.x == 0) active[0] = n;
}

// [count, (block, slot) ...] of the active list, in its order.
__global__ void __launch_bounds__(256) k_active_pairs(const int32_t* active, const int32_t* slot_to_block,
                                                      int32_t* pairs) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = active[0];
  if (k == 0) pairs[0] = n;
  if (k < n) {
    const int s = active[1 + k];
    pairs[1 + 2 * k] = slot_to_block[s];
    pairs[2 + 2 * k] = s;
  }
}

// ---------------------------------------------------------------------------
// FP64 totals of the composite state: unrefined L0 cells + interior fine
// cells weighted 1/8 (a fine cell is 1/8 of a coarse cell's volume)
// ---------------------------------------------------------------------------
// Adds weight x [f, f ex, f ey, f ez] over the interior cells of a pool
// level's active tiles that no finer tile covers (grid-stride over the
// launch).
template <int Q, int RB>
__device__ __forceinline__ void level_sums(const float* __restrict__ gp, const PoolGeom& G,
                                           const int32_t* __restrict__ active, double weight, double (&v)[4]) {
  using L = Lattice<Q>;
  constexpr int D = L::D;
  constexpr uint32_t rb2 = 2 * RB, per = D == 3 ? rb2 * rb2 * rb2 : rb2 * rb2;
  const uint32_t stride = gridDim.x * blockDim.x;
  const int nact = active[0];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < (uint32_t)nact * per; t += stride) {
    const uint32_t k = t / per, r = t - k * per;
    const int slot = active[1 + k];
    const int lx = r % rb2, ly = (r / rb2) % rb2, lz = r / (rb2 * rb2);
    if (G.child_block_slot) {
      const int id = G.slot_to_block[slot];
      const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
      if (covered_by_child<RB>(G, bx * (int)rb2 + lx, by * (int)rb2 + ly, bz * (int)rb2 + lz)) continue;
    }
    const uint32_t pc = pool_cell<RB, D>(slot, kGhost + lx, kGhost + ly, kGhost + lz);
    for (int i = 0; i < Q; i++) {
      const double f = weight * ((double)gp[(size_t)i * G.pool_plane + pc] + (double)L::w[i]);
      v[0] += f; v[1] += L::ex[i] * f; v[2] += L::ey[i] * f; v[3] += L::ez[i] * f;
    }
  }
}

template <int Q, int RB>
__global__ void __launch_bounds__(256)
k_level_totals(const float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ active, double weight,
               double* partials, unsigned* counter, double* out) {
  __shared__ double lds[16 * 4];
  double v[4] = {0, 0, 0, 0};
  level_sums<Q, RB>(gp, G, active, weight, v);
  block_sum<4>(v, lds);
  partials_then_fixed_order_sum<4>(v, partials, counter, [&](int j, double s) { out[j] = s; });
}

template <int Q, int RB>
__global__ void __launch_bounds__(256)
k_pool_totals(const float* __restrict__ gc, const float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ active,
              double* partials, unsigned* counter, double* out) {
  using L = Lattice<Q>;
  __shared__ double lds[16 * 4];
  double v[4] = {0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < G.ncells; c += stride) {
    if (G.ccls[c] & kCovered) continue;
    for (int i = 0; i < Q; i++) {
      const double f = (double)gc[lat_elem(Q, i, c)] + (double)L::w[i];
      v[0] += f; v[1] += L::ex[i] * f; v[2] += L::ey[i] * f; v[3] += L::ez[i] * f;
    }
  }
  level_sums<Q, RB>(gp, G, active, Lattice<Q>::D == 3 ? 0.125 : 0.25, v);
  block_sum<4>(v, lds);
  partials_then_fixed_order_sum<4>(v, partials, counter, [&](int j, double s) { out[j] = s; });
}

// ---------------------------------------------------------------------------
// launches: dispatch on (q, rb)
// ---------------------------------------------------------------------------
static inline dim3 grid_for(uint64_t n, unsigned rows = 1) { return dim3((unsigned)((n + 255) / 256), rows); }
// parent cells under one block (rb^D)
template <int Q, int RB>
static constexpr uint64_t parent_cells() { return Lattice<Q>::D == 3 ? (uint64_t)RB * RB * RB : (uint64_t)RB * RB; }

// The passes of the two-level explode interface also run the D2Q9 quadtree.
template <template <int, int> class Op, class... Args>
static hipError_t dispatch_q_rb(int q, int rb, Args&&... args) {
  switch (q * 100 + rb) {
    case 904: return Op<9, 4>::run(args...);
    case 908: return Op<9, 8>::run(args...);
    case 916: return Op<9, 16>::run(args...);
    case 1904: return Op<19, 4>::run(args...);
    case 1908: return Op<19, 8>::run(args...);
    case 1916: return Op<19, 16>::run(args...);
    case 2704: return Op<27, 4>::run(args...);
    case 2708: return Op<27, 8>::run(args...);
    case 2716: return Op<27, 16>::run(args...);
    default: return hipErrorInvalidValue;
  }
}
// The 3D-only passes (the body's force, the seam flux, levels >= 3).
template <template <int, int> class Op, class... Args>
static hipError_t dispatch_q3_rb(int q, int rb, Args&&... args) {
  switch (q * 100 + rb) {
    case 1904: return Op<19, 4>::run(args...);
    case 1908: return Op<19, 8>::run(args...);
    case 1916: return Op<19, 16>::run(args...);
    case 2704: return Op<27, 4>::run(args...);
    case 2708: return Op<27, 8>::run(args...);
    case 2716: return Op<27, 16>::run(args...);
    default: return hipErrorInvalidValue;
  }
}

// Runs f(std::true_type{}) for a pool parent, f(std::false_type{}) for the
// dense one.
template <class F>
static void with_parent(const ParentRef& p, F f) {
  if (p.block_slot) f(std::true_type{});
  else f(std::false_type{});
}

template <int Q, int RB>
struct OpCellClass {
  static hipError_t run(uint8_t* ccls, const PoolGeom& g, const SurfacePlan& plan, const int32_t* active, int nactive,
                        const ParentRef& P, hipStream_t s) {
    with_parent(P, [&](auto PP) {
      constexpr bool pp = decltype(PP)::value;
      hipLaunchKernelGGL((k_cell_class<Q, RB, pp>), grid_for(g.ncells), dim3(256), 0, s, ccls, g, plan, P);
      if (nactive > 0)
        hipLaunchKernelGGL((k_ring_plan<Q, RB, pp>), grid_for((uint64_t)nactive * TileQ<Q, RB>::RING), dim3(256), 0, s, g,
                           plan, active, nactive, P);
    });
    if (nactive > 0)
      hipLaunchKernelGGL(k_tile_table, grid_for((uint64_t)nactive), dim3(256), 0, s, g, active, nactive,
                         const_cast<int32_t*>(g.tile_table));
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpPoolStep {
  static hipError_t run(const float* gin, float* gout, float* mac_pool, const PoolGeom& g, const CollideParams& p,
                        int nactive, int collide_ring, int sub, hipStream_t s) {
    const dim3 grid = grid_for((uint64_t)nactive * TileQ<Q, RB>::CELLS);
    // a level without a body force runs the instantiation without Guo terms
    const bool frc = p.F[0] != 0.f || p.F[1] != 0.f || p.F[2] != 0.f;
    auto go = [&](auto B, auto F) {
      hipLaunchKernelGGL((k_pool_step<Q, RB, decltype(B)::value, decltype(F)::value>), grid, dim3(256), 0, s, gin,
                         gout, mac_pool, g, p, nactive, collide_ring, sub);
    };
    if (g.body.on) {
      if (frc) go(std::true_type{}, std::true_type{});
      else go(std::true_type{}, std::false_type{});
    } else {
      if (frc) go(std::false_type{}, std::true_type{});
      else go(std::false_type{}, std::false_type{});
    }
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpPoolForce {
  static hipError_t run(const float* gp, const PoolGeom& g, const int32_t* active, int nactive, double* partials,
                        hipStream_t s) {
    hipLaunchKernelGGL((k_pool_force<Q, RB>), dim3(pool_force_blocks(RB, nactive)), dim3(256), 0, s, gp, g, active,
                       nactive, partials);
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpFlux {
  static hipError_t run(const float* gp, float* acc, const PoolGeom& g, const SurfacePlan& plan, uint32_t n,
                        int accumulate, hipStream_t s) {
    hipLaunchKernelGGL((k_flux<Q, RB>), grid_for(n), dim3(256), 0, s, gp, reinterpret_cast<float4*>(acc), g, plan.flux,
                       n, accumulate);
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpReflux {
  static hipError_t run(const float* gct, float* gct1, float* mac, const PoolGeom& g, const SurfacePlan& plan,
                        uint32_t n, float* acc, const float* F, hipStream_t s) {
    hipLaunchKernelGGL((k_reflux<Q, RB>), grid_for(n), dim3(256), 0, s, gct, gct1, reinterpret_cast<float4*>(mac), g,
                       plan.flux, n, reinterpret_cast<float4*>(acc), F[0], F[1], F[2]);
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpExplode {
  static hipError_t run(const ParentRef& P, float* gp, const PoolGeom& g, const SurfacePlan& plan, uint32_t nring,
                        int linear, int dirs, hipStream_t s) {
    // about 2^20 threads: every direction its own thread for a small ring,
    // all Q in one thread for a ring of a million cells
    const int groups = (int)std::min<uint64_t>(Q, std::max<uint64_t>(1, ((1u << 20) + nring - 1) / nring));
    const int per = dirs > 0 ? std::min(dirs, Q) : (Q + groups - 1) / groups;
    with_parent(P, [&](auto PP) {
      hipLaunchKernelGGL((k_explode<Q, RB, decltype(PP)::value>), grid_for(nring, (Q + per - 1) / per), dim3(256), 0,
                         s, P, gp, g, plan.ring, nring, linear, per);
    });
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpCoalesce {
  static hipError_t run(const float* gp, const ParentRef& P, const float* mac_pool, const PoolGeom& g,
                        const int32_t* active, int nactive, const SurfacePlan& plan, uint32_t nseam, uint32_t nnear,
                        int orphans, hipStream_t s) {
    with_parent(P, [&](auto PP) {
      constexpr bool pp = decltype(PP)::value;
      if (nactive > 0)
        hipLaunchKernelGGL((k_coalesce_mac<Q, RB, pp>), grid_for((uint64_t)nactive * parent_cells<Q, RB>()), dim3(256), 0, s, P,
                           mac_pool, g, active, nactive);
      if (nseam > 0)
        hipLaunchKernelGGL((k_coalesce_seam<Q, RB, pp>), grid_for(nseam, Q - 1), dim3(256), 0, s, gp, P, g, plan.seam,
                           nseam);
      if (orphans && nnear > 0)
        hipLaunchKernelGGL((k_orphans<Q, RB, pp>), grid_for(nnear, Q - 1), dim3(256), 0, s, gp, P, g, plan.near,
                           nnear);
    });
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpInterp {
  static hipError_t run(const ParentRef& P, float* gp, const PoolGeom& g, const int32_t* pairs, int npairs,
                        float rescale, int ring_only, hipStream_t s) {
    with_parent(P, [&](auto PP) {
      hipLaunchKernelGGL((k_interp<Q, RB, decltype(PP)::value>), grid_for((uint64_t)npairs * TileQ<Q, RB>::CELLS),
                         dim3(256), 0, s, P, gp, g, pairs, npairs, rescale, ring_only);
    });
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpAverage {
  static hipError_t run(const float* gp, const ParentRef& P, const PoolGeom& g, const int32_t* pairs, int npairs,
                        float rescale, hipStream_t s) {
    with_parent(P, [&](auto PP) {
      hipLaunchKernelGGL((k_average<Q, RB, decltype(PP)::value>), grid_for((uint64_t)npairs * parent_cells<Q, RB>()),
                         dim3(256), 0, s, gp, P, g, pairs, npairs, rescale);
    });
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpTotals {
  static hipError_t run(const float* gc, const float* gp, const PoolGeom& g, const int32_t* active, double* partials,
                        unsigned* counter, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_pool_totals<Q, RB>), dim3(256), dim3(256), 0, s, gc, gp, g, active, partials, counter, out);
    return hipGetLastError();
  }
};
template <int Q, int RB>
struct OpLevelTotals {
  static hipError_t run(const float* gp, const PoolGeom& g, const int32_t* active, double weight, double* partials,
                        unsigned* counter, double* out, hipStream_t s) {
    hipLaunchKernelGGL((k_level_totals<Q, RB>), dim3(256), dim3(256), 0, s, gp, g, active, weight, partials, counter,
                       out);
    return hipGetLastError();
  }
};

hipError_t launch_cell_class(int q, uint8_t* ccls, const PoolGeom& g, const SurfacePlan& plan, const int32_t* active,
                             int nactive, const ParentRef& parent, hipStream_t s) {
  return dispatch_q_rb<OpCellClass>(q, g.rb, ccls, g, plan, active, nactive, parent, s);
}

hipError_t launch_pool_step(int q, const float* gin, float* gout, float* mac_pool, const PoolGeom& g,
                            const CollideParams& fine, int nactive, int collide_ring, int sub, hipStream_t s) {
  if (nactive <= 0) return hipSuccess;
  return dispatch_q_rb<OpPoolStep>(q, g.rb, gin, gout, mac_pool, g, fine, nactive, collide_ring, sub, s);
}
uint32_t pool_force_blocks(int rb, int nactive) {
  const uint64_t per = (uint64_t)(2 * rb) * (2 * rb) * (2 * rb);
  return (uint32_t)std::max<uint64_t>(1, ((uint64_t)std::max(0, nactive) * per + 255) / 256);
}
hipError_t launch_pool_force(int q, const float* gpool, const PoolGeom& g, const int32_t* active, int nactive,
                             double* partials, hipStream_t s) {
  return dispatch_q3_rb<OpPoolForce>(q, g.rb, gpool, g, active, nactive, partials, s);
}
hipError_t launch_flux(int q, const float* gp, float* acc, const PoolGeom& g, const SurfacePlan& plan, uint32_t nflux,
                       int accumulate, hipStream_t s) {
  if (nflux == 0) return hipSuccess;
  return dispatch_q3_rb<OpFlux>(q, g.rb, gp, acc, g, plan, nflux, accumulate, s);
}
hipError_t launch_reflux(int q, const float* gct, float* gct1, float* mac, const PoolGeom& g, const SurfacePlan& plan,
                         uint32_t nflux, float* acc, const float* force, hipStream_t s) {
  if (nflux == 0) return hipSuccess;
  return dispatch_q3_rb<OpReflux>(q, g.rb, gct, gct1, mac, g, plan, nflux, acc, force, s);
}

hipError_t launch_explode(int q, const ParentRef& parent, float* gp, const PoolGeom& g, const SurfacePlan& plan,
                          uint32_t nring, int linear, int dirs, hipStream_t s) {
  if (nring == 0) return hipSuccess;
  return dispatch_q_rb<OpExplode>(q, g.rb, parent, gp, g, plan, nring, linear, dirs, s);
}

hipError_t launch_coalesce(int q, const float* gp, const ParentRef& parent, const float* mac_pool, const PoolGeom& g,
                           const int32_t* active, int nactive, const SurfacePlan& plan, uint32_t nseam,
                           uint32_t nnear, int orphans, hipStream_t s) {
  return dispatch_q_rb<OpCoalesce>(q, g.rb, gp, parent, mac_pool, g, active, nactive, plan, nseam, nnear, orphans,
                                   s);
}

hipError_t launch_interp(int q, const ParentRef& parent, float* gp, const PoolGeom& g, const int32_t* pairs,
                         int npairs, float rescale, int ring_only, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  return dispatch_q_rb<OpInterp>(q, g.rb, parent, gp, g, pairs, npairs, rescale, ring_only, s);
}

hipError_t launch_average(int q, const float* gp, const ParentRef& parent, const PoolGeom& g, const int32_t* pairs,
                          int npairs, float rescale, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  return dispatch_q_rb<OpAverage>(q, g.rb, gp, parent, g, pairs, npairs, rescale, s);
}

hipError_t launch_criterion(const float* mac, float* crit, const PoolGeom& g, float lead, int kind, hipStream_t s) {
  hipLaunchKernelGGL(k_criterion, dim3(g.nbx * g.nby * g.nbz), dim3(64), 0, s, (const float4*)mac, crit, g,
                     (int)lead, kind);
  return hipGetLastError();
}

hipError_t launch_manage(const ManageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_manage, dim3(1), dim3(1024), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_decide(const DecideArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_decide, grid_for((uint64_t)a.nbx * a.nby * a.nbz), dim3(256), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_complete_octets(uint8_t* want, int nbx, int nby, int nbz, hipStream_t s) {
  hipLaunchKernelGGL(k_complete_octets, grid_for((uint64_t)(nbx / 2) * (nby / 2) * (nbz / 2)), dim3(256), 0, s,
                     want, nbx, nby, nbz);
  return hipGetLastError();
}

hipError_t launch_balance(uint8_t* want_parent, int pnbx, int pnby, int pnbz, const uint8_t* want_child,
                          hipStream_t s) {
  hipLaunchKernelGGL(k_balance, grid_for((uint64_t)pnbx * pnby * pnbz), dim3(256), 0, s, want_parent, pnbx, pnby,
                     pnbz, want_child);
  return hipGetLastError();
}

hipError_t launch_active_list(const int32_t* slot_to_block, int max_slots, int32_t* active, hipStream_t s) {
  hipLaunchKernelGGL(k_active_list, dim3(1), dim3(1024), 0, s, slot_to_block, max_slots, active);
  return hipGetLastError();
}

hipError_t launch_active_pairs(const int32_t* active, const int32_t* slot_to_block, int max_slots, int32_t* pairs,
                               hipStream_t s) {
  hipLaunchKernelGGL(k_active_pairs, grid_for((uint64_t)max_slots), dim3(256), 0, s, active, slot_to_block, pairs);
  return hipGetLastError();
}

hipError_t launch_pool_totals(int q, const float* gc, const float* gp, const PoolGeom& g, const int32_t* active,
                              double* partials, unsigned* counter, double* out, hipStream_t s) {
  return dispatch_q_rb<OpTotals>(q, g.rb, gc, gp, g, active, partials, counter, out, s);
}
hipError_t launch_level_totals(int q, const float* gp, const PoolGeom& g, const int32_t* active, double weight,
                               double* partials, unsigned* counter, double* out, hipStream_t s) {
  return dispatch_q3_rb<OpLevelTotals>(q, g.rb, gp, g, active, weight, partials, counter, out, s);
}

}  // namespace lbm

// Block-pool (AMR) kernels for gfx950, two levels (dense L0 + one pool of
// fine tiles).  Semantics follow the reference's pool shaders:
//   pool step      shaders/common_d3_amr_step1.wgsl:160-405 (DIRECT_GHOST
//                  gather, ring cells advect only, interior collides at
//                  omega_fine)
//   explode        shaders/common_d3_amr_explode.wgsl:98-214
//   coalesce       shaders/common_d3_amr_coalesce.wgsl:126-285 (+ orphans)
//   interp fill    shaders/common_d3_amr_interp.wgsl:142-185 with
//                  interpCoarseToFine3, common_d3_pool.wgsl:192-220
//   average/drain  shaders/common_d3_amr_average.wgsl:75-154
//   criterion      shaders/common_d3_criterion.wgsl:122-158, with the
//                  vorticity magnitude |curl u| in place of Q
//   manager        shaders/common_d3_manage.wgsl:430-546 and
//                  amr_active_list.wgsl:41-70, re-designed to be
//                  deterministic: ordered compactions in one workgroup
//                  instead of an atomic free list.
// Populations are stored shifted (g = f - w); the surface passes convert to
// f where their arithmetic is not linear in f with unit total weight.
#include <hip/hip_runtime.h>

#include "amr.h"
#include "collide.h"
#include "device_common.h"
#include "lattice_pairs.h"
#include "lattice_tables.h"

namespace lbm {

__device__ __forceinline__ int wrapi(int v, int n) { return ((v % n) + n) % n; }

__device__ __forceinline__ int block_id(const PoolGeom& G, int bx, int by, int bz) {
  return (bz * G.nby + by) * G.nbx + bx;
}
__device__ __forceinline__ bool coarse_refined(const PoolGeom& G, int cx, int cy, int cz) {
  return G.block_slot[block_id(G, cx / G.rb, cy / G.rb, cz / G.rb)] >= 0;
}
__device__ __forceinline__ uint32_t coarse_cell(const PoolGeom& G, int x, int y, int z) {
  return (uint32_t)((z * G.ny + y) * G.nx + x);
}
__device__ __forceinline__ uint32_t pool_cell(const PoolGeom& G, int slot, int fx, int fy, int fz) {
  return (uint32_t)(((slot * G.fb + fz) * G.fb + fy) * G.fb + fx);
}
// Tile-local index of global fine coordinate gcoord relative to block b on
// one axis, taking the periodic image nearest the block (localInTile,
// common_d3_pool.wgsl:75-85).
__device__ __forceinline__ int local_in_tile(int gcoord, int b, int nfine, int rb2) {
  int d = gcoord - b * rb2;
  if (d < -nfine / 2) d += nfine;
  if (d > nfine / 2) d -= nfine;
  return d + kGhost;
}

// ---------------------------------------------------------------------------
// pool step
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(256)
k_pool_step(const float* __restrict__ gin, float* __restrict__ gout, float* __restrict__ mac_pool, PoolGeom G,
            CollideParams P, const int32_t* __restrict__ active, int nactive) {
  using L = Lattice<Q>;
  const int FB = G.fb;
  const uint32_t tile = (uint32_t)FB * FB * FB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / tile;
  if (k >= (uint32_t)nactive) return;
  const uint32_t r = t - k * tile;
  const int slot = active[1 + k];
  const int fx = r % FB, fy = (r / FB) % FB, fz = r / (FB * FB);
  const int id = G.slot_to_block[slot];
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  const int RB2 = 2 * G.rb;
  const int lo = kGhost, hi = kGhost + RB2;
  const int offx = fx <= lo ? -1 : (fx + 1 >= hi ? 1 : 0);
  const int offy = fy <= lo ? -1 : (fy + 1 >= hi ? 1 : 0);
  const int offz = fz <= lo ? -1 : (fz + 1 >= hi ? 1 : 0);
  const int nbxw = wrapi(bx + offx, G.nbx), nbyw = wrapi(by + offy, G.nby), nbzw = wrapi(bz + offz, G.nbz);
  int nX = -1, nY = -1, nZ = -1, nXY = -1, nXZ = -1, nYZ = -1, nXYZ = -1;
  if (offx) nX = G.block_slot[block_id(G, nbxw, by, bz)];
  if (offy) nY = G.block_slot[block_id(G, bx, nbyw, bz)];
  if (offz) nZ = G.block_slot[block_id(G, bx, by, nbzw)];
  if (offx && offy) nXY = G.block_slot[block_id(G, nbxw, nbyw, bz)];
  if (offx && offz) nXZ = G.block_slot[block_id(G, nbxw, by, nbzw)];
  if (offy && offz) nYZ = G.block_slot[block_id(G, bx, nbyw, nbzw)];
  if (offx && offy && offz) nXYZ = G.block_slot[block_id(G, nbxw, nbyw, nbzw)];

  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) {
    int sx = fx - L::ex[i], sy = fy - L::ey[i], sz = fz - L::ez[i];
    const int ox = (sx < lo || sx >= hi) ? offx : 0;
    const int oy = (sy < lo || sy >= hi) ? offy : 0;
    const int oz = (sz < lo || sz >= hi) ? offz : 0;
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
      src_slot = ns;
      sx -= ox * RB2;
      sy -= oy * RB2;
      sz -= oz * RB2;
    }
    sx = min(max(sx, 0), FB - 1);
    sy = min(max(sy, 0), FB - 1);
    sz = min(max(sz, 0), FB - 1);
    g[i] = gin[(size_t)i * G.pool_plane + pool_cell(G, src_slot, sx, sy, sz)];
  }

  float drho, rho, ux, uy, uz;
  moments<Q>(g, P.F, drho, rho, ux, uy, uz);
  const uint32_t cell = pool_cell(G, slot, fx, fy, fz);
  mac_pool[cell] = rho;
  mac_pool[(size_t)G.pool_plane + cell] = ux;
  mac_pool[2 * (size_t)G.pool_plane + cell] = uy;
  mac_pool[3 * (size_t)G.pool_plane + cell] = uz;
  const bool interior = fx >= lo && fx < hi && fy >= lo && fy < hi && fz >= lo && fz < hi;
  // Ring cells advect and store without colliding (common_d3_amr_step1.wgsl
  // COLLIDE_RING = 0): coalesce averages uncollided interface states.
  if (interior) collide<Q>(g, P, drho, rho, ux, uy, uz);
#pragma unroll
  for (int i = 0; i < Q; i++) gout[(size_t)i * G.pool_plane + cell] = g[i];
}

// ---------------------------------------------------------------------------
// explode: coarse state at t -> ring of the pool buffer the first substep reads
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(256)
k_explode(const float* __restrict__ gc, float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ active,
          int nactive, int linear) {
  using L = Lattice<Q>;
  const int FB = G.fb;
  const uint32_t tile = (uint32_t)FB * FB * FB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / tile;
  if (k >= (uint32_t)nactive) return;
  const uint32_t r = t - k * tile;
  const int slot = active[1 + k];
  const int fx = r % FB, fy = (r / FB) % FB, fz = r / (FB * FB);
  const int RB2 = 2 * G.rb;
  const int lo = kGhost, hi = kGhost + RB2;
  if (fx >= lo && fx < hi && fy >= lo && fy < hi && fz >= lo && fz < hi) return;  // ring only
  const int id = G.slot_to_block[slot];
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  const int NFx = 2 * G.nx, NFy = 2 * G.ny, NFz = 2 * G.nz;
  const int gx = wrapi(bx * RB2 + fx - kGhost, NFx);
  const int gy = wrapi(by * RB2 + fy - kGhost, NFy);
  const int gz = wrapi(bz * RB2 + fz - kGhost, NFz);
  const int vx = gx / 2, vy = gy / 2, vz = gz / 2;
  // a ring cell over a refined coarse cell is served by the neighbour tile
  if (coarse_refined(G, vx, vy, vz)) return;
  const uint32_t N = G.ncells;
  const uint32_t vc = coarse_cell(G, vx, vy, vz);
  const int xm = wrapi(vx - 1, G.nx), xp = wrapi(vx + 1, G.nx);
  const int ym = wrapi(vy - 1, G.ny), yp = wrapi(vy + 1, G.ny);
  const int zm = wrapi(vz - 1, G.nz), zp = wrapi(vz + 1, G.nz);
  const bool okx = !coarse_refined(G, xm, vy, vz) && !coarse_refined(G, xp, vy, vz);
  const bool oky = !coarse_refined(G, vx, ym, vz) && !coarse_refined(G, vx, yp, vz);
  const bool okz = !coarse_refined(G, vx, vy, zm) && !coarse_refined(G, vx, vy, zp);
  const uint32_t cmx = coarse_cell(G, xm, vy, vz), cpx = coarse_cell(G, xp, vy, vz);
  const uint32_t cmy = coarse_cell(G, vx, ym, vz), cpy = coarse_cell(G, vx, yp, vz);
  const uint32_t cmz = coarse_cell(G, vx, vy, zm), cpz = coarse_cell(G, vx, vy, zp);
  // position of the fine cell inside its coarse cell, in coarse units
  const float dx = ((float)(gx - 2 * vx) - 0.5f) * 0.5f;
  const float dy = ((float)(gy - 2 * vy) - 0.5f) * 0.5f;
  const float dz = ((float)(gz - 2 * vz) - 0.5f) * 0.5f;
  const uint32_t cell = pool_cell(G, slot, fx, fy, fz);
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
    const int tx = wrapi(vx + ex, G.nx), ty = wrapi(vy + ey, G.ny), tz = wrapi(vz + ez, G.nz);
    const float* plane = gc + (size_t)i * N;
    float out;
    if (!coarse_refined(G, tx, ty, tz)) {
      // The coarse target is unrefined: only a population that streams from
      // this ring cell into this tile's interior is kept (in-orphan, uniform
      // injection); everything else is zeroed (f = 0, i.e. g = -w).
      const int dxg = wrapi(gx + ex, NFx), dyg = wrapi(gy + ey, NFy), dzg = wrapi(gz + ez, NFz);
      const bool into = coarse_refined(G, dxg / 2, dyg / 2, dzg / 2) && dxg / RB2 == bx && dyg / RB2 == by &&
                        dzg / RB2 == bz;
      out = into ? plane[vc] : -L::w[i];
    } else if (!linear) {
      out = plane[vc];
    } else {
      // linear reconstruction d . G with the gradient projected
      // perpendicular to e_i (common_d3_amr_explode.wgsl:197-212)
      float Gx = okx ? 0.5f * (plane[cpx] - plane[cmx]) : 0.f;
      float Gy = oky ? 0.5f * (plane[cpy] - plane[cmy]) : 0.f;
      float Gz = okz ? 0.5f * (plane[cpz] - plane[cmz]) : 0.f;
      const float ee = (float)(ex * ex + ey * ey + ez * ez);
      const float proj = (Gx * ex + Gy * ey + Gz * ez) / ee;
      Gx -= ex * proj;
      Gy -= ey * proj;
      Gz -= ez * proj;
      out = plane[vc] + (dx * Gx + dy * Gy + dz * Gz);
    }
    gp[(size_t)i * G.pool_plane + cell] = out;
  }
}

// ---------------------------------------------------------------------------
// coalesce: fine ring outflux -> time-t coarse cells under refined blocks,
// plus the orphan correction on unrefined cells next to them
// ---------------------------------------------------------------------------
template <int Q>
__global__ void __launch_bounds__(256)
k_coalesce(const float* __restrict__ gp, float* __restrict__ gc, float* __restrict__ mac,
           const float* __restrict__ mac_pool, PoolGeom G, int orphans) {
  using L = Lattice<Q>;
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= G.ncells) return;
  const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
  const int RB2 = 2 * G.rb;
  const int NFx = 2 * G.nx, NFy = 2 * G.ny, NFz = 2 * G.nz;
  const uint32_t N = G.ncells;
  const int bx = cx / G.rb, by = cy / G.rb, bz = cz / G.rb;
  const int slot = G.block_slot[block_id(G, bx, by, bz)];
  if (slot >= 0) {
    // moments projected from the 8 children (mass-weighted velocity)
    const int f0x = kGhost + 2 * (cx - bx * G.rb), f0y = kGhost + 2 * (cy - by * G.rb),
              f0z = kGhost + 2 * (cz - bz * G.rb);
    float rs = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
    for (int k = 0; k < 8; k++) {
      const uint32_t pc = pool_cell(G, slot, f0x + (k & 1), f0y + ((k >> 1) & 1), f0z + ((k >> 2) & 1));
      const float rho = mac_pool[pc];
      rs += rho;
      mx += rho * mac_pool[(size_t)G.pool_plane + pc];
      my += rho * mac_pool[2 * (size_t)G.pool_plane + pc];
      mz += rho * mac_pool[3 * (size_t)G.pool_plane + pc];
    }
    const float inv = 1.f / fmaxf(rs, 1e-6f);
    reinterpret_cast<float4*>(mac)[c] = make_float4(rs * 0.125f, mx * inv, my * inv, mz * inv);
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const int tx = wrapi(cx + L::ex[i], G.nx), ty = wrapi(cy + L::ey[i], G.ny), tz = wrapi(cz + L::ez[i], G.nz);
      if (coarse_refined(G, tx, ty, tz)) continue;
      float s = 0.f;
      for (int k = 0; k < 8; k++) {
        const int px = 2 * tx + (k & 1), py = 2 * ty + ((k >> 1) & 1), pz = 2 * tz + ((k >> 2) & 1);
        s += gp[(size_t)i * G.pool_plane +
                pool_cell(G, slot, local_in_tile(px, bx, NFx, RB2), local_in_tile(py, by, NFy, RB2),
                          local_in_tile(pz, bz, NFz, RB2))];
      }
      gc[(size_t)i * N + c] = s * 0.125f;  // an average of g is the average of f minus w
    }
    return;
  }
  if (!orphans) return;
  bool near = false;
  for (int i = 1; i < Q; i++)
    near = near || coarse_refined(G, wrapi(cx + L::ex[i], G.nx), wrapi(cy + L::ey[i], G.ny), wrapi(cz + L::ez[i], G.nz));
  if (!near) return;
#pragma unroll
  for (int i = 1; i < Q; i++) {
    const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
    const int tx = wrapi(cx + ex, G.nx), ty = wrapi(cy + ey, G.ny), tz = wrapi(cz + ez, G.nz);
    if (coarse_refined(G, tx, ty, tz)) continue;
    float s = 0.f;   // sum of f over the children found in refined tiles
    float nIn = 0.f;
    for (int k = 0; k < 8; k++) {
      const int ox = k & 1, oy = (k >> 1) & 1, oz = (k >> 2) & 1;
      const int px = wrapi(2 * tx + ox, NFx), py = wrapi(2 * ty + oy, NFy), pz = wrapi(2 * tz + oz, NFz);
      const int qx = wrapi(px - ex, NFx), qy = wrapi(py - ey, NFy), qz = wrapi(pz - ez, NFz);
      const int qbx = qx / RB2, qby = qy / RB2, qbz = qz / RB2;
      const int qs = G.block_slot[block_id(G, qbx, qby, qbz)];
      if (qs >= 0) {
        s += gp[(size_t)i * G.pool_plane +
                pool_cell(G, qs, local_in_tile(px, qbx, NFx, RB2), local_in_tile(py, qby, NFy, RB2),
                          local_in_tile(pz, qbz, NFz, RB2))] + L::w[i];
      }
      const int hx = wrapi(2 * cx + ox + ex, NFx), hy = wrapi(2 * cy + oy + ey, NFy), hz = wrapi(2 * cz + oz + ez, NFz);
      if (coarse_refined(G, hx / 2, hy / 2, hz / 2)) nIn += 1.f;
    }
    const float fc = gc[(size_t)i * N + c] + L::w[i];
    const float net = s - nIn * fc;
    if (net != 0.f) gc[(size_t)i * N + c] = (fc + net * 0.125f) - L::w[i];
  }
}

// ---------------------------------------------------------------------------
// interp fill of newly refined tiles, and average/drain of dying ones
// ---------------------------------------------------------------------------
template <int Q>
__device__ __forceinline__ void coarse_sample(const float* __restrict__ gc, uint32_t N, uint32_t cell, float& rho,
                                              float& ux, float& uy, float& uz, float (&fneq)[Q]) {
  float g[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) g[i] = gc[(size_t)i * N + cell];
  const float F0[3] = {0.f, 0.f, 0.f};
  float drho;
  moments<Q>(g, F0, drho, rho, ux, uy, uz);
#pragma unroll
  for (int i = 0; i < Q; i++) fneq[i] = g[i] - geq<Q>(i, rho, ux, uy, uz);
}

template <int Q>
__global__ void __launch_bounds__(256)
k_interp(const float* __restrict__ gc, float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ pairs,
         int npairs, float rescale) {
  const int FB = G.fb;
  const uint32_t tile = (uint32_t)FB * FB * FB;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / tile;
  if (k >= (uint32_t)npairs) return;
  const uint32_t r = t - k * tile;
  const int id = pairs[1 + 2 * k], slot = pairs[2 + 2 * k];
  const int fx = r % FB, fy = (r / FB) % FB, fz = r / (FB * FB);
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  // coarse-unit position origin - 1/4 + (j - GHOST)/2 (fineToCoarseUnit3)
  const float px = (float)(bx * G.rb) - 0.25f + 0.5f * (float)(fx - kGhost);
  const float py = (float)(by * G.rb) - 0.25f + 0.5f * (float)(fy - kGhost);
  const float pz = (float)(bz * G.rb) - 0.25f + 0.5f * (float)(fz - kGhost);
  const int c0x = (int)floorf(px), c0y = (int)floorf(py), c0z = (int)floorf(pz);
  const float tx = px - c0x, ty = py - c0y, tz = pz - c0z;
  float rho = 0.f, ux = 0.f, uy = 0.f, uz = 0.f, fneq[Q];
#pragma unroll
  for (int i = 0; i < Q; i++) fneq[i] = 0.f;
  for (int s = 0; s < 8; s++) {
    const int i = s & 1, j = (s >> 1) & 1, kk = (s >> 2) & 1;
    const float w = (i ? tx : 1.f - tx) * (j ? ty : 1.f - ty) * (kk ? tz : 1.f - tz);
    const uint32_t cell = coarse_cell(G, wrapi(c0x + i, G.nx), wrapi(c0y + j, G.ny), wrapi(c0z + kk, G.nz));
    float sr, sx, sy, sz, sf[Q];
    coarse_sample<Q>(gc, G.ncells, cell, sr, sx, sy, sz, sf);
    rho += w * sr;
    ux += w * sx;
    uy += w * sy;
    uz += w * sz;
#pragma unroll
    for (int q = 0; q < Q; q++) fneq[q] += w * sf[q];
  }
  const uint32_t cell = pool_cell(G, slot, fx, fy, fz);
#pragma unroll
  for (int q = 0; q < Q; q++) gp[(size_t)q * G.pool_plane + cell] = geq<Q>(q, rho, ux, uy, uz) + rescale * fneq[q];
}

template <int Q>
__global__ void __launch_bounds__(256)
k_average(const float* __restrict__ gp, float* __restrict__ gc, float* __restrict__ mac, PoolGeom G,
          const int32_t* __restrict__ pairs, int npairs, float rescale) {
  const uint32_t per = (uint32_t)G.rb * G.rb * G.rb;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t k = t / per;
  if (k >= (uint32_t)npairs) return;
  const uint32_t r = t - k * per;
  const int id = pairs[1 + 2 * k], slot = pairs[2 + 2 * k];
  const int lx = r % G.rb, ly = (r / G.rb) % G.rb, lz = r / (G.rb * G.rb);
  const int bx = id % G.nbx, by = (id / G.nbx) % G.nby, bz = id / (G.nbx * G.nby);
  float rs = 0.f, mx = 0.f, my = 0.f, mz = 0.f, sneq[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) sneq[q] = 0.f;
  for (int c = 0; c < 8; c++) {
    const uint32_t pc = pool_cell(G, slot, kGhost + 2 * lx + (c & 1), kGhost + 2 * ly + ((c >> 1) & 1),
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
  const float rho = rs * 0.125f;
  const float inv = 1.f / fmaxf(rs, 1e-6f);
  const float ux = mx * inv, uy = my * inv, uz = mz * inv;
  const uint32_t cc = coarse_cell(G, bx * G.rb + lx, by * G.rb + ly, bz * G.rb + lz);
#pragma unroll
  for (int q = 0; q < Q; q++) gc[(size_t)q * G.ncells + cc] = geq<Q>(q, rho, ux, uy, uz) + rescale * (sneq[q] * 0.125f);
  reinterpret_cast<float4*>(mac)[cc] = make_float4(rho, ux, uy, uz);
}

// ---------------------------------------------------------------------------
// criterion: max |curl u| over each block's L0 box dilated by `lead`, from
// the macroscopic field, central differences, one rounding per operation so
// the host twin reproduces it bit for bit
// ---------------------------------------------------------------------------
__device__ __forceinline__ float3 mac_u(const float4* mac, const PoolGeom& G, int x, int y, int z) {
  const float4 m = mac[coarse_cell(G, wrapi(x, G.nx), wrapi(y, G.ny), wrapi(z, G.nz))];
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

__global__ void __launch_bounds__(64) k_criterion(const float4* __restrict__ mac, float* __restrict__ crit, PoolGeom G,
                                                  int lead) {
  const int b = blockIdx.x;
  const int bx = b % G.nbx, by = (b / G.nbx) % G.nby, bz = b / (G.nbx * G.nby);
  const int x0 = bx * G.rb - lead, y0 = by * G.rb - lead, z0 = bz * G.rb - lead;
  const int n = G.rb + 2 * lead;
  const int total = n * n * n;
  float best = -1e30f;
  for (int k = threadIdx.x; k < total; k += 64) {
    const int kx = k % n, ky = (k / n) % n, kz = k / (n * n);
    best = fmaxf(best, vort_mag(mac, G, x0 + kx, y0 + ky, z0 + kz));
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
  //    is judged against the lower coarsen threshold)
  for (int b = tid; b < a.nblocks; b += 1024) {
    const bool held = a.block_slot[b] >= 0;
    const float c = a.crit[b];
    a.want[b] = held ? (c >= a.keep_at ? 1 : 0) : (c >= a.refine_at ? 1 : 0);
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
template <int Q>
__global__ void __launch_bounds__(256)
k_pool_totals(const float* __restrict__ gc, const float* __restrict__ gp, PoolGeom G, const int32_t* __restrict__ active,
              double* partials, unsigned* counter, double* out) {
  using L = Lattice<Q>;
  __shared__ double lds[16 * 4];
  double v[4] = {0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < G.ncells; c += stride) {
    const int cx = c % G.nx, cy = (c / G.nx) % G.ny, cz = c / (G.nx * G.ny);
    if (coarse_refined(G, cx, cy, cz)) continue;
    for (int i = 0; i < Q; i++) {
      const double f = (double)gc[(size_t)i * G.ncells + c] + (double)L::w[i];
      v[0] += f; v[1] += L::ex[i] * f; v[2] += L::ey[i] * f; v[3] += L::ez[i] * f;
    }
  }
  const int nact = active[0];
  const uint32_t rb2 = 2 * G.rb;
  const uint32_t per = rb2 * rb2 * rb2;
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < (uint32_t)nact * per; t += stride) {
    const uint32_t k = t / per, r = t - k * per;
    const int slot = active[1 + k];
    const uint32_t pc = pool_cell(G, slot, kGhost + r % rb2, kGhost + (r / rb2) % rb2, kGhost + r / (rb2 * rb2));
    for (int i = 0; i < Q; i++) {
      const double f = 0.125 * ((double)gp[(size_t)i * G.pool_plane + pc] + (double)L::w[i]);
      v[0] += f; v[1] += L::ex[i] * f; v[2] += L::ey[i] * f; v[3] += L::ez[i] * f;
    }
  }
  block_sum<4>(v, lds);
  partials_then_fixed_order_sum<4>(v, partials, counter, [&](int j, double s) { out[j] = s; });
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
static inline dim3 grid_for(uint64_t n) { return dim3((unsigned)((n + 255) / 256)); }

hipError_t launch_pool_step(int q, const float* gin, float* gout, float* mac_pool, const PoolGeom& g,
                            const CollideParams& fine, const int32_t* active, int nactive, hipStream_t s) {
  if (nactive <= 0) return hipSuccess;
  const dim3 grid = grid_for((uint64_t)nactive * g.fb * g.fb * g.fb);
  if (q == 19) hipLaunchKernelGGL(k_pool_step<19>, grid, dim3(256), 0, s, gin, gout, mac_pool, g, fine, active, nactive);
  else hipLaunchKernelGGL(k_pool_step<27>, grid, dim3(256), 0, s, gin, gout, mac_pool, g, fine, active, nactive);
  return hipGetLastError();
}

hipError_t launch_explode(int q, const float* gc, float* gp, const PoolGeom& g, const int32_t* active, int nactive,
                          int linear, hipStream_t s) {
  if (nactive <= 0) return hipSuccess;
  const dim3 grid = grid_for((uint64_t)nactive * g.fb * g.fb * g.fb);
  if (q == 19) hipLaunchKernelGGL(k_explode<19>, grid, dim3(256), 0, s, gc, gp, g, active, nactive, linear);
  else hipLaunchKernelGGL(k_explode<27>, grid, dim3(256), 0, s, gc, gp, g, active, nactive, linear);
  return hipGetLastError();
}

hipError_t launch_coalesce(int q, const float* gp, float* gc, float* mac, const float* mac_pool, const PoolGeom& g,
                           int orphans, hipStream_t s) {
  const dim3 grid = grid_for(g.ncells);
  if (q == 19) hipLaunchKernelGGL(k_coalesce<19>, grid, dim3(256), 0, s, gp, gc, mac, mac_pool, g, orphans);
  else hipLaunchKernelGGL(k_coalesce<27>, grid, dim3(256), 0, s, gp, gc, mac, mac_pool, g, orphans);
  return hipGetLastError();
}

hipError_t launch_interp(int q, const float* gc, float* gp, const PoolGeom& g, const int32_t* pairs, int npairs,
                         float rescale, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  const dim3 grid = grid_for((uint64_t)npairs * g.fb * g.fb * g.fb);
  if (q == 19) hipLaunchKernelGGL(k_interp<19>, grid, dim3(256), 0, s, gc, gp, g, pairs, npairs, rescale);
  else hipLaunchKernelGGL(k_interp<27>, grid, dim3(256), 0, s, gc, gp, g, pairs, npairs, rescale);
  return hipGetLastError();
}

hipError_t launch_average(int q, const float* gp, float* gc, float* mac, const PoolGeom& g, const int32_t* pairs,
                          int npairs, float rescale, hipStream_t s) {
  if (npairs <= 0) return hipSuccess;
  const dim3 grid = grid_for((uint64_t)npairs * g.rb * g.rb * g.rb);
  if (q == 19) hipLaunchKernelGGL(k_average<19>, grid, dim3(256), 0, s, gp, gc, mac, g, pairs, npairs, rescale);
  else hipLaunchKernelGGL(k_average<27>, grid, dim3(256), 0, s, gp, gc, mac, g, pairs, npairs, rescale);
  return hipGetLastError();
}

hipError_t launch_criterion(const float* mac, float* crit, const PoolGeom& g, float lead, hipStream_t s) {
  hipLaunchKernelGGL(k_criterion, dim3(g.nbx * g.nby * g.nbz), dim3(64), 0, s, (const float4*)mac, crit, g,
                     (int)lead);
  return hipGetLastError();
}

hipError_t launch_manage(const ManageArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_manage, dim3(1), dim3(1024), 0, s, a);
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
  const dim3 grid(256);
  if (q == 19) hipLaunchKernelGGL(k_pool_totals<19>, grid, dim3(256), 0, s, gc, gp, g, active, partials, counter, out);
  else hipLaunchKernelGGL(k_pool_totals<27>, grid, dim3(256), 0, s, gc, gp, g, active, partials, counter, out);
  return hipGetLastError();
}

}  // namespace lbm

// Dense-lattice kernels for gfx950: fused pull-stream + collide (D2Q9, D3Q19,
// D3Q27; BGK or TRT; Guo forcing; halfway bounce-back on walls and a pinned
// body; ALBC sponge; SOLID_EQ) with the FP64 momentum-exchange force fused
// in, equilibrium initialisation and FP64 population totals.
//
// Semantics follow the reference kernels
//   shaders/common_d3_step.wgsl:206-222 (initEq), :224-400 (step)
//   shaders/lbm_step.wgsl:137-270 (D2Q9 main)
//   plans/TRT.md:333-349 (TRT delta form)
//   shaders/common_d3_force.wgsl:171-216, shaders/lbm_force.wgsl:128-147 (MEM)
// The arithmetic is rewritten in shifted variables g_i = f_i - w_i:
//   rho - 1 = sum g_i (+ (sum w_i - 1), nonzero only for D2Q9's weights)
//   rho u   = sum g_i e_i + F/2          (sum w_i e_i = 0 exactly)
//   g_eq    = w_i ((rho - 1) + rho (3 e.u + 4.5 (e.u)^2 - 1.5 u.u))
// which is the same collision with the large constant w_i taken out before
// rounding.
//
// Layout: blocks of kLatBlock (64) consecutive cells, each holding its Q
// populations as Q runs of 64 elements (dense.h kLatBlock, lat_elem).
// One thread owns one cell; a wave covers 64 consecutive x, so every plane
// load and store of a wave is one contiguous 256 B (f32) or 128 B (16-bit)
// span, shifted by at most one element for the x-moving directions.  All Q
// loads are issued back to back (no branch between them); the rare cells next
// to a wall or the body then patch their reflected directions.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <type_traits>
#include "collide.h"
#include "dense.h"
#include "device_common.h"
#include "lattice_pairs.h"
#include "lattice_tables.h"
#include "rigid_device.h"

namespace lbm {

// One wave (64 lanes, all active): lane l sums slots l, l+64, ... of `p` in
// order, a fixed butterfly finishes, lane 0 appends the record.  Used by
// k_force_finalize and, for the previous step, inside k_step.
__device__ __forceinline__ void sum_force_slots(const double* __restrict__ p, uint32_t nslots, double* history,
                                                unsigned long long* written, uint32_t cap, int lane) {
  double v[6] = {0, 0, 0, 0, 0, 0};
  for (uint32_t s = lane; s < nslots; s += 64) {
#pragma unroll
    for (int k = 0; k < 6; k++) v[k] += p[(size_t)s * 6 + k];
  }
#pragma unroll
  for (int k = 0; k < 6; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
    const unsigned long long n = *written;
#pragma unroll
    for (int k = 0; k < 6; k++) history[(size_t)(n % cap) * 6 + k] = v[k];
    *written = n + 1;
  }
}

// First element of the block holding cell c (its population i is then at
// element i * kLatBlock + c % kLatBlock of the block).
template <int Q, class T>
__device__ __forceinline__ T* block_of(T* g, uint32_t c) {
  return g + (size_t)(c / kLatBlock) * (Q * kLatBlock);
}

// Lattice offset of cell x of a row, relative to the row's first element
// (rows are whole layout blocks).
template <int Q>
__device__ __forceinline__ uint32_t row_off(int x) {
  // a 24-bit multiply is full rate (a 32-bit one a quarter); x < 2^24
  return (uint32_t)__umul24((uint32_t)x / kLatBlock, Q * kLatBlock) + (uint32_t)x % kLatBlock;
}

// Signed distance of the pinned body at cell (x, y, z), binary32 with one
// rounding per operation (no contraction), the same value k_classify and the
// oracle compute:
//   sphere  phi = sqrt((x-cx)^2 + (y-cy)^2 + (z-cz)^2) - a
//           (common_d3_geometry.wgsl:119, :174-187)
//   circle  nearest periodic image of (x-cx, y-cy), then length - a
//           (common_geometry.wgsl:94-108)
__device__ __forceinline__ float body_phi(int shape, float radius, float cx, float cy, float cz, int nx, int ny,
                                          int x, int y, int z) {
  float dx = sub_rn((float)x, cx);
  float dy = sub_rn((float)y, cy);
  if (shape == 2) {
    const float W = (float)nx, H = (float)ny;
    dx = sub_rn(dx, mul_rn(W, rintf(div_rn(dx, W))));
    dy = sub_rn(dy, mul_rn(H, rintf(div_rn(dy, H))));
    return sub_rn(sqrt_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy))), radius);
  }
  const float dz = sub_rn((float)z, cz);
  const float s = add_rn(add_rn(mul_rn(dx, dx), mul_rn(dy, dy)), mul_rn(dz, dz));
  return sub_rn(sqrt_rn(s), radius);
}

// One step of tile b of a solve: every thread of the block pulls, collides
// and stores its cell; a tile with a force slot leaves its FP64 force there.
// `salt` seeds the dither of 16-bit stores; `mac` (or null) receives the
// macroscopic field.  Shared by the single-solve and the batched kernels.
// ROW (step_row_mode): each wave lies in one row, so y, z and every source
// row's base address are wave-uniform and only the x part of an address is
// per lane; otherwise each lane locates its sources' blocks itself.
// RIGID (with GEN): a moving rigid body, a.rigid (its own instantiations, so
// that the kernels without one keep their register budget).
// DIFF (with GEN): the pinned body's diffuse penalty coupling in place of
// bounce-back (common_d3_step.wgsl:288-300; the reference's default without
// ?bounceback): every cell takes the Guo force F + rho chi (0 - u*), u* = m /
// rho, chi = 1/2 (1 - tanh(clamp(phi / eps, -20, 20))), and a cell with chi >=
// 1e-6 adds rho chi u* to the body's force (common_d3_force.wgsl:242-268).
template <int Q, class ST, bool GEN, bool NT, bool ROW, bool RIGID = false, bool DIFF = false, bool FRC = true>
__device__ __forceinline__ void step_tile(const typename ST::T* __restrict__ gin, typename ST::T* __restrict__ gout,
                                          const StepArgs& a, uint32_t b, uint32_t salt, float* mac) {
  static_assert(!DIFF || FRC, "the diffuse coupling forces every cell");
  using L = Lattice<Q>;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  int x, y, z;
  bool active;
  if (a.linear) {
    const uint32_t c = b * (blockDim.x * blockDim.y) + threadIdx.y * blockDim.x + threadIdx.x;
    const uint32_t row = div_invariant(c, a.div_nx);
    x = (int)(c - row * nx);
    const uint32_t r = ROW ? __builtin_amdgcn_readfirstlane(row) : row;
    z = (int)div_invariant(r, a.div_ny);
    y = (int)(r - (uint32_t)z * ny);
    active = c < a.ncells;
  } else {
    const int tx = b % a.ntx;
    const int rest = b / a.ntx;
    const int ty = rest % a.nty;
    z = rest / a.nty;
    x = tx * blockDim.x + threadIdx.x;
    y = ty * blockDim.y + threadIdx.y;
    if constexpr (ROW) y = __builtin_amdgcn_readfirstlane(y);
    active = x < nx && y < ny;
  }
  if constexpr (GEN) {
    if (active && a.cover)
      active = a.cover[((z / a.cover_rb) * a.cover_nby + y / a.cover_rb) * a.cover_nbx + x / a.cover_rb] < 0;
  }
  double fv[6] = {0, 0, 0, 0, 0, 0};  // this cell's force/torque on the body

  if (active) {
    const uint32_t cell = (uint32_t)((z * ny + y) * nx + x);

    // This cell's populations: population i at self + i * kLatBlock, lane
    // offset self_lo.
    using T = typename ST::T;
    size_t self;
    uint32_t self_lo;
    if constexpr (ROW) {
      self = (size_t)(z * ny + y) * nx * Q;
      self_lo = row_off<Q>(x);
    } else {
      self = (size_t)(cell / kLatBlock) * (Q * kLatBlock);
      self_lo = cell % kLatBlock;
    }

    // 1. periodic pull-stream, all Q loads straight-line.  The source of
    // direction i is cell - e_i, wrapping only on the domain faces.
    float g[Q];
    if constexpr (ROW) {
      const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
      const uint32_t lo_m = row_off<Q>(xm), lo_p = row_off<Q>(xp);
      const int ym = y == 0 ? ny - 1 : y - 1, yp = y == ny - 1 ? 0 : y + 1;
      const int zm = z == 0 ? nz - 1 : z - 1, zp = z == nz - 1 ? 0 : z + 1;
#pragma unroll
      for (int i = 0; i < Q; i++) {
        const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
        const int sy = ey == 1 ? ym : (ey == -1 ? yp : y);
        const int sz = ez == 1 ? zm : (ez == -1 ? zp : z);
        const T* src = gin + (size_t)(sz * ny + sy) * nx * Q + i * kLatBlock;
        g[i] = ld<ST, NT>(src, ex == 1 ? lo_m : (ex == -1 ? lo_p : self_lo));
      }
    } else {
      const int nxy = nx * ny;
      const int ox_m = x == 0 ? nx - 1 : -1, ox_p = x == nx - 1 ? 1 - nx : 1;
      const int oy_m = y == 0 ? (ny - 1) * nx : -nx, oy_p = y == ny - 1 ? (1 - ny) * nx : nx;
      const int oz_m = z == 0 ? (nz - 1) * nxy : -nxy, oz_p = z == nz - 1 ? (1 - nz) * nxy : nxy;
#pragma unroll
      for (int i = 0; i < Q; i++) {
        const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
        const int o = (ex == 1 ? ox_m : (ex == -1 ? ox_p : 0)) + (ey == 1 ? oy_m : (ey == -1 ? oy_p : 0)) +
                      (ez == 1 ? oz_m : (ez == -1 ? oz_p : 0));
        const uint32_t sc = (uint32_t)((int)cell + o);
        const T* src = gin + (size_t)(sc / kLatBlock) * (Q * kLatBlock) + i * kLatBlock;
        g[i] = ld<ST, NT>(src, sc % kLatBlock);
      }
    }
    const T* in_self = gin + self;
    T* out_self = gout + self;

    uint8_t c = 0;
    if constexpr (GEN) {
      // under the diffuse coupling the body has no links and no solid cells
      if (a.cls && !DIFF) c = a.cls[cell];
      const bool onFace = (a.wall_mask & 1 && (x == 0 || x == nx - 1)) ||
                          (a.wall_mask & 2 && (y == 0 || y == ny - 1)) ||
                          (a.wall_mask & 4 && (z == 0 || z == nz - 1));
      if ((c & 2) || onFace) {
        // 1b. halfway bounce-back: the body first, then the walls
        // (common_d3_step.wgsl:250-276).  w_i = w_opp(i), so a reflected
        // population keeps its shift.  Which sources are solid comes from the
        // cell's precomputed link mask (one load instead of a class byte per
        // neighbour, which would keep Q addresses live in this rare branch and
        // cost the whole kernel occupancy).
        double fx = 0, fy = 0, fz = 0;
        const uint32_t lm = (c & 2) ? a.links[cell] : 0u;
        // a moving body's velocity here and the wall density: this cell's
        // own pre-stream density (common_d3_step.wgsl:236-242)
        float wux = 0.f, wuy = 0.f, wuz = 0.f, rho_w = 1.f;
        if (RIGID && lm) {
          body_velocity_f(a.rigid, x, y, z, wux, wuy, wuz);
          if (a.rigid->rho_w_local) {
            rho_w = 0.f;
            for (int i = 0; i < Q; i++) rho_w += ST::load(in_self + i * kLatBlock, self_lo) + L::w[i];
          }
        }
#pragma unroll
        for (int i = 1; i < Q; i++) {
          const int ex = L::ex[i], ey = L::ey[i], ez = L::ez[i];
          const T* oplane = in_self + L::opp[i] * kLatBlock;
          if ((lm >> i) & 1u) {
            g[i] = ST::load(oplane, self_lo);
            // Ladd's moving-wall term 2 w rho_w (e . u_body) / cs2
            if constexpr (RIGID) g[i] += 2.f * L::w[i] * rho_w * ((float)ex * wux + (float)ey * wuy + (float)ez * wuz) * 3.f;
            // momentum exchange on a fluid cell: -e_i (2 f_opp), pinned body
            if (!(c & 1)) {
              const double m = 2.0 * ((double)g[i] + (double)L::w[i]);
              fx -= ex * m;
              fy -= ey * m;
              fz -= ez * m;
            }
          } else if (onFace) {
            const bool xo = (ex == 1 && x == 0) || (ex == -1 && x == nx - 1);
            const bool yo = (ey == 1 && y == 0) || (ey == -1 && y == ny - 1);
            const bool zo = (ez == 1 && z == 0) || (ez == -1 && z == nz - 1);
            const bool yWall = (a.wall_mask & 2) && yo;
            if (((a.wall_mask & 1) && xo) || yWall || ((a.wall_mask & 4) && zo)) {
              float corr = 0.f;
              if (ex != 0 && yWall) {
                // moving y-wall (common_walls.wgsl:27-31): 2 w e_x u_wall / cs2
                const float uw = (ey == 1) ? a.wall_u0 : a.wall_u1;
                corr = 2.f * L::w[i] * (float)ex * uw / (1.0f / 3.0f);
              }
              g[i] = ST::load(oplane, self_lo) + corr;
            }
          }
        }
        if (a.force_slot && (fx != 0.0 || fy != 0.0 || fz != 0.0)) {
          double rx = x - a.body_c[0], ry = y - a.body_c[1], rz = z - a.body_c[2];
          if (a.periodic_arm) {
            rx -= nx * rint(rx / nx);
            ry -= ny * rint(ry / ny);
            rz = 0.0;
          }
          fv[0] = fx; fv[1] = fy; fv[2] = fz;
          fv[3] = ry * fz - rz * fy;
          fv[4] = rz * fx - rx * fz;
          fv[5] = rx * fy - ry * fx;
        }
      }
    }

    // 2. pair sums and differences; moments from them.  The momentum sums
    // start from -0, the identity of floating-point addition, so that their
    // first term costs no instruction (0 + x is not x when x is -0).
    constexpr int NP = Pairs<Q>::NP;
    float S[NP], D[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int i = Pairs<Q>::lo(k);
      S[k] = g[i] + g[L::opp[i]];
      D[k] = g[i] - g[L::opp[i]];
    }
    // D2Q9's f32 weights do not sum to 1 (wsum_m1); the 3D sets' do exactly
    float drho = Q == 9 ? g[0] + a.wsum_m1 : g[0], mx = -0.f, my = -0.f, mz = -0.f;
#pragma unroll
    for (int k = 0; k < NP; k++) {
      const int i = Pairs<Q>::lo(k);
      drho += S[k];
      if (L::ex[i] == 1) mx += D[k]; else if (L::ex[i] == -1) mx -= D[k];
      if (L::ey[i] == 1) my += D[k]; else if (L::ey[i] == -1) my -= D[k];
      if (L::ez[i] == 1) mz += D[k]; else if (L::ez[i] == -1) mz -= D[k];
    }
    const float rho = 1.f + drho;
    const float rinv = recip_nr(fmaxf(rho, 1e-6f));
    // FRC false: no body force (F = 0), so every Guo term vanishes
    float Fx = FRC ? a.F[0] : 0.f, Fy = FRC ? a.F[1] : 0.f, Fz = FRC ? a.F[2] : 0.f;
    if constexpr (DIFF) {
      const float phi = body_phi(a.body_shape, a.body_r, a.body_cf[0], a.body_cf[1], a.body_cf[2], nx, ny, x, y, z);
      const float chi = 0.5f * (1.f - tanhf(fminf(fmaxf(phi / a.chi_eps, -20.f), 20.f)));
      const float usx = mx * rinv, usy = my * rinv, usz = mz * rinv;
      const float rc = rho * chi;
      Fx += rc * (0.f - usx);
      Fy += rc * (0.f - usy);
      Fz += rc * (0.f - usz);
      if (a.force_slot && chi >= 1e-6f) {
        // the body feels the penalty's negative, rho chi (u* - U_s)
        const double fx = (double)rc * usx, fy = (double)rc * usy, fz = (double)rc * usz;
        double rx = x - a.body_c[0], ry = y - a.body_c[1], rz = z - a.body_c[2];
        if (a.periodic_arm) {
          rx -= nx * rint(rx / nx);
          ry -= ny * rint(ry / ny);
          rz = 0.0;
        }
        fv[0] = fx; fv[1] = fy; fv[2] = fz;
        fv[3] = ry * fz - rz * fy;
        fv[4] = rz * fx - rx * fz;
        fv[5] = rx * fy - ry * fx;
      }
    }
    // Guo: u = u* + F / (2 rho)
    const float ux = FRC ? (mx + 0.5f * Fx) * rinv : mx * rinv;
    const float uy = FRC ? (my + 0.5f * Fy) * rinv : my * rinv;
    const float uz = FRC ? (mz + 0.5f * Fz) * rinv : mz * rinv;
    const float usq = ux * ux + uy * uy + uz * uz;
    const float uF = FRC ? ux * Fx + uy * Fy + uz * Fz : 0.f;

    bool solid = false;
    if constexpr (GEN) solid = a.solid_eq && (c & 1);
    // a moving body's interior velocity (zero for a pinned body)
    float sux = 0.f, suy = 0.f, suz = 0.f;
    if constexpr (RIGID) {
      if (solid) body_velocity_f(a.rigid, x, y, z, sux, suy, suz);
    }

    if (mac) {
      float4 m = make_float4(rho, ux, uy, uz);
      if (solid && a.solid_mac_rest) m = make_float4(1.f, sux, suy, suz);
      reinterpret_cast<float4*>(mac)[cell] = m;
    }

    if (solid) {
      // SOLID_EQ: feq(1, u_body), g = feq - w; for a pinned body g = 0,
      // exact in every format
      const float uu = sux * sux + suy * suy + suz * suz;
#pragma unroll
      for (int i = 0; i < Q; i++) {
        const float eu = (float)L::ex[i] * sux + (float)L::ey[i] * suy + (float)L::ez[i] * suz;
        st<ST, NT>(out_self + i * kLatBlock, self_lo, L::w[i] * (3.f * eu + 4.5f * eu * eu - 1.5f * uu));
      }
    } else {
      float sp = 0.f;
      if constexpr (GEN) {
        if (a.sponge_w > 0.f) {
          const float big = 1e30f;
          // under a moving window, distances in window coordinates
          // (common_d3_step.wgsl:332-337)
          float px = (float)x, py = (float)y, pz = (float)z;
          if constexpr (RIGID) {
            const RigidBody* B = a.rigid;
            px = (float)win_coord(x, B->c[0], B->win_a[0], B->win_n[0]);
            py = (float)win_coord(y, B->c[1], B->win_a[1], B->win_n[1]);
            pz = (float)win_coord(z, B->c[2], B->win_a[2], B->win_n[2]);
          }
          const float dX = (a.sponge_mask & 1) ? fminf(px, (float)(nx - 1) - px) : big;
          const float dY = (a.sponge_mask & 2) ? fminf(py, (float)(ny - 1) - py) : big;
          const float dZ = (a.sponge_mask & 4) ? fminf(pz, (float)(nz - 1) - pz) : big;
          const float d = fminf(dX, fminf(dY, dZ));
          // beyond the band d / w >= 1 rounds to >= 1, so t clamps to 0 there:
          // only cells inside the band pay for the division
          if (d < a.sponge_w) {
            float t = 1.f - d / a.sponge_w;
            t = fminf(fmaxf(t, 0.f), 1.f);
            sp = t * t * (3.f - 2.f * t);
          }
        }
      }

      // 3. collide pair by pair.  With g+ = (g_i + g_j)/2, g- = (g_i - g_j)/2
      // (j = opp i) and the equilibrium split into its even part
      //   A = w (rho - 1 + rho (4.5 (e.u)^2 - 1.5 u.u))
      // and odd part B = 3 w rho (e.u), BGK/TRT with Guo forcing reads
      //   P = g+ - omega   (g+ - A) + (1 - omega/2)   w (9 (e.u)(e.F) - 3 u.F)
      //   M = g- - omega_m (g- - B) + (1 - omega_m/2) 3 w (e.F)
      //   g_i' = P + M,  g_j' = P - M,
      // which is the BGK line plus the TRT delta terms of plans/TRT.md:333-349
      // regrouped (omega_m = omega gives BGK).  The sponge blends P and M
      // toward the even and odd parts of feq(1, U_far) - w.
      //
      // The kernel evaluates it doubled: P2 = 2P and M2 = 2M from S = 2 g+
      // and D = 2 g- with 2w in place of w (and the host's pair constants
      // doubled), then g_i' = fma(1/2, P2, M2/2), g_j' = fma(1/2, P2, -M2/2).
      // Scaling by two commutes with rounding, so every intermediate is
      // exactly twice its undoubled value and the outputs are bit for bit
      // P + M and P - M, without the two halvings per pair.
      //
      // Rounding: every relaxation keeps the form h - omega (h - eq), and the
      // exact f32 weight w is the LAST factor of A and B.  A pre-rounded
      // constant such as f32(3 w) or f32((1 - omega_m)/2) would carry the
      // same relative bias (~1e-8) into every cell and every step, which
      // breaks mass/momentum conservation systematically: the error then
      // grows linearly in time instead of as a random walk.
      const float om = a.omega, omm = a.omega_m;
      const float c0 = drho - 1.5f * rho * usq;
      // random bits of this cell and step for the 16-bit stores
      const uint32_t hsh = ST::kBytes == 2 ? cell_hash(cell, salt) : 0u;
      const float g3uF2 = FRC ? 6.f * a.guo * uF : 0.f;  // 2 (3 (1 - omega/2) u.F)
      // The sponge blend only where a lane of the wave lies in the band.
      auto collide_pairs = [&](auto SPONGE) {
        constexpr bool SP = decltype(SPONGE)::value;
        {
          const float w = L::w[0];
          float p = g[0] - om * (g[0] - w * c0);
          if constexpr (FRC) p -= w * (0.5f * g3uF2);
          if constexpr (SP) p += sp * (a.ae[13] - p);
          st_sr<ST, NT>(out_self, self_lo, p, ST::dither(hsh, 0));
        }
#pragma unroll
        for (int k = 0; k < NP; k++) {
          const int i = Pairs<Q>::lo(k);
          const int j = L::opp[i];
          const float w = L::w[i], w2 = 2.f * L::w[i];
          const float eu = edot<Q>(i, ux, uy, uz);
          const float t = rho * eu;
          const float X = c0 + 4.5f * (t * eu);
          float P2 = S[k] - om * (S[k] - w2 * X);
          float M2 = D[k] - omm * (D[k] - w2 * (3.f * t));
          if constexpr (FRC) {
            // Guo source, even and odd parts (doubled): (1 - omega/2) w (9
            // (e.u)(e.F) - 3 u.F) and (1 - omega_m/2) 3 w (e.F); a uniform F
            // has them per pair
            float c1k = a.c1[k], sok = a.so[k];
            if constexpr (DIFF) {
              const float eF = edot<Q>(i, Fx, Fy, Fz);
              c1k = 18.f * a.guo * w * eF;
              sok = (1.f - 0.5f * omm) * 6.f * w * eF;
            }
            P2 += c1k * eu - w * g3uF2;
            M2 += sok;
          }
          if constexpr (SP) {
            P2 += sp * (a.ae[k] - P2);
            M2 += sp * (a.bo[k] - M2);
          }
          const float hM = 0.5f * M2;
          st_sr_pair<ST, NT>(out_self + i * kLatBlock, out_self + j * kLatBlock, self_lo,
                             __builtin_fmaf(0.5f, P2, hM), __builtin_fmaf(0.5f, P2, -hM),
                             ST::dither(hsh, (11 * (2 * k + 1)) % 17), ST::dither(hsh, (11 * (2 * k + 2)) % 17));
        }
      };
      collide_pairs(std::integral_constant<bool, GEN>{});
    }
  }

  if constexpr (GEN) {
    // Tiles touching the body reduce their cells' force in FP64 and park it
    // in the tile's slot; k_force_finalize sums the slots in order.
    if (a.force_slot) {
      const int slot = a.force_slot[b];
      if (slot >= 0) {
        __shared__ double lds[16 * 6];
        block_sum<6>(fv, lds);
        if (threadIdx.x == 0 && threadIdx.y == 0) {
#pragma unroll
          for (int k = 0; k < 6; k++) a.force_partials[(size_t)slot * 6 + k] = fv[k];
        }
      }
    }
  }
}

template <int Q, class ST, bool GEN, bool NT, bool ROW, bool RIGID = false, bool DIFF = false, bool FRC = true>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Q == 27 ? 6 : 8)))
k_step(const typename ST::T* __restrict__ gin, typename ST::T* __restrict__ gout, StepArgs a) {
  if constexpr (GEN) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (a.finalize_prev && blockIdx.x == 0 && tid < 64)
      sum_force_slots(a.prev_partials, a.nslots, a.history, a.written, a.hist_cap, tid);
  }
  uint32_t b = blockIdx.x;
  if (a.swizzle) {
    // Hardware hands block b to XCD b % 8.  Re-number so each XCD walks one
    // contiguous eighth of the tiles: neighbouring rows then share an L2.
    const uint32_t per = a.nblocks >> 3;
    if (b < per * 8u) b = (b & 7u) * per + (b >> 3);
  }
  uint32_t salt = 0;
  if constexpr (ST::kBytes == 2) {
    salt = *a.salt_in;
    if (blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) *a.salt_out = salt + 1;
  }
  step_tile<Q, ST, GEN, NT, ROW, RIGID, DIFF, FRC>(gin, gout, a, b, salt, a.mac);
}

// Many solves in one launch (the serve loop's packed pool): block k steps
// tile tiles[k].tile of solve tiles[k].solve.  A solve steps while
// launch < remaining; its lattice parity is its base parity flipped by the
// launch index, so every launch of a captured chunk is one step of every
// active solve.
// Tile k of the batch table at launch `launch` of the chunk.
template <int Q, class ST, bool GEN, bool NT, bool ROW, bool FRC>
__device__ __forceinline__ void batch_tile(const BatchArgs& B, uint32_t k, int32_t launch) {
  // The tables are written before the launch and only read during it: read
  // them through the constant address space, so their fields are scalar loads
  // that may be repeated at will instead of values held in registers.
  using CSolve = const __attribute__((address_space(4))) BatchSolve;
  using CTile = const __attribute__((address_space(4))) BatchTile;
  CTile* tp = ((CTile*)B.tiles) + k;
  const uint32_t tile = tp->tile;
  CSolve* sp = ((CSolve*)B.solves) + tp->solve;
  CSolve& S = *sp;
  if (launch >= S.remaining) return;
  const int p = S.parity ^ (launch & 1);
  const StepArgs& a = *(const StepArgs*)(&sp->args[p]);
  if constexpr (GEN) {
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    if (tile == 0 && (launch > 0 || S.pending) && a.force_slot && tid < 64)
      sum_force_slots(a.prev_partials, a.nslots, a.history, a.written, a.hist_cap, tid);
  }
  // the dither counter of the solve is its step count
  const uint32_t salt = S.step0 + launch;
  float* mac = launch + 1 == S.remaining ? S.mac : nullptr;
  step_tile<Q, ST, GEN, NT, ROW, false, false, FRC>(static_cast<const typename ST::T*>(S.f[p]),
                                                    static_cast<typename ST::T*>(S.f[p ^ 1]), a, tile, salt, mac);
}

template <int Q, class ST, bool GEN, bool NT, bool ROW, bool FRC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Q == 27 ? 6 : 8))) k_step_batch(BatchArgs B) {
  batch_tile<Q, ST, GEN, NT, ROW, FRC>(B, blockIdx.x, B.launch);
}

// Grid barrier of the persistent chunk, with a bounded wait.  Arrivals count
// up monotonically in one word (the host zeroes it before the launch):
// barrier j is passed when it reaches (j + 1) * gridDim.x.  Each block's
// first thread, after the block's own barrier, arrives with an agent-scope
// release (its XCD's L2 writes the block's populations back, so other XCDs
// can read them) and polls the same word with relaxed loads; an agent-scope
// acquire after the wait drops this XCD's stale lines.  A wait that exceeds
// kBarrierSpins polls (seconds, far beyond any step) sets the word's top bit,
// which ends every block's wait and fails every later barrier: the launch
// always ends, and the host reports the failure.
constexpr uint32_t kBarrierSpins = 1u << 20;
constexpr uint32_t kBarrierFailed = 1u << 31;

__device__ __forceinline__ bool grid_barrier(PersistBarrier* bar, uint32_t target) {
  __shared__ uint32_t ok;
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    uint32_t v = __hip_atomic_fetch_add(&bar->arrived, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT) + 1u;
    for (uint32_t spins = 0; v < target; spins++) {
      if (spins >= kBarrierSpins) {
        __hip_atomic_fetch_or(&bar->arrived, kBarrierFailed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        v = kBarrierFailed;
        break;
      }
      __builtin_amdgcn_s_sleep(8);
      v = __hip_atomic_load(&bar->arrived, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok = (v & kBarrierFailed) ? 0u : 1u;
  }
  __syncthreads();
  return ok != 0u;
}

// The persistent form of a chunk: one cooperative launch whose blocks stay
// resident for all `steps` launches of the chunk.  Block b steps tiles b, b +
// gridDim.x, ... of the batch table, then the grid meets at grid_barrier
// before the next step (the step reads every neighbour's populations of the
// previous one).  Every block runs the same `steps` iterations, so each
// barrier is reached by all; the cooperative launch keeps them all resident.
template <int Q, class ST, bool GEN, bool NT, bool ROW, bool FRC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Q == 27 ? 6 : 8)))
k_step_persistent(BatchArgs B, uint32_t ntiles, int32_t steps, PersistBarrier* bar) {
  for (int32_t j = 0; j < steps; j++) {
    for (uint32_t k = blockIdx.x; k < ntiles; k += gridDim.x) batch_tile<Q, ST, GEN, NT, ROW, FRC>(B, k, j);
    if (j + 1 < steps && !grid_barrier(bar, (uint32_t)(j + 1) * gridDim.x)) return;
  }
}

// One wave: lane l sums slots l, l+64, ... in order, then a fixed butterfly.
__global__ void __launch_bounds__(64) k_force_finalize(ForceFinalizeArgs a) {
  sum_force_slots(a.partials, a.nslots, a.history, a.written, a.capacity, threadIdx.x);
}

template <int Q, class ST>
__global__ void __launch_bounds__(256)
k_init_eq(const float* __restrict__ mac, typename ST::T* __restrict__ g, uint32_t N) {
  using L = Lattice<Q>;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= N) return;
  const float4 m = reinterpret_cast<const float4*>(mac)[cell];
  const float usq = m.y * m.y + m.z * m.z + m.w * m.w;
  const float drho = m.x - 1.f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    const float eu = (float)L::ex[i] * m.y + (float)L::ey[i] * m.z + (float)L::ez[i] * m.w;
    const float geq = L::w[i] * (drho + m.x * (3.f * eu + 4.5f * eu * eu - 1.5f * usq));
    ST::store(block_of<Q>(g, cell) + i * kLatBlock, cell % kLatBlock, geq);
  }
}

// Wire format (f32 plane-major f) <-> storage (shifted g), f-pack.mjs:93-124.
template <int Q, class ST>
__global__ void __launch_bounds__(256)
k_pack(const float* __restrict__ wire, typename ST::T* __restrict__ g, uint32_t N) {
  using L = Lattice<Q>;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= N) return;
#pragma unroll
  for (int i = 0; i < Q; i++)
    ST::store(block_of<Q>(g, cell) + i * kLatBlock, cell % kLatBlock, wire[(size_t)i * N + cell] - L::w[i]);
}

template <int Q, class ST>
__global__ void __launch_bounds__(256)
k_unpack(const typename ST::T* __restrict__ g, float* __restrict__ wire, uint32_t N) {
  using L = Lattice<Q>;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= N) return;
#pragma unroll
  for (int i = 0; i < Q; i++)
    wire[(size_t)i * N + cell] = ST::load(block_of<Q>(g, cell) + i * kLatBlock, cell % kLatBlock) + L::w[i];
}

template <int Q, class ST>
__global__ void __launch_bounds__(256)
k_totals(const typename ST::T* __restrict__ g, TotalsArgs a) {
  using L = Lattice<Q>;
  __shared__ double lds[16 * 4];
  double v[4] = {0, 0, 0, 0};
  for (uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x; cell < a.ncells; cell += gridDim.x * blockDim.x) {
#pragma unroll
    for (int i = 0; i < Q; i++) {
      const double fi = (double)ST::load(block_of<Q>(g, cell) + i * kLatBlock, cell % kLatBlock) + (double)L::w[i];
      v[0] += fi;
      v[1] += L::ex[i] * fi;
      v[2] += L::ey[i] * fi;
      v[3] += L::ez[i] * fi;
    }
  }
  block_sum<4>(v, lds);
  partials_then_fixed_order_sum<4>(v, a.partials, a.counter, [&](int j, double s) { a.out[j] = s; });
}

// Per-cell class byte from the signed distance of the pinned body (body_phi),
// which the host oracle reproduces exactly.
__global__ void __launch_bounds__(256) k_classify(ClassifyArgs a) {
  const uint32_t n = (uint32_t)a.nx * a.ny * a.nz;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= n) return;
  const int x = cell % a.nx;
  const int y = (cell / a.nx) % a.ny;
  const int z = cell / (a.nx * a.ny);
  const float phi = body_phi(a.shape, a.radius, a.cx, a.cy, a.cz, a.nx, a.ny, x, y, z);
  uint8_t c = 0;
  if (phi < 0.f) c |= kClsSolid;
  if (phi < 2.f) c |= kClsNear;
  if (phi < a.band) c |= kClsBand;
  a.cls[cell] = c;
}

// Link masks for the bounce-back branch of the step: bit i of a cell next to
// the body is set when its pull source cell - e_i (periodic) is solid.
template <int Q>
__global__ void __launch_bounds__(256) k_links(ClassifyArgs a) {
  using L = Lattice<Q>;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const uint32_t n = (uint32_t)nx * ny * nz;
  const uint32_t cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= n) return;
  uint32_t m = 0;
  if (a.cls[cell] & 2) {
    const int x = cell % nx, y = (cell / nx) % ny, z = cell / (nx * ny);
#pragma unroll
    for (int i = 1; i < Q; i++) {
      const int sx = x - L::ex[i], sy = y - L::ey[i], sz = z - L::ez[i];
      const int wx = sx < 0 ? sx + nx : (sx >= nx ? sx - nx : sx);
      const int wy = sy < 0 ? sy + ny : (sy >= ny ? sy - ny : sy);
      const int wz = sz < 0 ? sz + nz : (sz >= nz ? sz - nz : sz);
      if (a.cls[(uint32_t)((wz * ny + wy) * nx + wx)] & 1) m |= 1u << i;
    }
  }
  a.links[cell] = m;
}

// 16-byte streaming copy, one element per thread with non-temporal loads and
// stores: the practical HBM ceiling the roofline fraction is read against.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_copy_probe(const u32x4* __restrict__ src, u32x4* __restrict__ dst, size_t n) {
  const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// ---------------------------------------------------------------------------
// dispatch: one Op<Q, Storage>::run per kernel family
// ---------------------------------------------------------------------------
template <template <int, class> class Op, class... A>
static hipError_t dispatch(int q, int storage, A... args) {
  switch (q) {
    case 9:
      if (storage == kF32) return Op<9, StoreF32>::run(args...);
      if (storage == kF16) return Op<9, StoreF16>::run(args...);
      return Op<9, StoreBF16>::run(args...);
    case 19:
      if (storage == kF32) return Op<19, StoreF32>::run(args...);
      if (storage == kF16) return Op<19, StoreF16>::run(args...);
      return Op<19, StoreBF16>::run(args...);
    default:
      if (storage == kF32) return Op<27, StoreF32>::run(args...);
      if (storage == kF16) return Op<27, StoreF16>::run(args...);
      return Op<27, StoreBF16>::run(args...);
  }
}

static inline dim3 cells_grid(uint32_t n) { return dim3((n + 255) / 256); }

// LDS bytes per CU of the current device (160 KiB on gfx950).
static size_t lds_per_cu() {
  static const size_t bytes = [] {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess || v <= 0)
      v = 160 * 1024;
    return (size_t)v;
  }();
  return bytes;
}

// Blocks per CU that give `occ` waves on each of the CU's 4 SIMDs, and the
// LDS share A = L / blocks that admits exactly that many: blocks * A <= L <
// (blocks + 1) * A.  A is a multiple of 1 KiB for the shapes used here, so
// the hardware's allocation granule does not round it up.
static size_t occ_share(int occ, int threads) {
  const int wpb = threads / 64;
  if (occ <= 0 || wpb <= 0 || (4 * occ) % wpb) return 0;
  const size_t blocks = (size_t)(4 * occ / wpb), L = lds_per_cu();
  const size_t A = L / blocks / 1024 * 1024;
  if (A == 0 || (blocks + 1) * A <= L) return 0;
  return A;
}
bool occupancy_cap_ok(int occ, int threads) { return occ == 0 || occ_share(occ, threads) != 0; }

// Dynamic LDS for a launch of kernel K under the cap: its share minus the
// kernel's own static LDS (queried once per kernel).
template <auto K>
static size_t occ_dynamic_lds(int occ, int threads) {
  const size_t A = occ_share(occ, threads);
  if (!A) return 0;
  static const size_t static_bytes = [] {
    hipFuncAttributes fa{};
    return hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(K)) == hipSuccess ? fa.sharedSizeBytes
                                                                                     : (size_t)0;
  }();
  return A > static_bytes ? A - static_bytes : 0;
}

// Runtime flags -> template arguments of the step kernels.
template <class F>
static void with_flags(bool general, bool nt, bool rows, F f) {
  auto r = [&](auto G, auto N) {
    if (rows) f(G, N, std::true_type{});
    else f(G, N, std::false_type{});
  };
  auto n = [&](auto G) {
    if (nt) r(G, std::true_type{});
    else r(G, std::false_type{});
  };
  if (general) n(std::true_type{});
  else n(std::false_type{});
}

template <int Q, class ST> struct OpStep {
  static hipError_t run(bool general, bool nt, int occ, const void* fin, void* fout, StepArgs a, int bx, int by,
                        hipStream_t s, bool diffuse) {
    using T = typename ST::T;
    const bool rows = step_row_mode(a.nx, a.linear != 0, bx, by);
    if (diffuse) {
      // the pinned body's diffuse coupling runs its own general instantiation
      with_flags(true, nt, rows, [&](auto, auto N, auto R) {
        constexpr auto k = k_step<Q, ST, true, decltype(N)::value, decltype(R)::value, false, true>;
        hipLaunchKernelGGL(k, dim3(a.nblocks), dim3(bx, by), occ_dynamic_lds<k>(occ, bx * by), s, (const T*)fin,
                           (T*)fout, a);
      });
      return hipGetLastError();
    }
    if constexpr (Q != 9) {
      if (a.rigid) {
        // a moving body runs the general kernel's rigid instantiation
        with_flags(true, nt, rows, [&](auto, auto N, auto R) {
          constexpr auto k = k_step<Q, ST, true, decltype(N)::value, decltype(R)::value, true>;
          hipLaunchKernelGGL(k, dim3(a.nblocks), dim3(bx, by), occ_dynamic_lds<k>(occ, bx * by), s, (const T*)fin,
                             (T*)fout, a);
        });
        return hipGetLastError();
      }
    }
    // a solve without a body force runs the instantiation without Guo terms
    const bool frc = a.F[0] != 0.f || a.F[1] != 0.f || a.F[2] != 0.f;
    with_flags(general, nt, rows, [&](auto G, auto N, auto R) {
      auto go = [&](auto F) {
        constexpr auto k = k_step<Q, ST, decltype(G)::value, decltype(N)::value, decltype(R)::value, false, false,
                                  decltype(F)::value>;
        hipLaunchKernelGGL(k, dim3(a.nblocks), dim3(bx, by), occ_dynamic_lds<k>(occ, bx * by), s, (const T*)fin,
                           (T*)fout, a);
      };
      if (frc) go(std::true_type{});
      else go(std::false_type{});
    });
    return hipGetLastError();
  }
};
template <int Q, class ST> struct OpStepBatch {
  static hipError_t run(bool general, bool nt, int occ, bool rows, bool force, BatchArgs b, uint32_t ntiles, int bx,
                        int by, int persistent_steps, PersistBarrier* bar, hipStream_t s) {
    hipError_t err = hipSuccess;
    with_flags(general, nt, rows, [&](auto G, auto N, auto R) {
      auto go = [&](auto F) {
        constexpr bool g = decltype(G)::value, n = decltype(N)::value, r = decltype(R)::value, f = decltype(F)::value;
        if (persistent_steps <= 0) {
          constexpr auto k = k_step_batch<Q, ST, g, n, r, f>;
          hipLaunchKernelGGL(k, dim3(ntiles), dim3(bx, by), occ_dynamic_lds<k>(occ, bx * by), s, b);
          err = hipGetLastError();
          return;
        }
        // at most 4 blocks per CU (and never more than the device holds at
        // once: the cooperative launch requires every block resident), at
        // most one per tile.  Each block's arrival at the grid barrier is an
        // agent-scope atomic on one word, so the barrier grows with the block
        // count: on a 192^3 f16 request the step took 0.432, 0.293, 0.287 and
        // 0.366 ms at 8, 4, 2 and 1 blocks per CU (0.110 ms as a graph of
        // batched launches).
        constexpr auto k = k_step_persistent<Q, ST, g, n, r, f>;
        const size_t lds = occ_dynamic_lds<k>(occ, bx * by);
        int per_cu = 0, dev = 0, cus = 0;
        if ((err = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(k), bx * by,
                                                                lds)) != hipSuccess ||
            (err = hipGetDevice(&dev)) != hipSuccess ||
            (err = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
          return;
        if (per_cu <= 0) {
          err = hipErrorInvalidConfiguration;
          return;
        }
        per_cu = std::min(per_cu, 4);
        const uint32_t grid = std::min<uint32_t>(ntiles, (uint32_t)per_cu * (uint32_t)cus);
        uint32_t nt_arg = ntiles;
        int32_t steps_arg = persistent_steps;
        if ((err = hipMemsetAsync(bar, 0, sizeof(PersistBarrier), s)) != hipSuccess) return;
        void* args[] = {&b, &nt_arg, &steps_arg, &bar};
        err = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k), dim3(grid), dim3(bx, by), args, lds, s);
      };
      if (force) go(std::true_type{});
      else go(std::false_type{});
    });
    return err;
  }
};
template <int Q, class ST> struct OpInit {
  static hipError_t run(const float* mac, void* f, uint32_t N, hipStream_t s) {
    hipLaunchKernelGGL((k_init_eq<Q, ST>), cells_grid(N), dim3(256), 0, s, mac, (typename ST::T*)f, N);
    return hipGetLastError();
  }
};
template <int Q, class ST> struct OpPack {
  static hipError_t run(const float* wire, void* f, uint32_t N, hipStream_t s) {
    hipLaunchKernelGGL((k_pack<Q, ST>), cells_grid(N), dim3(256), 0, s, wire, (typename ST::T*)f, N);
    return hipGetLastError();
  }
};
template <int Q, class ST> struct OpUnpack {
  static hipError_t run(const void* f, float* wire, uint32_t N, hipStream_t s) {
    hipLaunchKernelGGL((k_unpack<Q, ST>), cells_grid(N), dim3(256), 0, s, (const typename ST::T*)f, wire, N);
    return hipGetLastError();
  }
};
template <int Q, class ST> struct OpTotals {
  static hipError_t run(const void* f, TotalsArgs a, int nblocks, hipStream_t s) {
    hipLaunchKernelGGL((k_totals<Q, ST>), dim3(nblocks), dim3(256), 0, s, (const typename ST::T*)f, a);
    return hipGetLastError();
  }
};

hipError_t launch_step(int q, int storage, bool general, bool nontemporal, int occ, const void* fin, void* fout,
                       const StepArgs& a, int bx, int by, hipStream_t s, bool diffuse) {
  return dispatch<OpStep>(q, storage, general, nontemporal, occ, fin, fout, a, bx, by, s, diffuse);
}
hipError_t launch_step_batch(int q, int storage, bool general, bool nontemporal, int occ, bool rows, bool force,
                             const BatchArgs& b, uint32_t ntiles, int bx, int by, hipStream_t s) {
  if (ntiles == 0) return hipSuccess;
  return dispatch<OpStepBatch>(q, storage, general, nontemporal, occ, rows, force, b, ntiles, bx, by, 0,
                               (PersistBarrier*)nullptr, s);
}
hipError_t launch_step_persistent(int q, int storage, bool general, bool nontemporal, int occ, bool rows, bool force,
                                  const BatchArgs& b, uint32_t ntiles, int bx, int by, int steps, PersistBarrier* bar,
                                  hipStream_t s) {
  if (ntiles == 0 || steps <= 0) return hipSuccess;
  return dispatch<OpStepBatch>(q, storage, general, nontemporal, occ, rows, force, b, ntiles, bx, by, steps, bar, s);
}
hipError_t launch_init_eq(int q, int storage, const float* mac, void* f, uint32_t N, hipStream_t s) {
  return dispatch<OpInit>(q, storage, mac, f, N, s);
}
hipError_t launch_pack(int q, int storage, const float* wire, void* f, uint32_t N, hipStream_t s) {
  return dispatch<OpPack>(q, storage, wire, f, N, s);
}
hipError_t launch_unpack(int q, int storage, const void* f, float* wire, uint32_t N, hipStream_t s) {
  return dispatch<OpUnpack>(q, storage, f, wire, N, s);
}
hipError_t launch_totals(int q, int storage, const void* f, const TotalsArgs& a, int nblocks, hipStream_t s) {
  return dispatch<OpTotals>(q, storage, f, a, nblocks, s);
}
hipError_t launch_force_finalize(const ForceFinalizeArgs& a, hipStream_t s) {
  hipLaunchKernelGGL(k_force_finalize, dim3(1), dim3(64), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_classify(const ClassifyArgs& a, hipStream_t s) {
  const uint32_t n = (uint32_t)a.nx * a.ny * a.nz;
  hipLaunchKernelGGL(k_classify, dim3((n + 255) / 256), dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_links(int q, const ClassifyArgs& a, hipStream_t s) {
  const uint32_t n = (uint32_t)a.nx * a.ny * a.nz;
  const dim3 grid((n + 255) / 256);
  if (q == 9) hipLaunchKernelGGL(k_links<9>, grid, dim3(256), 0, s, a);
  else if (q == 19) hipLaunchKernelGGL(k_links<19>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_links<27>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, (const u32x4*)src, (u32x4*)dst, n);
  return hipGetLastError();
}

}  // namespace lbm

// Rigid-body kernels of a dense 3D solve (gfx950).
//   k_rigid_force     momentum exchange on the bounce-back links with
//                     Ladd's moving-wall term at the local wall density, and
//                     the swept-cell term (common_d3_force.wgsl:171-222),
//                     over a box around the body; FP64 partials per block,
//                     summed in block order by the last block, which then
//                     appends the force record and integrates the body
//                     (d3_physics.wgsl:56-158, d3-body.mjs:234-297)
//   k_rigid_classify  solid bit and pull-link mask of every cell near the
//                     body at its new pose, over the union of the old and new
//                     neighbourhoods, so that cells the body left are cleared
// Cells farther than rc + 2 from the centre can be neither solid nor next to
// the body (a link is at most sqrt(3) long), so both kernels leave them out.
#include <hip/hip_runtime.h>
#include "dense.h"
#include "device_common.h"
#include "lattice_tables.h"
#include "rigid.h"
#include "rigid_device.h"

namespace lbm {

namespace {

// Cell coordinate k of a box edge of `edge` cells starting `half` below
// floor(c) on an axis of n cells (the whole axis when edge >= n).
__device__ __forceinline__ int box_coord(double c, int half, int edge, int n, int k) {
  if (edge >= n) return k;
  int x = (int)floor(c) - half + k;
  x %= n;
  return x < 0 ? x + n : x;
}

template <class ST, int Q>
__device__ __forceinline__ const typename ST::T* cell_block(const typename ST::T* f, uint32_t cell) {
  return f + (size_t)(cell / kLatBlock) * (Q * kLatBlock);
}

__device__ __forceinline__ void omega_from_L(const RigidBody* B, const double q[4], const double L[3], double w[3]) {
#pragma clang fp contract(off)
  double Lb[3];
  q_rot_inv(q, L, Lb);
  const double t[3] = {Lb[0] / B->I[0], Lb[1] / B->I[1], Lb[2] / B->I[2]};
  q_rot(q, t, w);
}

__device__ __forceinline__ void q_normalize(double q[4]) {
#pragma clang fp contract(off)
  const double n = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
  if (!(n > 1e-12)) { q[0] = 1.0; q[1] = q[2] = q[3] = 0.0; return; }
  for (int k = 0; k < 4; k++) q[k] = q[k] / n;
}

// One integrator step with the measured force and torque.
__device__ void rigid_physics(RigidBody* B, const double F[3], const double T[3]) {
#pragma clang fp contract(off)
  for (int k = 0; k < 3; k++) { B->force[k] = F[k]; B->torque[k] = T[k]; }
  if (B->pinned) return;
  const double use = B->no_fluid ? 0.0 : 1.0;
  const double m = B->mass;
  double v[3];
  for (int k = 0; k < 3; k++) v[k] = B->v[k] + (use * F[k] / m + B->g[k]);
  if (B->planar) v[2] = 0.0;
  const double vmag = len3(v[0], v[1], v[2]);
  if (vmag > B->v_max) for (int k = 0; k < 3; k++) v[k] = v[k] * (B->v_max / vmag);
  double L[3];
  for (int k = 0; k < 3; k++) L[k] = B->L[k] + use * T[k];
  if (B->planar) { L[0] = 0.0; L[1] = 0.0; }
  double q[4] = {B->q[0], B->q[1], B->q[2], B->q[3]};
  double w[3];
  omega_from_L(B, q, L, w);
  if (B->planar) { w[0] = 0.0; w[1] = 0.0; }
  const double wmag = len3(w[0], w[1], w[2]);
  if (wmag > B->o_max) {
    const double s = B->o_max / wmag;
    for (int k = 0; k < 3; k++) { w[k] = w[k] * s; L[k] = L[k] * s; }
  }
  double e[4], qh[4];
  exp_map(w, 0.5, e);
  q_mul(e, q, qh);
  q_normalize(qh);
  double wh[3];
  omega_from_L(B, qh, L, wh);
  if (B->planar) { wh[0] = 0.0; wh[1] = 0.0; }
  double qn[4];
  exp_map(wh, 1.0, e);
  q_mul(e, q, qn);
  q_normalize(qn);
  for (int k = 0; k < 3; k++) {
    B->d[k] = B->d[k] + v[k];
    B->c[k] = win_wrap_pos(B->c[k] + v[k], B->win_n[k]);
    B->v[k] = v[k];
    B->L[k] = L[k];
    B->w[k] = wh[k];
  }
  for (int k = 0; k < 4; k++) B->q[k] = qn[k];
}

}  // namespace

template <int Q, class ST>
__global__ void __launch_bounds__(256) k_rigid_force(const typename ST::T* __restrict__ f, RigidArgs a) {
  using L = Lattice<Q>;
  using T = typename ST::T;
  __shared__ double lds[16 * 6];
  __shared__ bool last;
  __shared__ double ahead[7];  // the pose one step ahead: centre, orientation
  const RigidBody* B = a.body;
  const int tid = threadIdx.x;
  if (tid == 0 && B->swept) {
    // by v and the cached omega (common_d3_geometry.wgsl get_phi3Ahead)
    const double w[3] = {B->w[0], B->w[1], B->w[2]};
    const double q[4] = {B->q[0], B->q[1], B->q[2], B->q[3]};
    double e[4], qa[4];
    exp_map(w, 1.0, e);
    q_mul(e, q, qa);
    for (int k = 0; k < 3; k++) ahead[k] = B->c[k] + B->v[k];
    for (int k = 0; k < 4; k++) ahead[3 + k] = qa[k];
  }
  __syncthreads();
  double fv[6] = {0, 0, 0, 0, 0, 0};
  const uint32_t t = blockIdx.x * blockDim.x + tid;
  const uint32_t nbox = (uint32_t)a.fbox[0] * a.fbox[1] * a.fbox[2];
  if (t < nbox) {
    const int kx = t % a.fbox[0], ky = (t / a.fbox[0]) % a.fbox[1], kz = t / (a.fbox[0] * a.fbox[1]);
    const int x = box_coord(B->c[0], a.half, a.fbox[0], a.nx, kx);
    const int y = box_coord(B->c[1], a.half, a.fbox[1], a.ny, ky);
    const int z = box_coord(B->c[2], a.half, a.fbox[2], a.nz, kz);
    double r[3];
    body_delta(B, x, y, z, r);
    if (len3(r[0], r[1], r[2]) <= B->rc + 2.0) {
      const uint32_t cell = (uint32_t)((z * a.ny + y) * a.nx + x);
      const T* blk = cell_block<ST, Q>(f, cell);
      const uint32_t lo = cell % kLatBlock;
      const bool solid_now = a.cls[cell] & 1;
      const uint32_t lm = a.links[cell];
      double fx = 0, fy = 0, fz = 0;
      if (!solid_now && lm) {
        double us[3];
        body_velocity(B, x, y, z, us);
        double rho_w = 1.0;
        if (B->rho_w_local) {
          rho_w = 0.0;
          for (int i = 0; i < Q; i++) rho_w += (double)ST::load(blk + i * kLatBlock, lo) + (double)L::w[i];
        }
#pragma unroll
        for (int i = 1; i < Q; i++) {
          if (!((lm >> i) & 1u)) continue;
          const double eu = L::ex[i] * us[0] + L::ey[i] * us[1] + L::ez[i] * us[2];
          const int j = L::opp[i];
          const double m = 2.0 * ((double)ST::load(blk + j * kLatBlock, lo) + (double)L::w[j]) +
                           2.0 * (double)L::w[i] * rho_w * eu * 3.0;
          fx -= L::ex[i] * m;
          fy -= L::ey[i] * m;
          fz -= L::ez[i] * m;
        }
      }
      if (B->swept) {
        const double ca[3] = {ahead[0], ahead[1], ahead[2]};
        const double qa[4] = {ahead[3], ahead[4], ahead[5], ahead[6]};
        if (solid_now != (body_phi(B, ca, qa, x, y, z) < 0.0)) {
          double mx = 0, my = 0, mz = 0;
          for (int i = 0; i < Q; i++) {
            const double fi = (double)ST::load(blk + i * kLatBlock, lo) + (double)L::w[i];
            mx += fi * L::ex[i];
            my += fi * L::ey[i];
            mz += fi * L::ez[i];
          }
          const double s = solid_now ? -1.0 : 1.0;
          fx += s * mx;
          fy += s * my;
          fz += s * mz;
        }
      }
      fv[0] = fx; fv[1] = fy; fv[2] = fz;
      fv[3] = r[1] * fz - r[2] * fy;
      fv[4] = r[2] * fx - r[0] * fz;
      fv[5] = r[0] * fy - r[1] * fx;
    }
  }
  block_sum<6>(fv, lds);
  const unsigned nblocks = gridDim.x;
  if (tid == 0) {
    for (int k = 0; k < 6; k++) a.partials[(size_t)blockIdx.x * 6 + k] = fv[k];
    __threadfence();
    last = atomicAdd(a.counter, 1u) == nblocks - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // wave 0: lane l sums blocks l, l + 64, ... in order, then a fixed
  // butterfly, so the totals repeat bit for bit
  if (tid >= 64) return;
  double tot[6];
  for (int k = 0; k < 6; k++) {
    double sk = 0.0;
    for (unsigned j = tid; j < nblocks; j += 64)
      sk += __hip_atomic_load(&a.partials[(size_t)j * 6 + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tot[k] = wave_sum(sk);
  }
  if (tid == 0) {
    *a.counter = 0u;
    const unsigned long long n = *a.written;
    for (int k = 0; k < 6; k++) a.history[(size_t)(n % a.hist_cap) * 6 + k] = tot[k];
    *a.written = n + 1;
    const double F[3] = {tot[0], tot[1], tot[2]}, Tq[3] = {tot[3], tot[4], tot[5]};
    rigid_physics(a.body, F, Tq);
  }
}

template <int Q>
__global__ void __launch_bounds__(256) k_rigid_classify(RigidArgs a) {
  using L = Lattice<Q>;
  const RigidBody* B = a.body;
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t nbox = (uint32_t)a.cbox[0] * a.cbox[1] * a.cbox[2];
  if (t >= nbox) return;
  const int kx = t % a.cbox[0], ky = (t / a.cbox[0]) % a.cbox[1], kz = t / (a.cbox[0] * a.cbox[1]);
  // the box starts below both the new centre and the previous one (c - v)
  const int x = box_coord(fmin(B->c[0], B->c[0] - B->v[0]), a.half, a.cbox[0], a.nx, kx);
  const int y = box_coord(fmin(B->c[1], B->c[1] - B->v[1]), a.half, a.cbox[1], a.ny, ky);
  const int z = box_coord(fmin(B->c[2], B->c[2] - B->v[2]), a.half, a.cbox[2], a.nz, kz);
  const uint32_t cell = (uint32_t)((z * a.ny + y) * a.nx + x);
  double r[3];
  body_delta(B, x, y, z, r);
  uint8_t cl = 0;
  uint32_t lm = 0;
  if (len3(r[0], r[1], r[2]) <= B->rc + 2.0) {
    const double c[3] = {B->c[0], B->c[1], B->c[2]};
    const double q[4] = {B->q[0], B->q[1], B->q[2], B->q[3]};
    const double phi = body_phi(B, c, q, x, y, z);
    if (phi < 0.0) {
      cl = 1;
    } else if (phi < 1.75) {
      // a signed distance is 1-Lipschitz and a link at most sqrt(3) long, so
      // only cells this close can pull from a solid cell
      for (int i = 1; i < Q; i++)
        if (body_phi(B, c, q, x - L::ex[i], y - L::ey[i], z - L::ez[i]) < 0.0) lm |= 1u << i;
      if (lm) cl = 2;
    }
  }
  a.cls[cell] = cl;
  a.links[cell] = lm;
}

uint32_t rigid_force_blocks(const RigidArgs& a) {
  return ((uint32_t)a.fbox[0] * a.fbox[1] * a.fbox[2] + 255) / 256;
}

template <int Q>
static hipError_t force_q(int storage, const void* f, const RigidArgs& a, hipStream_t s) {
  const dim3 grid(rigid_force_blocks(a));
  if (storage == kF32) hipLaunchKernelGGL((k_rigid_force<Q, StoreF32>), grid, dim3(256), 0, s, (const float*)f, a);
  else if (storage == kF16)
    hipLaunchKernelGGL((k_rigid_force<Q, StoreF16>), grid, dim3(256), 0, s, (const uint16_t*)f, a);
  else hipLaunchKernelGGL((k_rigid_force<Q, StoreBF16>), grid, dim3(256), 0, s, (const uint16_t*)f, a);
  return hipGetLastError();
}

hipError_t launch_rigid_force(int q, int storage, const void* f, const RigidArgs& a, hipStream_t s) {
  return q == 19 ? force_q<19>(storage, f, a, s) : force_q<27>(storage, f, a, s);
}

hipError_t launch_rigid_classify(int q, const RigidArgs& a, hipStream_t s) {
  const dim3 grid(((uint32_t)a.cbox[0] * a.cbox[1] * a.cbox[2] + 255) / 256);
  if (q == 19) hipLaunchKernelGGL(k_rigid_classify<19>, grid, dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_rigid_classify<27>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}

}  // namespace lbm

// The collision of one cell in shifted variables g_i = f_i - w_i, shared by
// the pool kernels (the dense kernel inlines the same arithmetic).  Pair form:
// with j = opp(i), S = g_i + g_j, D = g_i - g_j and the equilibrium split
// into A = w (rho - 1 + rho (4.5 (e.u)^2 - 1.5 u.u)) and B = 3 w rho (e.u),
//   P = S/2 - omega   (S/2 - A) + (1 - omega/2)   w (9 (e.u)(e.F) - 3 u.F)
//   M = D/2 - omega_m (D/2 - B) + (1 - omega_m/2) 3 w (e.F)
//   g_i' = P + M,  g_j' = P - M
// (BGK for omega_m = omega; TRT per plans/TRT.md:333-349).
#pragma once
#include <hip/hip_runtime.h>
#include "lattice_pairs.h"
#include "lattice_tables.h"

namespace lbm {

struct CollideParams {
  float omega, omega_m, guo;  // guo = 1 - omega/2
  float F[3];
  float c1[13];               // 9 (1 - omega/2) w (e.F) per pair
  float so[13];               // (1 - omega_m/2) 3 w (e.F) per pair
};

// e_i . (x, y, z) for a direction known at compile time: the zero components
// are left out and the sum starts from -0, the identity of floating-point
// addition, so an axis direction costs no instruction and a diagonal one add.
template <int Q>
__device__ __forceinline__ float edot(int i, float x, float y, float z) {
  using L = Lattice<Q>;
  float s = -0.f;
  if (L::ex[i] == 1) s += x; else if (L::ex[i] == -1) s -= x;
  if (L::ey[i] == 1) s += y; else if (L::ey[i] == -1) s -= y;
  if (L::ez[i] == 1) s += z; else if (L::ez[i] == -1) s -= z;
  return s;
}

// 1 / x for a normal x: the hardware reciprocal (1 ulp) refined by one
// Newton step, r + r (1 - x r), whose error is a small fraction of an ulp
// and of either sign, so it leaves no systematic bias in rho u / rho.  A
// correctly rounded division costs about three times the instructions.
__device__ __forceinline__ float recip_nr(float x) {
  const float r = __builtin_amdgcn_rcpf(x);
  return __builtin_fmaf(__builtin_fmaf(-x, r, 1.f), r, r);
}

// Moments of g: rho - 1, and rho u with the Guo half-force added.
template <int Q>
__device__ __forceinline__ void moments(const float (&g)[Q], const float* F, float& drho, float& rho, float& ux,
                                        float& uy, float& uz) {
  using L = Lattice<Q>;
  float d = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
#pragma unroll
  for (int i = 0; i < Q; i++) {
    d += g[i];
    if (L::ex[i] == 1) mx += g[i]; else if (L::ex[i] == -1) mx -= g[i];
    if (L::ey[i] == 1) my += g[i]; else if (L::ey[i] == -1) my -= g[i];
    if (L::ez[i] == 1) mz += g[i]; else if (L::ez[i] == -1) mz -= g[i];
  }
  drho = d;
  rho = 1.f + d;
  const float rinv = 1.0f / fmaxf(rho, 1e-6f);
  ux = (mx + 0.5f * F[0]) * rinv;
  uy = (my + 0.5f * F[1]) * rinv;
  uz = (mz + 0.5f * F[2]) * rinv;
}

// Collides g in place.
template <int Q>
__device__ __forceinline__ void collide(float (&g)[Q], const CollideParams& p, float drho, float rho, float ux,
                                        float uy, float uz) {
  using L = Lattice<Q>;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float uF = ux * p.F[0] + uy * p.F[1] + uz * p.F[2];
  const float om = p.omega, omm = p.omega_m;
  const float c0 = drho - 1.5f * rho * usq;
  const float g3uF = 3.f * p.guo * uF;
  g[0] = g[0] - om * (g[0] - L::w[0] * c0) - L::w[0] * g3uF;
#pragma unroll
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    const int j = L::opp[i];
    const float w = L::w[i];
    const float eu = (float)L::ex[i] * ux + (float)L::ey[i] * uy + (float)L::ez[i] * uz;
    const float t = rho * eu;
    const float A = w * (c0 + 4.5f * (t * eu));
    const float B = w * (3.f * t);
    const float hS = 0.5f * (g[i] + g[j]), hD = 0.5f * (g[i] - g[j]);
    const float P = hS - om * (hS - A) + (p.c1[k] * eu - w * g3uF);
    const float M = hD - omm * (hD - B) + p.so[k];
    g[i] = P + M;
    g[j] = P - M;
  }
}

// feq(rho, u) - w for direction i (the shifted equilibrium).
template <int Q>
__device__ __forceinline__ float geq(int i, float rho, float ux, float uy, float uz) {
  using L = Lattice<Q>;
  const float eu = (float)L::ex[i] * ux + (float)L::ey[i] * uy + (float)L::ez[i] * uz;
  const float usq = ux * ux + uy * uy + uz * uz;
  return L::w[i] * ((rho - 1.f) + rho * (3.f * eu + 4.5f * eu * eu - 1.5f * usq));
}

}  // namespace lbm

// The collision of one cell in shifted variables g_i = f_i - w_i, shared by
// the pool kernels (the dense kernel inlines the same arithmetic).  Pair form:
// with j = opp(i), S = g_i + g_j, D = g_i - g_j and the equilibrium split
// into A = w (rho - 1 + rho (4.5 (e.u)^2 - 1.5 u.u)) and B = 3 w rho (e.u),
//   P = S/2 - omega   (S/2 - A) + (1 - omega/2)   w (9 (e.u)(e.F) - 3 u.F)
//   M = D/2 - omega_m (D/2 - B) + (1 - omega_m/2) 3 w (e.F)
//   g_i' = P + M,  g_j' = P - M
// (BGK for omega_m = omega; TRT per plans/TRT.md:333-349), evaluated in
// doubled variables (collide_pairs).
#pragma once
#include <hip/hip_runtime.h>
#include "lattice_pairs.h"
#include "lattice_tables.h"

namespace lbm {

struct CollideParams {
  float omega, omega_m, guo;  // guo = 1 - omega/2
  float F[3];
  float c1[13];               // 2 x 9 (1 - omega/2) w (e.F) per pair (doubled)
  float so[13];               // 2 x (1 - omega_m/2) 3 w (e.F) per pair (doubled)
  float wsum_m1;              // (sum of the f32 weights) - 1: nonzero for D2Q9 alone
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

// Pair sums S_k = g_i + g_j and differences D_k = g_i - g_j (j = opp i).
template <int Q>
__device__ __forceinline__ void pair_sums(const float (&g)[Q], float (&S)[Pairs<Q>::NP], float (&D)[Pairs<Q>::NP]) {
  using L = Lattice<Q>;
#pragma unroll
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    S[k] = g[i] + g[L::opp[i]];
    D[k] = g[i] - g[L::opp[i]];
  }
}

// Moments from the pair sums: rho - 1, and rho u with the Guo half-force
// added (FRC false: no body force).  The momentum sums start from -0, the
// identity of floating-point addition.
template <int Q, bool FRC = true>
__device__ __forceinline__ void pair_moments(float g0, const float (&S)[Pairs<Q>::NP], const float (&D)[Pairs<Q>::NP],
                                             const float* F, float& drho, float& rho, float& ux, float& uy,
                                             float& uz) {
  using L = Lattice<Q>;
  float d = g0, mx = -0.f, my = -0.f, mz = -0.f;
#pragma unroll
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    d += S[k];
    if (L::ex[i] == 1) mx += D[k]; else if (L::ex[i] == -1) mx -= D[k];
    if (L::ey[i] == 1) my += D[k]; else if (L::ey[i] == -1) my -= D[k];
    if (L::ez[i] == 1) mz += D[k]; else if (L::ez[i] == -1) mz -= D[k];
  }
  drho = d;
  rho = 1.f + d;
  const float rinv = recip_nr(fmaxf(rho, 1e-6f));
  if constexpr (FRC) {
    ux = (mx + 0.5f * F[0]) * rinv;
    uy = (my + 0.5f * F[1]) * rinv;
    uz = (mz + 0.5f * F[2]) * rinv;
  } else {
    ux = mx * rinv;
    uy = my * rinv;
    uz = mz * rinv;
  }
}

// Moments of g: rho - 1, and rho u with the Guo half-force added.
template <int Q>
__device__ __forceinline__ void moments(const float (&g)[Q], const float* F, float& drho, float& rho, float& ux,
                                        float& uy, float& uz) {
  float S[Pairs<Q>::NP], D[Pairs<Q>::NP];
  pair_sums<Q>(g, S, D);
  pair_moments<Q>(g[0], S, D, F, drho, rho, ux, uy, uz);
}

// Collides g in place from its pair sums, in doubled variables: P2 = 2P and
// M2 = 2M from S and D with 2w in place of w and the pair constants doubled
// (CollideParams holds them so), then g_i' = fma(1/2, P2, M2/2) and g_j' =
// fma(1/2, P2, -M2/2).  Scaling by two commutes with rounding, so these are
// bit for bit P + M and P - M.  FRC false: no body force, no Guo terms.
template <int Q, bool FRC = true>
__device__ __forceinline__ void collide_pairs(float (&g)[Q], const float (&S)[Pairs<Q>::NP],
                                              const float (&D)[Pairs<Q>::NP], const CollideParams& p, float drho,
                                              float rho, float ux, float uy, float uz) {
  using L = Lattice<Q>;
  const float usq = ux * ux + uy * uy + uz * uz;
  const float om = p.omega, omm = p.omega_m;
  const float c0 = drho - 1.5f * rho * usq;
  // 2 (3 (1 - omega/2) u.F)
  const float g3uF2 = FRC ? 6.f * p.guo * (ux * p.F[0] + uy * p.F[1] + uz * p.F[2]) : 0.f;
  g[0] = g[0] - om * (g[0] - L::w[0] * c0);
  if constexpr (FRC) g[0] -= L::w[0] * (0.5f * g3uF2);
#pragma unroll
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    const int j = L::opp[i];
    const float w = L::w[i], w2 = 2.f * L::w[i];
    const float eu = edot<Q>(i, ux, uy, uz);
    const float t = rho * eu;
    const float X = c0 + 4.5f * (t * eu);
    float P2 = S[k] - om * (S[k] - w2 * X);
    float M2 = D[k] - omm * (D[k] - w2 * (3.f * t));
    if constexpr (FRC) {
      P2 += p.c1[k] * eu - w * g3uF2;
      M2 += p.so[k];
    }
    const float hM = 0.5f * M2;
    g[i] = __builtin_fmaf(0.5f, P2, hM);
    g[j] = __builtin_fmaf(0.5f, P2, -hM);
  }
}

// Collides g in place.
template <int Q>
__device__ __forceinline__ void collide(float (&g)[Q], const CollideParams& p, float drho, float rho, float ux,
                                        float uy, float uz) {
  float S[Pairs<Q>::NP], D[Pairs<Q>::NP];
  pair_sums<Q>(g, S, D);
  collide_pairs<Q>(g, S, D, p, drho, rho, ux, uy, uz);
}

// feq(rho, u) - w for direction i (the shifted equilibrium).
template <int Q>
__device__ __forceinline__ float geq(int i, float rho, float ux, float uy, float uz) {
  using L = Lattice<Q>;
  const float eu = edot<Q>(i, ux, uy, uz);
  const float usq = ux * ux + uy * uy + uz * uz;
  return L::w[i] * ((rho - 1.f) + rho * (3.f * eu + 4.5f * eu * eu - 1.5f * usq));
}

}  // namespace lbm

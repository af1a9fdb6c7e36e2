// Device geometry of the rigid body, FP64 throughout: the moving window's
// wraps (d3-window.mjs:96-162, common_d3_window.wgsl), the body-relative
// position and the rigid velocity (common_d3_geometry.wgsl bodyDelta3 and
// bodyVelocity3), quaternions [w, x, y, z] (d3-body.mjs:161-218) and the
// signed distances (d3-body.mjs:56-101).  Evaluated without contraction, so
// each operation rounds as the FP64 host restatement's does.
#pragma once
#include <hip/hip_runtime.h>
#include "rigid.h"

namespace lbm {

// First statement of every function below: no contraction into FMAs.
#define LBM_NO_CONTRACT _Pragma("clang fp contract(off)")

// Nearest image along a windowed axis (n > 0); JS Math.round semantics.
__device__ __forceinline__ double win_wrap_delta(double d, double n) {
  LBM_NO_CONTRACT
  return n > 0.0 ? d - n * floor(d / n + 0.5) : d;
}
__device__ __forceinline__ double win_wrap_pos(double c, double n) {
  LBM_NO_CONTRACT
  return n > 0.0 ? c - n * floor(c / n) : c;
}
// Window coordinate of a cell coordinate p (the sponge's frame).
__device__ __forceinline__ double win_coord(double p, double c, double anchor, double n) {
  LBM_NO_CONTRACT
  if (!(n > 0.0)) return p;
  const double w = p - (floor(c) - anchor);
  return w - n * floor(w / n);
}

__device__ __forceinline__ void body_delta(const RigidBody* B, double x, double y, double z, double r[3]) {
  LBM_NO_CONTRACT
  r[0] = win_wrap_delta(x - B->c[0], B->win_n[0]);
  r[1] = win_wrap_delta(y - B->c[1], B->win_n[1]);
  r[2] = win_wrap_delta(z - B->c[2], B->win_n[2]);
}

// u_body = v + w x r at cell (x, y, z).
__device__ __forceinline__ void body_velocity(const RigidBody* B, double x, double y, double z, double u[3]) {
  LBM_NO_CONTRACT
  double r[3];
  body_delta(B, x, y, z, r);
  u[0] = B->v[0] + (B->w[1] * r[2] - B->w[2] * r[1]);
  u[1] = B->v[1] + (B->w[2] * r[0] - B->w[0] * r[2]);
  u[2] = B->v[2] + (B->w[0] * r[1] - B->w[1] * r[0]);
}

// The step kernel's binary32 u_body (the offset from the centre in FP64).
__device__ __forceinline__ void body_velocity_f(const RigidBody* B, int x, int y, int z, float& ux, float& uy,
                                                float& uz) {
  const float rx = (float)win_wrap_delta((double)x - B->c[0], B->win_n[0]);
  const float ry = (float)win_wrap_delta((double)y - B->c[1], B->win_n[1]);
  const float rz = (float)win_wrap_delta((double)z - B->c[2], B->win_n[2]);
  const float wx = (float)B->w[0], wy = (float)B->w[1], wz = (float)B->w[2];
  ux = (float)B->v[0] + (wy * rz - wz * ry);
  uy = (float)B->v[1] + (wz * rx - wx * rz);
  uz = (float)B->v[2] + (wx * ry - wy * rx);
}
__device__ __forceinline__ void q_mul(const double p[4], const double q[4], double o[4]) {
  LBM_NO_CONTRACT
  o[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  o[1] = p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2];
  o[2] = p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1];
  o[3] = p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0];
}
// v_body = q* v q
__device__ __forceinline__ void q_rot_inv(const double q[4], const double v[3], double o[3]) {
  LBM_NO_CONTRACT
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  const double vq[4] = {0.0, v[0], v[1], v[2]};
  double t[4], u[4];
  q_mul(qc, vq, t);
  q_mul(t, q, u);
  o[0] = u[1]; o[1] = u[2]; o[2] = u[3];
}
// v_world = q v q*
__device__ __forceinline__ void q_rot(const double q[4], const double v[3], double o[3]) {
  LBM_NO_CONTRACT
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  const double vq[4] = {0.0, v[0], v[1], v[2]};
  double t[4], u[4];
  q_mul(q, vq, t);
  q_mul(t, qc, u);
  o[0] = u[1]; o[1] = u[2]; o[2] = u[3];
}
__device__ __forceinline__ double len3(double x, double y, double z) {
  LBM_NO_CONTRACT
  return sqrt(x * x + y * y + z * z);
}
// Rotation by |w| dt about w (identity for a vanishing rate).
__device__ __forceinline__ void exp_map(const double w[3], double dt, double o[4]) {
  LBM_NO_CONTRACT
  const double m = len3(w[0], w[1], w[2]);
  if (!(m > 1e-12)) { o[0] = 1.0; o[1] = o[2] = o[3] = 0.0; return; }
  const double ang = m * dt;
  const double s = sin(ang / 2.0) / m;
  o[0] = cos(ang / 2.0); o[1] = w[0] * s; o[2] = w[1] * s; o[3] = w[2] * s;
}

__device__ __forceinline__ double sd_round_box(const double p[3], double a, double b, double c, double r) {
  LBM_NO_CONTRACT
  const double qx = fabs(p[0]) - (a - r), qy = fabs(p[1]) - (b - r), qz = fabs(p[2]) - (c - r);
  const double outside = len3(fmax(qx, 0.0), fmax(qy, 0.0), fmax(qz, 0.0));
  const double inside = fmin(fmax(qx, fmax(qy, qz)), 0.0);
  return outside + inside - r;
}
// Meridional-ellipse closest point by three clamped Newton steps on the
// angle parameter, seeded at atan2(y a, x c).
__device__ __forceinline__ double sd_spheroid(const double p[3], double a, double c) {
  LBM_NO_CONTRACT
  const double r0 = sqrt(p[0] * p[0] + p[1] * p[1]);
  if (fabs(a - c) <= 1e-6 * fmax(a, c)) return sqrt(r0 * r0 + p[2] * p[2]) - a;
  const double x = fabs(r0), y = fabs(p[2]);
  const bool inside = sqrt((x * x) / (a * a) + (y * y) / (c * c)) < 1.0;
  double t = atan2(y * a, x * c);
  for (int k = 0; k < 3; k++) {
    const double ct = cos(t), st = sin(t);
    const double ex = x - a * ct, ey = y - c * st;
    const double F = ex * (-a * st) + ey * (c * ct);
    const double Fp = -(a * a * st * st + c * c * ct * ct) + ex * (-a * ct) + ey * (-c * st);
    t -= F / (fabs(Fp) < 1e-9 ? -1e-9 : Fp);
    t = fmin(fmax(t, 0.0), 1.5707963267948966);
  }
  const double dx = x - a * cos(t), dy = y - c * sin(t);
  const double d = sqrt(dx * dx + dy * dy);
  return inside ? -d : d;
}
__device__ __forceinline__ double sdf_body(const RigidBody* B, const double p[3]) {
  LBM_NO_CONTRACT
  if (B->shape == kRigidSphere) return len3(p[0], p[1], p[2]) - B->a;
  if (B->shape == kRigidSpheroid) return sd_spheroid(p, B->a, B->cz);
  return sd_round_box(p, B->a, B->b, B->cz, B->r);
}

// Signed distance of cell coordinate (x, y, z) at pose (c, q).
__device__ __forceinline__ double body_phi(const RigidBody* B, const double c[3], const double q[4], double x,
                                           double y, double z) {
  LBM_NO_CONTRACT
  const double d[3] = {win_wrap_delta(x - c[0], B->win_n[0]), win_wrap_delta(y - c[1], B->win_n[1]),
                       win_wrap_delta(z - c[2], B->win_n[2])};
  double lp[3];
  q_rot_inv(q, d, lp);
  return sdf_body(B, lp);
}

#undef LBM_NO_CONTRACT

}  // namespace lbm

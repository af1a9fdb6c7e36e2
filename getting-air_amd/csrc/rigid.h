// Moving rigid body of a dense 3D solve: the device record (FP64 state and
// parameters) and the launch interface of rigid_kernels.hip.
//
// Per step, in stream order: k_rigid_force (momentum exchange with the
// moving-wall term and the swept-cell term on the pre-stream populations at
// the current pose; its last block sums the blocks' partials in a fixed
// order, appends the force record and integrates the body), k_rigid_classify
// (solid bits and pull-link masks at the new pose, over the union of the old
// and new neighbourhoods), then the fused step kernel, which reads the body
// for the moving-wall term, the solid interior's feq(1, u_body) and the
// sponge's window coordinates.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

enum RigidShape { kRigidSphere = 0, kRigidSpheroid = 1, kRigidRoundBox = 2 };

struct RigidBody {
  // state (the integrator writes these)
  double c[3];        // centre of mass, wrapped into the buffer on windowed axes
  double q[4];        // orientation [w, x, y, z], body to world
  double v[3];        // velocity, cells per step
  double L[3];        // angular momentum (world)
  double w[3];        // angular velocity cached at the half step (world)
  double d[3];        // displacement since attach (unwrapped)
  double force[3];    // last measured fluid force and torque
  double torque[3];
  // parameters
  double mass, I[3];  // principal moments
  double a, b, cz, r; // semi-axes / half-extents (x, y, z) and corner radius
  double rc;          // circumradius
  double g[3];        // acceleration (gravity, buoyancy-corrected)
  double v_max, o_max;
  double win_n[3];    // windowed axes: domain size, else 0
  double win_a[3];    // window anchors
  int32_t shape, pinned, planar, no_fluid, swept, rho_w_local;
  int32_t pad[2];
};

struct RigidArgs {
  int32_t nx, ny, nz;
  int32_t half;            // neighbourhood half-width: ceil(rc) + 3
  int32_t fbox[3];         // force box edge per axis (at most the domain)
  int32_t cbox[3];         // classification box edge per axis
  RigidBody* body;
  uint8_t* cls;            // bit 0 solid, bit 1 fluid with a solid pull source
  uint32_t* links;         // bit i: the pull source of direction i is solid
  double* partials;        // 6 per force block
  unsigned* counter;       // zero between launches
  double* history;         // force record ring [fx, fy, fz, tx, ty, tz]
  unsigned long long* written;
  uint32_t hist_cap;
};

hipError_t launch_rigid_force(int q, int storage, const void* f, const RigidArgs& a, hipStream_t s);
hipError_t launch_rigid_classify(int q, const RigidArgs& a, hipStream_t s);
// Blocks of the force launch (the partials buffer holds 6 doubles per block).
uint32_t rigid_force_blocks(const RigidArgs& a);

}  // namespace lbm

// Launch interface between the dense-lattice kernels (dense_kernels.hip) and
// the host engine (engine.cpp).  Plain structs passed by value as kernel
// arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

enum Storage { kF32 = 0, kF16 = 1, kBF16 = 2 };

// Everything one fused stream-collide launch needs besides the two lattices.
struct StepArgs {
  int32_t nx, ny, nz;
  uint32_t ncells;
  float omega;       // even (BGK) rate
  float dw;          // TRT: omega_minus - omega; 0 selects plain BGK
  float guo;         // 1 - omega/2
  float wsum_m1;     // (sum of the f32 weights) - 1: 0 for D3Q19/D3Q27, ~1e-8 for D2Q9
  float F[3];        // uniform Guo body force
  int32_t wall_mask; // bit a: halfway walls on the two faces normal to axis a
  float wall_u0, wall_u1;
  float sponge_w;    // <= 0 disables
  int32_t sponge_mask;
  float su[3];
  const uint8_t* cls;  // per-cell class (bit0 solid, bit1 near); null: no body
  int32_t solid_eq;
  int32_t solid_mac_rest;  // 3D: a solid cell reports [1, 0, 0, 0]
  float* mac;          // [rho, ux, uy, uz] per cell, or null (not observed)
  // tiling of the 1-D grid onto (x-tiles, y-tiles, z)
  int32_t ntx, nty;
  uint32_t nblocks;
  int32_t swizzle;     // 1: consecutive tiles on one XCD
};

// Momentum-exchange force over the list of fluid cells that have at least
// one link whose streaming source is solid.
struct ForceArgs {
  const uint32_t* cells;   // ascending cell index
  const uint32_t* masks;   // bit i: the source of direction i is solid
  uint32_t count;
  uint32_t ncells;
  int32_t nx, ny, nz;
  double cx, cy, cz;
  int32_t periodic_arm;    // 2D: torque arm is the nearest periodic image
  double* partials;
  unsigned* counter;
  double* history;         // ring of [fx, fy, fz, tx, ty, tz]
  unsigned long long* written;  // records written since reset
  uint32_t capacity;
};

struct TotalsArgs {
  uint32_t ncells;
  double* partials;
  unsigned* counter;
  double* out;             // [mass, px, py, pz]
};

struct ClassifyArgs {
  int32_t nx, ny, nz;
  int32_t shape;           // 1 sphere, 2 circle (periodic nearest image)
  float radius;
  float cx, cy, cz;
  uint8_t* cls;
};

hipError_t launch_step(int q, int storage, bool general, const void* fin, void* fout,
                       const StepArgs& a, int bx, int by, hipStream_t s);
hipError_t launch_force(int q, int storage, const void* fin, const ForceArgs& a, hipStream_t s);
hipError_t launch_totals(int q, int storage, const void* f, const TotalsArgs& a, int nblocks, hipStream_t s);
hipError_t launch_init_eq(int q, int storage, const float* mac, void* f, uint32_t ncells, hipStream_t s);
hipError_t launch_pack(int q, int storage, const float* wire, void* f, uint32_t ncells, hipStream_t s);
hipError_t launch_unpack(int q, int storage, const void* f, float* wire, uint32_t ncells, hipStream_t s);
hipError_t launch_classify(const ClassifyArgs& a, hipStream_t s);

}  // namespace lbm

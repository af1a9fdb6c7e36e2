// Launch interface between the dense-lattice kernels (dense_kernels.hip) and
// the host engine (engine.cpp).  Plain structs passed by value as kernel
// arguments.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

enum Storage { kF32 = 0, kF16 = 1, kBF16 = 2 };

struct RigidBody;  // rigid.h

// floor(n / d) for 0 <= n < 2^31 as a multiply and a shift: with
// l = ceil(log2 d) and m = ceil(2^(31+l) / d), 2^(31+l) <= m d <= 2^(31+l) + 2^l,
// which is the condition of Granlund & Montgomery ("Division by invariant
// integers using multiplication", PLDI 1994, Theorem 4.2) for N = 31.
// m < 2^32 for every d <= 2^31, so n m fits 64 bits.
struct InvariantDiv {
  uint32_t m, shift;
};
inline InvariantDiv make_invariant_div(uint32_t d) {
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  return InvariantDiv{(uint32_t)(((1ull << (31 + l)) + d - 1) / d), 31 + l};
}
__host__ __device__ inline uint32_t div_invariant(uint32_t n, InvariantDiv v) {
  return (uint32_t)(((uint64_t)n * v.m) >> v.shift);
}

// Lattice layout.  Cells (flat index c = (z ny + y) nx + x) are grouped into
// blocks of kLatBlock consecutive cells, and a block stores its Q populations
// as Q runs of kLatBlock elements: population i of cell c lies at element
//   ((c / kLatBlock) Q + i) kLatBlock + c % kLatBlock.
// A step's 2Q streams then walk a few contiguous regions of HBM (the blocks
// of the source rows and of the own row) instead of 2Q planes a lattice-size
// apart, whose concurrent accesses crowd the same DRAM channels and banks.
// A block is one wave's cells: a wave's Q stores fill one contiguous span of
// Q * 256 B (f32), and each of its loads is one 256 B (f32) or 128 B (16-bit)
// span.  Measured with a stream-only probe of the D3Q19 pull on 256x128x128
// (19 x-shifted loads, 19 stores per cell): blocks of 32 / 64 / 128 / 256 /
// 1024 cells moved 6.29 / 6.31 / 6.12 / 6.09 / 6.13 TB/s, plane-major 5.26.
constexpr uint32_t kLatBlock = 64;
__host__ __device__ constexpr size_t lattice_elems(int q, uint32_t ncells) {
  return (size_t)((ncells + kLatBlock - 1) / kLatBlock) * (size_t)q * kLatBlock;
}
// Element of population i of cell c.
__host__ __device__ inline size_t lat_elem(int q, int i, uint32_t c) {
  return ((size_t)(c / kLatBlock) * q + i) * kLatBlock + c % kLatBlock;
}

// Everything one fused stream-collide launch needs besides the two lattices.
struct StepArgs {
  int32_t nx, ny, nz;
  uint32_t ncells;
  float omega;       // even (BGK) rate
  float guo;         // 1 - omega/2
  float wsum_m1;     // (sum of the f32 weights) - 1: 0 for D3Q19/D3Q27, ~1e-8 for D2Q9
  float F[3];        // uniform Guo body force
  // Collision constants, uniform over the grid (computed on the host).
  // Pair k couples direction i = lo(k) with opp(i); see k_step.  The pair
  // constants are doubled, as the kernel's doubled pair form uses them.
  float omega_m;     // odd rate (TRT omega_minus; = omega for BGK)
  float c1[13];      // 2 x 9 (1 - omega/2) w_i (e_i . F)
  float so[13];      // 2 x (1 - omega_m/2) 3 w_i (e_i . F): odd Guo source
  float ae[14];      // sponge target, even part 2 x w_i (4.5 (e_i.U)^2 - 1.5 U.U); [13] = rest, not doubled
  float bo[13];      // sponge target, odd part 2 x 3 w_i (e_i . U)
  int32_t wall_mask; // bit a: halfway walls on the two faces normal to axis a
  float wall_u0, wall_u1;
  float sponge_w;    // <= 0 disables
  int32_t sponge_mask;
  float su[3];
  const uint8_t* cls;  // per-cell class (bit0 solid, bit1 near); null: no body
  // Per-cell link mask, read only where cls bit1 is set: bit i is set when
  // the streaming source of direction i (cell - e_i, periodic) is solid.
  const uint32_t* links;
  int32_t solid_eq;
  int32_t solid_mac_rest;  // 3D: a solid cell reports [1, 0, 0, 0]
  // Moving rigid body (3D; null otherwise): cls/links hold its pose, the
  // body links take Ladd's moving-wall term at the local wall density, the
  // solid interior is held at feq(1, u_body) and reports [1, u_body], and
  // the sponge is measured in window coordinates.
  const RigidBody* rigid;
  float* mac;          // [rho, ux, uy, uz] per cell, or null (not observed)
  // Block refinement map of an AMR hierarchy: a cell whose block holds a
  // slot is solved by the finer level and skipped here
  // (coveredByFiner, common_d3_step.wgsl:104-114).  Null without AMR.
  const int32_t* cover;
  int32_t cover_rb, cover_nbx, cover_nby;
  // Momentum-exchange force, fused into the step: tiles holding a fluid cell
  // with a solid streaming source own a slot of 6 FP64 partials.
  const int32_t* force_slot;  // per logical tile: slot index or -1; null: no force
  double* force_partials;     // [slot][6] of this step (lattice parity buffer)
  // The previous step's partials are summed by the first wave of block 0 of
  // this launch (in the same fixed order as k_force_finalize), so a run of
  // steps needs no separate reduction launch per step.
  int32_t finalize_prev;
  const double* prev_partials;
  uint32_t nslots;
  double* history;
  unsigned long long* written;
  uint32_t hist_cap;
  double body_c[3];           // torque arm origin
  int32_t periodic_arm;       // 2D: nearest periodic image for the arm
  // The diffuse penalty coupling of the pinned body (the DIFF kernels, in
  // place of bounce-back): its signed distance as k_classify evaluates it and
  // chi = 1/2 (1 - tanh(clamp(phi / chi_eps, -20, 20))).
  int32_t body_shape;         // 1 sphere, 2 circle (periodic nearest image)
  float body_r, body_cf[3];
  float chi_eps;
  // tiling of the 1-D grid onto (x-tiles, y-tiles, z), or, with `linear`,
  // block b covering cells [b T, (b + 1) T) of the flattened index (T threads
  // per block), so that no lane idles when nx is not a multiple of the block
  // width; x, y, z then come from two divisions by invariant integers
  int32_t ntx, nty;
  uint32_t nblocks;
  int32_t linear;
  InvariantDiv div_nx, div_ny;
  int32_t swizzle;     // 1: consecutive tiles on one XCD
  // Step counter of the dither of 16-bit stores: this launch reads *salt_in
  // and block 0 writes *salt_out = *salt_in + 1 for the next (the two words
  // alternate with the lattice parity, so graph replays keep counting).
  const uint32_t* salt_in;
  uint32_t* salt_out;
};

// The serve loop's batched step: per-solve state in device memory.
struct BatchSolve {
  StepArgs args[2];      // by lattice parity p: f[p] -> f[p ^ 1], force buffers of that parity
  void* f[2];            // the two lattices
  float* mac;            // written on the solve's last step of the chunk
  int32_t parity;        // lattice parity at launch 0 of the chunk
  int32_t remaining;     // steps left at launch 0 (the solve idles once launch >= remaining)
  int32_t pending;       // force partials of the step before launch 0 not yet summed
  uint32_t step0;        // steps done at launch 0 (seeds the 16-bit dither)
};
struct BatchTile {
  uint32_t solve;        // index into the solve table
  uint32_t tile;         // tile of that solve, numbered as k_step numbers its blocks
};
struct BatchArgs {
  const BatchSolve* solves;
  const BatchTile* tiles;
  int32_t launch;        // index of this launch within the chunk
};

// Sums the force partial slots in slot order and appends one record
// [fx, fy, fz, tx, ty, tz] to the history ring.
struct ForceFinalizeArgs {
  const double* partials;
  uint32_t nslots;
  double* history;
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
  float band;              // class bit 2 where phi < band (diffuse coupling's force cells)
  uint8_t* cls;
  uint32_t* links;         // filled by launch_links from cls
};
// Class byte bits: solid (phi < 0), near (phi < 2: a link can reach the
// solid), band (phi < ClassifyArgs::band).
constexpr uint8_t kClsSolid = 1, kClsNear = 2, kClsBand = 4;

// `occ` caps residency at that many waves per SIMD (0: no cap) by giving
// each block its share of the CU's LDS; the step kernels stream HBM, and the
// best number of loads in flight depends on the element width and the grid.
// `diffuse`: the pinned body's diffuse penalty coupling (with `general`).
hipError_t launch_step(int q, int storage, bool general, bool nontemporal, int occ, const void* fin, void* fout,
                       const StepArgs& a, int bx, int by, hipStream_t s, bool diffuse = false);
// `rows`: step_row_mode holds for every solve of the batch; `force`: some
// solve of the batch has a nonzero body force.
hipError_t launch_step_batch(int q, int storage, bool general, bool nontemporal, int occ, bool rows, bool force,
                             const BatchArgs& b, uint32_t ntiles, int bx, int by, hipStream_t s);
// Device word of the persistent chunk's grid barrier (zeroed on the stream
// before each launch): the arrivals, with bit 31 set after the launch when a
// barrier wait ran out (the launch then stopped early; its results are void).
struct PersistBarrier {
  uint32_t arrived;
  uint32_t pad;
};
// The same `steps` launches as one cooperative persistent launch (b.launch is
// ignored): the grid holds as many blocks as fit on the device at once and
// meets at a grid barrier between steps.  Not capturable in a graph.
hipError_t launch_step_persistent(int q, int storage, bool general, bool nontemporal, int occ, bool rows, bool force,
                                  const BatchArgs& b, uint32_t ntiles, int bx, int by, int steps, PersistBarrier* bar,
                                  hipStream_t s);
// Whether `occ` waves per SIMD can be set exactly for blocks of `threads`.
bool occupancy_cap_ok(int occ, int threads);
hipError_t launch_force_finalize(const ForceFinalizeArgs& a, hipStream_t s);
hipError_t launch_totals(int q, int storage, const void* f, const TotalsArgs& a, int nblocks, hipStream_t s);
// The wire format is plane-major f ([Q][ncells] float); the lattice uses the
// block layout below.
hipError_t launch_init_eq(int q, int storage, const float* mac, void* f, uint32_t ncells, hipStream_t s);
hipError_t launch_pack(int q, int storage, const float* wire, void* f, uint32_t ncells, hipStream_t s);
hipError_t launch_unpack(int q, int storage, const void* f, float* wire, uint32_t ncells, hipStream_t s);

// Row mode of the step kernels: every wave of a tile lies in one grid row and
// rows are whole layout blocks, so a source row's address is wave-uniform.
// Waves are 64 consecutive threads; a linear tile's waves start at multiples
// of 64 cells, a 2D tile's at multiples of 64 in x when bx is one.
inline bool step_row_mode(int nx, bool linear, int bx, int by) {
  if (nx % (int)kLatBlock != 0 || nx % 64 != 0 || (bx * by) % 64 != 0) return false;
  return linear || bx % 64 == 0;
}
hipError_t launch_classify(const ClassifyArgs& a, hipStream_t s);
// Link masks of the cells next to the body (StepArgs::links) from the class
// bytes of launch_classify; zero for every other cell.
hipError_t launch_links(int q, const ClassifyArgs& a, hipStream_t s);
// Streaming copy of `bytes` (multiple of 16) used to calibrate HBM bandwidth
// and the PMC byte counters.
hipError_t launch_copy_probe(const void* src, void* dst, size_t bytes, hipStream_t s);

}  // namespace lbm

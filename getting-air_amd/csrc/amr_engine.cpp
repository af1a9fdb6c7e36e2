// Host side of the block-structured AMR engine: the lbm_amr object behind the
// lbm_amr_* entry points of include/lbm.h.  It owns a dense L0 solve (its
// lattices, macroscopic field and stream) and a pool per finer level: two
// population buffers (A holds the state between macro-steps), the planar
// pool macroscopic field, the block maps and the surface work lists; level 1
// also carries the manager's lists.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/lbm.h"
#include "amr.h"
#include "engine.h"
#include "lattice_pairs.h"

using namespace lbm;

#define HIPCHK(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

// A pool level m >= 2, whose parent is the pool of level m - 1.  Level 1
// (the L0 parent) lives in lbm_amr's own fields.
struct DeepLevel {
  int m = 0;
  PoolGeom G{};
  int nblocks = 0;
  size_t pool_cells = 0;
  float* pool[2] = {nullptr, nullptr};
  float* mac_pool = nullptr;
  int32_t* block_slot = nullptr;
  int32_t* slot_to_block = nullptr;
  int32_t* active = nullptr;
  int32_t* pairs = nullptr;
  uint8_t* ccls = nullptr;
  int32_t* tile_table = nullptr;
  SurfacePlan plan{};
  uint32_t nseam = 0, nnear = 0, nring = 0;
  CollideParams fine{};
  float c2f = 0.f;  // non-equilibrium rescale from level m - 1
  float f2c = 0.f;  // and back to it (the restriction)
  int nactive = 0;
  std::vector<int32_t> h_block_slot, h_slot_to_block;  // host copies (topology checks)
  // the manager (manage_every > 0)
  uint8_t* want = nullptr;
  uint8_t* geom = nullptr;     // level_mask
  int32_t* dying = nullptr;
  int32_t* fresh = nullptr;
  int32_t* free_list = nullptr;
  int32_t* status = nullptr;
};

struct lbm_amr {
  lbm_amr_desc d{};
  int levels = 2;                // including L0
  std::vector<DeepLevel> deep;   // levels 2 .. levels - 1
  std::vector<int32_t> h_block_slot1;  // level 1's uploaded block map (host copy)
  lbm_solve* l0 = nullptr;
  hipStream_t stream = nullptr;
  PoolGeom G{};
  int fb = 0;
  int nblocks = 0;
  size_t pool_cells = 0;  // max_slots * fb^3: the compact plane of the host arrays
  float* pool[2] = {nullptr, nullptr};
  float* mac_pool = nullptr;
  int32_t* block_slot = nullptr;
  int32_t* slot_to_block = nullptr;
  uint8_t* want = nullptr;
  float* crit = nullptr;
  int32_t* dying = nullptr;
  int32_t* fresh = nullptr;
  int32_t* active = nullptr;
  int32_t* free_list = nullptr;
  int32_t* status = nullptr;
  int32_t* pairs = nullptr;
  uint8_t* ccls = nullptr;
  int32_t* tile_table = nullptr;  // kTileWords per active tile
  SurfacePlan plan{};
  uint32_t nseam = 0, nnear = 0, nring = 0, nflux = 0;
  float* flux_acc = nullptr;        // [dm, dpx, dpy, dpz] per flux cell (reflux only)
  double* partials = nullptr;
  unsigned* counter = nullptr;
  double* totals = nullptr;
  double* force_partials = nullptr;       // one FP64 [6] per pool force block
  double* history = nullptr;              // force records, a ring of kForceCap
  unsigned long long* written = nullptr;  // records written since the seed
  CollideParams fine{};
  float c2f = 0.f, f2c = 0.f;
  int nactive = 0;
  int exhausted = 0;
  long long step = 0;
  bool seeded = false;
  bool body_covered = false;  // the uploaded topology holds the body (check_topology)
  bool has_body = false;      // a pinned sphere, solved on the finest level
  bool managed = false;       // the manager has changed the block maps since the host copies
  uint8_t* geom = nullptr;    // per level-1 block (the manager): level_mask
};

namespace {

const uint32_t kForceCap = 1u << 16;

// The f32 weight of direction i of the pool's lattice.
float lattice_w(int q, int i) {
  return q == 9 ? Lattice<9>::w[i] : (q == 19 ? Lattice<19>::w[i] : Lattice<27>::w[i]);
}

template <int Q>
void fine_params(CollideParams& p, float omega, float omega_m, const float* F) {
  using L = Lattice<Q>;
  p.omega = omega;
  p.omega_m = omega_m;
  p.guo = 1.f - 0.5f * omega;
  for (int k = 0; k < 3; k++) p.F[k] = F[k];
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    const float eF = L::ex[i] * F[0] + L::ey[i] * F[1] + L::ez[i] * F[2];
    // doubled, as collide_pairs uses them
    p.c1[k] = 2.f * (9.f * (1.f - 0.5f * omega) * L::w[i] * eF);
    p.so[k] = 2.f * ((1.f - 0.5f * omega_m) * 3.f * L::w[i] * eF);
  }
  // D2Q9's f32 weights sum to 1 + this (the dense step's constant)
  double ws = 0.0;
  for (int i = 0; i < Q; i++) ws += (double)L::w[i];
  p.wsum_m1 = Q == 9 ? (float)(ws - 1.0) : 0.f;
}

int validate(const lbm_amr_desc* d) {
  if (!d) return fail(LBM_E_INVALID, "null descriptor");
  if (d->abi_version != LBM_ABI_VERSION) return fail(LBM_E_INVALID, "descriptor abi_version %d", d->abi_version);
  if (d->q != 9 && d->q != 19 && d->q != 27)
    return fail(LBM_E_INVALID, "AMR needs q = 9 (the 2D quadtree), 19 or 27 (got %d)", d->q);
  const bool two_d = d->q == 9;
  if (two_d && d->nz != 1) return fail(LBM_E_INVALID, "the D2Q9 quadtree needs nz = 1 (got %d)", d->nz);
  if (!supported_rb(d->rb)) return fail(LBM_E_INVALID, "rb %d: the pool kernels are built for rb = 4, 8, 16", d->rb);
  if (d->nx < d->rb || d->ny < d->rb || d->nx % d->rb || d->ny % d->rb ||
      (!two_d && (d->nz < d->rb || d->nz % d->rb)))
    return fail(LBM_E_INVALID, "grid %dx%dx%d is not a multiple of rb=%d", d->nx, d->ny, d->nz, d->rb);
  // the nearest-periodic-image tile lookup needs three blocks per axis
  // (localInTile, common_d3_pool.wgsl:81-93)
  if (d->nx / d->rb < 3 || d->ny / d->rb < 3 || (!two_d && d->nz / d->rb < 3))
    return fail(LBM_E_INVALID, "grid %dx%dx%d has fewer than 3 blocks of %d on an axis", d->nx, d->ny, d->nz, d->rb);
  if (d->max_slots < 1) return fail(LBM_E_INVALID, "max_slots %d", d->max_slots);
  // explode's ring plan packs L0 coordinates in 10 bits each
  if (d->nx > 1024 || d->ny > 1024 || d->nz > 1024)
    return fail(LBM_E_INVALID, "grid %dx%dx%d: the AMR L0 grid is at most 1024 per axis", d->nx, d->ny, d->nz);
  if (d->criterion != 0 && d->criterion != 1) return fail(LBM_E_INVALID, "criterion %d: 0 (|curl u|) or 1 (Q)", d->criterion);
  if (d->explode_dirs < 0 || d->explode_dirs > d->q)
    return fail(LBM_E_INVALID, "explode_dirs %d outside 0..q", d->explode_dirs);
  const double tile = std::pow(2.0 * d->rb + 2 * kGhost, two_d ? 2 : 3);
  if ((double)d->max_slots * tile * 4.0 >= 4294967296.0)
    return fail(LBM_E_INVALID, "pool plane of %d tiles exceeds 4 GiB", d->max_slots);
  if (!(d->omega > 0.f && d->omega < 2.f)) return fail(LBM_E_INVALID, "omega %g outside (0, 2)", d->omega);
  const double tau = 1.0 / d->omega;
  if (std::fabs(tau - 1.0) < 1e-3)
    return fail(LBM_E_REFUSED, "tau_c = %g: the fine/coarse non-equilibrium rescale is singular at tau = 1", tau);
  if (d->collision == LBM_TRT && !(d->omega_minus > 0.f && d->omega_minus < 2.f))
    return fail(LBM_E_INVALID, "omega_minus %g outside (0, 2)", d->omega_minus);
  if (d->manage_every < 0 || d->lead < 0) return fail(LBM_E_INVALID, "manage_every / lead must be >= 0");
  if (!(d->ladder_octaves <= 16.f)) return fail(LBM_E_INVALID, "ladder_octaves %g outside [0, 16]", d->ladder_octaves);
  // the flux correction measures a mismatch that explode/coalesce does not
  // create (main-3d.js:1333-1336)
  if (d->reflux && !d->interface_interp)
    return fail(LBM_E_INVALID, "reflux needs the interp interface: explode/coalesce is conservative by construction");
  const int levels = d->levels == 0 ? 2 : d->levels;
  if (levels < 2 || levels > 4) return fail(LBM_E_INVALID, "levels %d: 2, 3 or 4", d->levels);
  if (two_d) {
    // the D2Q9 quadtree: two levels on the explode/coalesce interface, no body
    if (levels != 2) return fail(LBM_E_REFUSED, "levels %d: the D2Q9 quadtree runs two levels", levels);
    if (d->interface_interp) return fail(LBM_E_REFUSED, "the D2Q9 quadtree runs the explode interface");
    if (d->body.shape != LBM_BODY_NONE) return fail(LBM_E_REFUSED, "the D2Q9 quadtree takes no body");
  }
  if (levels >= 3) {
    // the pool-parent coupling is explode/coalesce (main-3d.js:1283-1288)
    if (d->interface_interp) return fail(LBM_E_REFUSED, "levels %d needs the explode interface", levels);
    // explode's ring plan packs parent-grid coordinates in 10 bits each
    const int s = 1 << (levels - 2);
    if ((long)d->nx * s > 1024 || (long)d->ny * s > 1024 || (long)d->nz * s > 1024)
      return fail(LBM_E_INVALID, "levels %d: level %d's parent grid exceeds 1024 cells per axis", levels, levels - 1);
    // every level but the finest is a parent whose rescale divides by tau - 1
    double tau = 1.0 / d->omega;
    for (int m = 1; m < levels - 1; m++) {
      tau = 2.0 * tau - 0.5;
      if (std::fabs(tau - 1.0) < 1e-3)
        return fail(LBM_E_REFUSED, "tau at level %d = %g: the non-equilibrium rescale is singular at tau = 1", m, tau);
    }
  }
  if (d->body.shape != LBM_BODY_NONE) {
    if (d->body.shape != LBM_BODY_SPHERE) return fail(LBM_E_INVALID, "AMR body shape %d: only a sphere", d->body.shape);
    if (!(d->body.radius > 0.f)) return fail(LBM_E_INVALID, "AMR body radius %g", d->body.radius);
    if (d->interface_interp) return fail(LBM_E_INVALID, "an AMR body needs the explode interface");
  }
  return LBM_OK;
}

// The manager's geometry term for the pinned sphere (blockWanted,
// common_d3_manage.wgsl:257-302; a pinned body has no look-ahead): a block
// whose box of L0 cells comes within kBodyMargin of the surface is always
// refined.  The reference's default margin (main-3d.js:485), which covers the
// fine step's bounce-back and force reach of one L0 cell.
constexpr double kBodyMargin = 2.0;
// sqrt(3) rounded up, the literal of d3-amr.mjs:120 and
// common_d3_manage.wgsl:178 (a box's circumradius bound must not round down)
constexpr double kSqrt3Up = 1.7320509;

// Whether any point of the cube centred at c with half-edge h comes within
// `margin` of the sphere: the SDF is 1-Lipschitz, so phi(c) - h sqrt(3)
// bounds the cube's minimum from below; an ambiguous cube splits into octants
// down to `depth`, where it is accepted (nearBodyWant, d3-amr.mjs:139-159).
// Evaluated in FP64 without contraction, as the host twin evaluates it.
#pragma clang fp contract(off)
bool near_sphere(const double c[3], double h, int depth, const double ctr[3], double r, double margin) {
  const double dx = c[0] - ctr[0], dy = c[1] - ctr[1], dz = c[2] - ctr[2];
  const double phi = std::sqrt(dx * dx + dy * dy + dz * dz) - r;
  if (phi - h * kSqrt3Up > margin) return false;
  if (phi <= margin || depth == 0) return true;
  const double q = h / 2;
  for (int k = 0; k < 8; k++) {
    const double cc[3] = {c[0] + ((k & 1) ? q : -q), c[1] + ((k & 2) ? q : -q), c[2] + ((k & 4) ? q : -q)};
    if (near_sphere(cc, q, depth - 1, ctr, r, margin)) return true;
  }
  return false;
}

// Per block of level m (rb cells of level m - 1, an L0 box of edge
// rb 2^-(m-1)): bit 0 (the finest level, with a body) the body's geometry
// wants it refined, whatever the field says; bit 1 the field criterion may
// not refine it, because its L0 box comes within the sponge band (the fine
// step does not relax toward the far field) dilated by how far the 2:1
// closure can carry a level-m want toward level 1: octet completion and the
// 27-neighbour parent gather spread a want by at most three of its own
// blocks per level, 3 rb (1 - 2^-(m-1)) L0 cells down to level 1 in all.
// The box is scaled as refineHierarchy scales it (a pure 2^-(m-1), the
// BOX_SCALE of common_d3_manage.wgsl:150-163), so the geometry set equals
// the host hierarchy's.
std::vector<uint8_t> level_mask(const lbm_amr_desc& d, int m, bool body) {
  const int s = 1 << (m - 1), rb = d.rb;
  const int nbx = d.nx * s / rb, nby = d.ny * s / rb, nbz = d.q == 9 ? 1 : d.nz * s / rb;
  const int n[3] = {d.nx, d.ny, d.nz};
  const double sc = std::ldexp(1.0, -(m - 1)), edge = rb * sc, dil = 3.0 * rb * (1.0 - sc);
  const double ctr[3] = {d.body.centre[0], d.body.centre[1], d.body.centre[2]};
  std::vector<uint8_t> mask((size_t)nbx * nby * nbz, 0);
  for (size_t b = 0; b < mask.size(); b++) {
    const int bc[3] = {(int)(b % nbx), (int)((b / nbx) % nby), (int)(b / ((size_t)nbx * nby))};
    if (body) {
      const double c[3] = {(bc[0] * rb + rb / 2.0) * sc, (bc[1] * rb + rb / 2.0) * sc, (bc[2] * rb + rb / 2.0) * sc};
      if (near_sphere(c, edge / 2.0, 3, ctr, d.body.radius, kBodyMargin)) mask[b] |= 1;
    }
    if (d.sponge_width > 0.f) {
      for (int k = 0; k < (d.q == 9 ? 2 : 3); k++) {
        const double lo = bc[k] * edge;
        if (d.sponge_axes[k] && (lo < d.sponge_width + dil || lo + edge > n[k] - d.sponge_width - dil)) mask[b] |= 2;
      }
    }
  }
  return mask;
}
#pragma clang fp contract(on)

// Refuses a topology that leaves an L0 cell within kBodyReach (1 cell) of the
// body unrefined (the body lives on the fine level alone, main-3d.js:3127-3136)
// or refines a block that reaches into the sponge band (the fine step does
// not relax toward the far field).
int check_topology(const lbm_amr_desc& d, const int32_t* block_slot) {
  const int rb = d.rb, nbx = d.nx / rb, nby = d.ny / rb, nbz = d.q == 9 ? 1 : d.nz / rb;
  const int n[3] = {d.nx, d.ny, d.nz};
  for (int b = 0; b < nbx * nby * nbz; b++) {
    const int bc[3] = {b % nbx, (b / nbx) % nby, b / (nbx * nby)};
    if (block_slot[b] < 0 && d.body.shape == LBM_BODY_SPHERE) {
      // distance from the centre to the block's box of cell centres
      double s2 = 0.0;
      for (int k = 0; k < 3; k++) {
        const double lo = bc[k] * rb, hi = lo + rb - 1, c = d.body.centre[k];
        const double e = c < lo ? lo - c : (c > hi ? c - hi : 0.0);
        s2 += e * e;
      }
      if (std::sqrt(s2) - d.body.radius < 1.0)
        return fail(LBM_E_INVALID, "block %d lies within one cell of the body and is not refined", b);
    }
    if (block_slot[b] >= 0 && d.sponge_width > 0.f) {
      for (int k = 0; k < (d.q == 9 ? 2 : 3); k++) {
        const int lo = bc[k] * rb, hi = lo + rb - 1;
        if (d.sponge_axes[k] && (lo < d.sponge_width || (n[k] - 1) - hi < d.sponge_width))
          return fail(LBM_E_INVALID, "refined block %d reaches into the sponge band", b);
      }
    }
  }
  return LBM_OK;
}

void release(lbm_amr* a) {
  if (!a) return;
  void* ptrs[] = {a->pool[0], a->pool[1], a->mac_pool, a->block_slot, a->slot_to_block, a->want, a->crit,
                  a->dying, a->fresh, a->active, a->free_list, a->status, a->pairs, a->ccls, a->partials, a->counter, a->totals,
                  a->plan.seam, a->plan.near, a->plan.ring, a->plan.flux, a->plan.count, a->flux_acc,
                  a->force_partials, a->history, a->written, a->geom, a->tile_table};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (DeepLevel& L : a->deep) {
    void* dp[] = {L.pool[0], L.pool[1], L.mac_pool, L.block_slot, L.slot_to_block, L.active, L.pairs, L.ccls,
                  L.tile_table, L.plan.seam, L.plan.near, L.plan.ring, L.plan.count, L.want, L.geom,
                  L.dying, L.fresh, L.free_list, L.status};
    for (void* p : dp)
      if (p) (void)hipFree(p);
  }
  if (a->l0) lbm_solve_destroy(a->l0);
  delete a;
}

template <class T>
int dalloc(T** p, size_t bytes) {
  hipError_t e = hipMalloc((void**)p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(LBM_E_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  return LBM_OK;
}

// L0 as the parent of level 1: population buffer g, its moments.
ParentRef dense_parent(lbm_amr* a, float* g) {
  ParentRef P{};
  P.g = g;
  P.mac = solve_mac(a->l0);
  return P;
}

// Pool level m >= 1 as the parent of level m + 1, its population buffer p.
ParentRef pool_parent(lbm_amr* a, int m, int p) {
  ParentRef P{};
  const PoolGeom& G = m == 1 ? a->G : a->deep[m - 2].G;
  P.g = m == 1 ? a->pool[p] : a->deep[m - 2].pool[p];
  P.mac = m == 1 ? a->mac_pool : a->deep[m - 2].mac_pool;
  P.block_slot = G.block_slot;
  P.nbx = G.nbx;
  P.nby = G.nby;
  P.nbz = G.nbz;
  P.plane = G.pool_plane;
  return P;
}

// Fills the tiles listed in `pairs` ([count, (block, slot)...], device) from
// the current L0 state.
int fill_tiles(lbm_amr* a, const int32_t* pairs, int n) {
  HIPCHK(launch_interp(a->d.q, dense_parent(a, solve_lattice(a->l0, 0)), a->pool[0], a->G, pairs, n, a->c2f, 0,
                       a->stream));
  return LBM_OK;
}

// Rebuilds the L0 cell classes and the surface passes' work lists for the
// current block maps and active list (a->nactive must be current).
int rebuild_plan(lbm_amr* a) {
  HIPCHK(hipMemsetAsync(a->plan.count, 0, 16, a->stream));
  HIPCHK(launch_cell_class(a->d.q, a->ccls, a->G, a->plan, a->active, a->nactive, dense_parent(a, nullptr),
                           a->stream));
  uint32_t c[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(c, a->plan.count, 16, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  if (c[0] > a->plan.cap_seam || c[1] > a->plan.cap_near || c[3] > a->plan.cap_flux)
    return fail(LBM_E_DEVICE, "surface work lists overflow (%u/%u seam, %u/%u near, %u/%u flux)", c[0],
                a->plan.cap_seam, c[1], a->plan.cap_near, c[3], a->plan.cap_flux);
  a->nseam = c[0];
  a->nnear = c[1];
  a->nring = (uint32_t)a->nactive * ring_per_tile(a->d.rb, a->d.q == 9 ? 2 : 3);
  a->nflux = c[3];
  return LBM_OK;
}

int refresh_active(lbm_amr* a) {
  HIPCHK(launch_active_list(a->slot_to_block, a->d.max_slots, a->active, a->stream));
  int32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, a->active, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  a->nactive = n;
  int rc = rebuild_plan(a);
  if (rc) return rc;
  return LBM_OK;
}

// Rebuilds pool level L's cell classes over its parent grid and its surface
// work lists, for its current active list.
int rebuild_level_plan(lbm_amr* a, DeepLevel& L) {
  HIPCHK(hipMemsetAsync(L.plan.count, 0, 16, a->stream));
  HIPCHK(launch_cell_class(a->d.q, L.ccls, L.G, L.plan, L.active, L.nactive, pool_parent(a, L.m - 1, 0),
                           a->stream));
  uint32_t c[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(c, L.plan.count, 16, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  if (c[0] > L.plan.cap_seam || c[1] > L.plan.cap_near)
    return fail(LBM_E_DEVICE, "level %d surface work lists overflow (%u/%u seam, %u/%u near)", L.m, c[0],
                L.plan.cap_seam, c[1], L.plan.cap_near);
  L.nseam = c[0];
  L.nnear = c[1];
  L.nring = (uint32_t)L.nactive * ring_per_tile(a->d.rb, 3);
  return LBM_OK;
}

int refresh_level(lbm_amr* a, DeepLevel& L) {
  HIPCHK(launch_active_list(L.slot_to_block, a->d.max_slots, L.active, a->stream));
  int32_t n = 0;
  HIPCHK(hipMemcpyAsync(&n, L.active, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  L.nactive = n;
  return rebuild_level_plan(a, L);
}

// Fills every tile of pool level L from its parent level's state (buffer 0,
// which holds it between macro-steps), ring included.
int fill_level(lbm_amr* a, DeepLevel& L) {
  if (L.nactive <= 0) return LBM_OK;
  HIPCHK(launch_active_pairs(L.active, L.slot_to_block, a->d.max_slots, L.pairs, a->stream));
  HIPCHK(launch_interp(a->d.q, pool_parent(a, L.m - 1, 0), L.pool[0], L.G, L.pairs, L.nactive, L.c2f, 0, a->stream));
  return LBM_OK;
}

// Host checks of a level-m (m >= 2) block map against its parent's: every
// tile nests in a parent tile, and every ring cell's parent cell that no
// level-m tile covers lies in a parent tile (checkRingParentCoverage,
// d3-amr.mjs:348-390), so explode and the seam passes always find parent
// data.  On the finest level with a body, every block within one L0 cell of
// it must be refined (the coarser levels never see the body).
int check_level_topology(lbm_amr* a, int m, const int32_t* bs) {
  const lbm_amr_desc& d = a->d;
  const int rb = d.rb, rb2 = 2 * rb, fb = rb2 + 2 * kGhost;
  const DeepLevel& L = a->deep[m - 2];
  const std::vector<int32_t>& pbs = m == 2 ? a->h_block_slot1 : a->deep[m - 3].h_block_slot;
  if (pbs.empty()) return fail(LBM_E_STATE, "level %d uploaded before level %d", m, m - 1);
  const int nbx = L.G.nbx, nby = L.G.nby, nbz = L.G.nbz;
  const int px = L.G.nx, py = L.G.ny, pz = L.G.nz;         // parent grid
  const int pnbx = px / rb2, pnby = py / rb2;              // parent blocks per axis
  auto parent_has = [&](int vx, int vy, int vz) { return pbs[((vz / rb2) * pnby + vy / rb2) * pnbx + vx / rb2] >= 0; };
  auto covered = [&](int vx, int vy, int vz) { return bs[((vz / rb) * nby + vy / rb) * nbx + vx / rb] >= 0; };
  for (int b = 0; b < nbx * nby * nbz; b++) {
    if (bs[b] < 0) continue;
    const int bx = b % nbx, by = (b / nbx) % nby, bz = b / (nbx * nby);
    if (!parent_has(bx * rb, by * rb, bz * rb))
      return fail(LBM_E_INVALID, "level %d block %d does not nest in a level %d tile", m, b, m - 1);
    for (int f = 0; f < fb * fb * fb; f++) {
      const int fx = f % fb, fy = (f / fb) % fb, fz = f / (fb * fb);
      if (fx >= kGhost && fx < kGhost + rb2 && fy >= kGhost && fy < kGhost + rb2 && fz >= kGhost && fz < kGhost + rb2)
        continue;
      const int gx = ((bx * rb2 + fx - kGhost) % (2 * px) + 2 * px) % (2 * px);
      const int gy = ((by * rb2 + fy - kGhost) % (2 * py) + 2 * py) % (2 * py);
      const int gz = ((bz * rb2 + fz - kGhost) % (2 * pz) + 2 * pz) % (2 * pz);
      const int vx = gx / 2, vy = gy / 2, vz = gz / 2;
      if (!covered(vx, vy, vz) && !parent_has(vx, vy, vz))
        return fail(LBM_E_INVALID, "level %d block %d: a ring cell's parent cell lies in no level %d tile", m, b, m - 1);
    }
  }
  if (a->has_body && m == a->levels - 1) {
    // the block's box of parent-cell centres in L0 units, +- half a cell
    const double sc = std::ldexp(1.0, -(m - 1)), off = -0.5 * (1.0 - sc);
    for (int b = 0; b < nbx * nby * nbz; b++) {
      if (bs[b] >= 0) continue;
      const int bc[3] = {b % nbx, (b / nbx) % nby, b / (nbx * nby)};
      double s2 = 0.0;
      for (int k = 0; k < 3; k++) {
        const double lo = (bc[k] * rb) * sc + off - 0.5 * sc, hi = (bc[k] * rb + rb - 1) * sc + off + 0.5 * sc;
        const double c = d.body.centre[k];
        const double e = c < lo ? lo - c : (c > hi ? c - hi : 0.0);
        s2 += e * e;
      }
      if (std::sqrt(s2) - d.body.radius < 1.0)
        return fail(LBM_E_INVALID, "level %d block %d lies within one L0 cell of the body and is not refined", m, b);
    }
  }
  return LBM_OK;
}

// Host copies of every level's block map, from the device (after the
// manager has moved tiles; the level uploads check against them).
int sync_host_maps(lbm_amr* a) {
  a->h_block_slot1.resize(a->nblocks);
  HIPCHK(hipMemcpyAsync(a->h_block_slot1.data(), a->block_slot, (size_t)a->nblocks * 4, hipMemcpyDeviceToHost,
                        a->stream));
  for (DeepLevel& L : a->deep) {
    L.h_block_slot.resize(L.nblocks);
    L.h_slot_to_block.resize(a->d.max_slots);
    HIPCHK(hipMemcpyAsync(L.h_block_slot.data(), L.block_slot, (size_t)L.nblocks * 4, hipMemcpyDeviceToHost,
                          a->stream));
    HIPCHK(hipMemcpyAsync(L.h_slot_to_block.data(), L.slot_to_block, (size_t)a->d.max_slots * 4,
                          hipMemcpyDeviceToHost, a->stream));
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  a->managed = false;
  return LBM_OK;
}

// Restricts every tile in use onto the L0 cells under it (the average/drain
// pass).  Between manager passes the covered L0 step leaves those cells
// untouched, so this is what makes them a valid interpolation source for a
// neighbouring fresh tile, and it drains the tiles about to die.
int restrict_active(lbm_amr* a) {
  if (a->nactive <= 0) return LBM_OK;
  HIPCHK(launch_active_pairs(a->active, a->slot_to_block, a->d.max_slots, a->pairs, a->stream));
  HIPCHK(launch_average(a->d.q, a->pool[0], dense_parent(a, solve_lattice(a->l0, 0)), a->G, a->pairs, a->nactive,
                        a->f2c, a->stream));
  return LBM_OK;
}

// The field ladder's rung at level m: refine and keep thresholds scaled by
// r^(m-1), r = 2^octaves for a rotation rate, 4^octaves for Q.
void ladder_rung(const lbm_amr_desc& d, int m, float* refine_at, float* keep_at) {
  const double oct = d.ladder_octaves > 0.f ? d.ladder_octaves : 1.0;
  const double f = std::exp2(oct * (d.criterion == 1 ? 2.0 : 1.0) * (m - 1));
  *refine_at = (float)(d.refine_at * f);
  *keep_at = (float)(d.keep_at * f);
}

// The manager over a hierarchy of three or four levels (main-3d.js:3251-3321):
// the criterion on L0's field, every tile restricted into its parent
// finest-first (the drain; it also makes covered parent cells a valid fill
// source), every level's want (the body at the finest level, the field on
// its rung), the 2:1 closure finest-first, each level's slots (the ordered
// allocator of the two-level manager), the plans rebuilt, and the new tiles
// filled from their parents coarsest-first, so that a tile whose parent is
// also new reads the parent's fill.
int manage_deep(lbm_amr* a) {
  const lbm_amr_desc& d = a->d;
  const int D = a->levels - 1;
  HIPCHK(launch_criterion(solve_mac(a->l0), a->crit, a->G, (float)d.lead, d.criterion, a->stream));
  for (int m = D; m >= 2; m--) {
    DeepLevel& L = a->deep[m - 2];
    if (L.nactive <= 0) continue;
    HIPCHK(launch_active_pairs(L.active, L.slot_to_block, d.max_slots, L.pairs, a->stream));
    HIPCHK(launch_average(d.q, L.pool[0], pool_parent(a, m - 1, 0), L.G, L.pairs, L.nactive, L.f2c, a->stream));
  }
  int rc = restrict_active(a);
  if (rc) return rc;
  for (int m = 1; m <= D; m++) {
    DecideArgs x{};
    const PoolGeom& G = m == 1 ? a->G : a->deep[m - 2].G;
    x.nbx = G.nbx;
    x.nby = G.nby;
    x.nbz = G.nbz;
    x.shift = m - 1;
    x.cnbx = a->G.nbx;
    x.cnby = a->G.nby;
    x.crit = a->crit;
    ladder_rung(d, m, &x.refine_at, &x.keep_at);
    x.geom = m == 1 ? a->geom : a->deep[m - 2].geom;
    x.block_slot = G.block_slot;
    x.want = m == 1 ? a->want : a->deep[m - 2].want;
    HIPCHK(launch_decide(x, a->stream));
  }
  for (int m = D; m >= 2; m--) {
    DeepLevel& L = a->deep[m - 2];
    HIPCHK(launch_complete_octets(L.want, L.G.nbx, L.G.nby, L.G.nbz, a->stream));
    const PoolGeom& P = m == 2 ? a->G : a->deep[m - 3].G;
    HIPCHK(launch_balance(m == 2 ? a->want : a->deep[m - 3].want, P.nbx, P.nby, P.nbz, L.want, a->stream));
  }
  for (int m = 1; m <= D; m++) {
    ManageArgs x{};
    const bool top = m == 1;
    DeepLevel* L = top ? nullptr : &a->deep[m - 2];
    x.nblocks = top ? a->nblocks : L->nblocks;
    x.max_slots = d.max_slots;
    x.geom = top ? a->geom : nullptr;  // level 1: count wanted blocks in the sponge band
    x.block_slot = top ? a->block_slot : L->block_slot;
    x.slot_to_block = top ? a->slot_to_block : L->slot_to_block;
    x.want = top ? a->want : L->want;
    x.dying = top ? a->dying : L->dying;
    x.fresh = top ? a->fresh : L->fresh;
    x.active = top ? a->active : L->active;
    x.free_list = top ? a->free_list : L->free_list;
    x.status = top ? a->status : L->status;
    x.decided = 1;
    HIPCHK(launch_manage(x, a->stream));
  }
  std::vector<int32_t> counts(4 * D, 0);  // per level: fresh, active, status[0], status[1]
  for (int m = 1; m <= D; m++) {
    int32_t* c = &counts[4 * (m - 1)];
    const bool top = m == 1;
    HIPCHK(hipMemcpyAsync(&c[0], top ? a->fresh : a->deep[m - 2].fresh, 4, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipMemcpyAsync(&c[1], top ? a->active : a->deep[m - 2].active, 4, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipMemcpyAsync(&c[2], top ? a->status : a->deep[m - 2].status, 8, hipMemcpyDeviceToHost, a->stream));
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  a->nactive = counts[1];
  for (int m = 2; m <= D; m++) a->deep[m - 2].nactive = counts[4 * (m - 1) + 1];
  if ((rc = rebuild_plan(a))) return rc;
  for (int m = 2; m <= D; m++)
    if ((rc = rebuild_level_plan(a, a->deep[m - 2]))) return rc;
  if ((rc = fill_tiles(a, a->fresh, counts[0]))) return rc;
  for (int m = 2; m <= D; m++) {
    DeepLevel& L = a->deep[m - 2];
    const int n = counts[4 * (m - 1)];
    if (n > 0)
      HIPCHK(launch_interp(d.q, pool_parent(a, m - 1, 0), L.pool[0], L.G, L.fresh, n, L.c2f, 0, a->stream));
  }
  int exhausted = 0;
  for (int m = 1; m <= D; m++) exhausted += counts[4 * (m - 1) + 2];
  a->exhausted = exhausted;
  if (counts[3] > 0)
    return fail(LBM_E_INVALID, "the 2:1 closure carried %d level-1 blocks into the sponge band", counts[3]);
  if (exhausted > 0)
    return fail(LBM_E_NOMEM, "AMR pool exhausted: %d wanted tiles found no free slot (max_slots %d per level)",
                exhausted, d.max_slots);
  if (a->has_body) a->body_covered = true;
  return LBM_OK;
}

int manage(lbm_amr* a) {
  const lbm_amr_desc& d = a->d;
  a->managed = true;
  if (a->levels > 2) return manage_deep(a);
  HIPCHK(launch_criterion(solve_mac(a->l0), a->crit, a->G, (float)d.lead, d.criterion, a->stream));
  int rc = restrict_active(a);
  if (rc) return rc;
  ManageArgs m{};
  m.nblocks = a->nblocks;
  m.max_slots = d.max_slots;
  m.crit = a->crit;
  m.refine_at = d.refine_at;
  m.keep_at = d.keep_at;
  m.geom = a->geom;
  m.block_slot = a->block_slot;
  m.slot_to_block = a->slot_to_block;
  m.want = a->want;
  m.dying = a->dying;
  m.fresh = a->fresh;
  m.active = a->active;
  m.free_list = a->free_list;
  m.status = a->status;
  HIPCHK(launch_manage(m, a->stream));
  int32_t counts[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(&counts[0], a->dying, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipMemcpyAsync(&counts[1], a->fresh, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipMemcpyAsync(&counts[2], a->active, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipMemcpyAsync(&counts[3], a->status, 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  a->nactive = counts[2];
  rc = rebuild_plan(a);
  if (rc) return rc;
  // dying tiles were drained by restrict_active; fill the new ones
  rc = fill_tiles(a, a->fresh, counts[1]);
  if (rc) return rc;
  a->exhausted = counts[3];
  if (counts[3] > 0)
    return fail(LBM_E_NOMEM, "AMR pool exhausted: %d wanted tiles found no free slot (max_slots %d)", counts[3],
                d.max_slots);
  // every block the geometry wants is placed, and they cover each L0 cell
  // within kBodyReach of the body (the margin exceeds that reach)
  if (a->geom) a->body_covered = true;
  return LBM_OK;
}

// The interp interface (main-3d.js:3344-3357): the coarse step first (cells
// under a tile skipped), each tile's ring re-interpolated from the coarse
// state at t, two substeps in which the ring collides as well, every tile
// restricted onto the coarse state at t + 1, and with reflux the fine seam
// flux measured before each substep and the difference to the coarse one
// corrected onto the unrefined side.
int macro_step_interp(lbm_amr* a) {
  const lbm_amr_desc& d = a->d;
  int rc = solve_enqueue_covered_step(a->l0, a->block_slot, d.rb, a->G.nbx, a->G.nby);
  if (rc) return rc;
  const float* ct = solve_lattice(a->l0, 1);  // the state the coarse step read (t)
  float* ct1 = solve_lattice(a->l0, 0);       // the one it wrote (t + 1)
  if (a->nactive > 0) {
    HIPCHK(launch_active_pairs(a->active, a->slot_to_block, d.max_slots, a->pairs, a->stream));
    HIPCHK(launch_interp(d.q, dense_parent(a, const_cast<float*>(ct)), a->pool[0], a->G, a->pairs, a->nactive, a->c2f,
                         1, a->stream));
    if (d.reflux) HIPCHK(launch_flux(d.q, a->pool[0], a->flux_acc, a->G, a->plan, a->nflux, 0, a->stream));
    HIPCHK(launch_pool_step(d.q, a->pool[0], a->pool[1], nullptr, a->G, a->fine, a->nactive, 1, -1,
                            a->stream));
    if (d.reflux) HIPCHK(launch_flux(d.q, a->pool[1], a->flux_acc, a->G, a->plan, a->nflux, 1, a->stream));
    HIPCHK(launch_pool_step(d.q, a->pool[1], a->pool[0], a->mac_pool, a->G, a->fine, a->nactive, 1, -1,
                            a->stream));
    HIPCHK(launch_average(d.q, a->pool[0], dense_parent(a, ct1), a->G, a->pairs, a->nactive, a->f2c, a->stream));
    if (d.reflux)
      HIPCHK(launch_reflux(d.q, ct, ct1, solve_mac(a->l0), a->G, a->plan, a->nflux, a->flux_acc, d.force,
                           a->stream));
  }
  return LBM_OK;
}

// One substep of pool level m, reading its buffer p (advanceLevel,
// main-3d.js:3169-3180): with a finer level, that level's ring is exploded
// from this state, the finer level takes its two substeps and is coalesced
// back into this state; then this level steps.  Only the second substep's
// moments are written (the parent's coalesce and the readbacks take them).
int advance(lbm_amr* a, int m, int p) {
  const lbm_amr_desc& d = a->d;
  if (m + 1 < a->levels) {
    DeepLevel& c = a->deep[m - 1];
    if (c.nactive > 0) {
      const ParentRef P = pool_parent(a, m, p);
      HIPCHK(launch_explode(d.q, P, c.pool[0], c.G, c.plan, c.nring, d.explode_linear, d.explode_dirs, a->stream));
      int rc = advance(a, m + 1, 0);
      if (rc) return rc;
      rc = advance(a, m + 1, 1);
      if (rc) return rc;
      HIPCHK(launch_coalesce(d.q, c.pool[0], P, c.mac_pool, c.G, c.active, c.nactive, c.plan, c.nseam, c.nnear,
                             d.orphans, a->stream));
    }
  }
  if (m == 1) {
    HIPCHK(launch_pool_step(d.q, a->pool[p], a->pool[p ^ 1], p ? a->mac_pool : nullptr, a->G, a->fine, a->nactive, 0,
                            p, a->stream));
  } else {
    DeepLevel& L = a->deep[m - 2];
    HIPCHK(launch_pool_step(d.q, L.pool[p], L.pool[p ^ 1], p ? L.mac_pool : nullptr, L.G, L.fine, L.nactive, 0, p,
                            a->stream));
  }
  return LBM_OK;
}

int macro_step(lbm_amr* a) {
  const lbm_amr_desc& d = a->d;
  if (d.manage_every > 0 && a->step % d.manage_every == 0) {
    int rc = manage(a);
    if (rc) return rc;
  }
  if (d.interface_interp) {
    int rc = macro_step_interp(a);
    if (rc) return rc;
    a->step++;
    return LBM_OK;
  }
  float* l0 = solve_lattice(a->l0, 0);
  if (a->has_body) {
    // the one force pass of the macro-step, on the finest level's state at t
    // (main-3d.js:3127-3141, 1596-1597)
    const bool deep = a->levels > 2;
    const PoolGeom& G = deep ? a->deep.back().G : a->G;
    const int n = deep ? a->deep.back().nactive : a->nactive;
    HIPCHK(launch_pool_force(d.q, deep ? a->deep.back().pool[0] : a->pool[0], G,
                             deep ? a->deep.back().active : a->active, n, a->force_partials, a->stream));
    ForceFinalizeArgs fa{a->force_partials, pool_force_blocks(d.rb, n), a->history, a->written, kForceCap};
    HIPCHK(launch_force_finalize(fa, a->stream));
  }
  const ParentRef P0 = dense_parent(a, l0);
  HIPCHK(launch_explode(d.q, P0, a->pool[0], a->G, a->plan, a->nring, d.explode_linear, d.explode_dirs, a->stream));
  int rc0 = advance(a, 1, 0);
  if (rc0) return rc0;
  rc0 = advance(a, 1, 1);
  if (rc0) return rc0;
  HIPCHK(launch_coalesce(d.q, a->pool[0], P0, a->mac_pool, a->G, a->active, a->nactive, a->plan, a->nseam, a->nnear,
                         d.orphans, a->stream));
  int rc = solve_enqueue_covered_step(a->l0, a->block_slot, d.rb, a->G.nbx, a->G.nby);
  if (rc) return rc;
  a->step++;
  return LBM_OK;
}

// Pool levels 2 .. levels - 1: level m refines level m - 1's global grid
// (L0's times 2^(m-1)) in blocks of rb, with tau_m = 2 tau_(m-1) - 1/2 and
// the same tile shape and slot count as level 1.  Each level's step skips
// the interior cells its child covers.
int create_deep_levels(lbm_amr* a, double tau1) {
  const lbm_amr_desc& d = a->d;
  double tau_p = tau1;
  for (int m = 2; m < a->levels; m++) {
    a->deep.emplace_back();
    DeepLevel& L = a->deep.back();
    L.m = m;
    const int s = 1 << (m - 1);
    PoolGeom& G = L.G;
    G.nx = d.nx * s; G.ny = d.ny * s; G.nz = d.nz * s;
    G.rb = d.rb; G.fb = a->fb;
    G.nbx = G.nx / d.rb; G.nby = G.ny / d.rb; G.nbz = G.nz / d.rb;
    G.ncells = (uint32_t)((size_t)G.nx * G.ny * G.nz);
    G.max_slots = d.max_slots;
    G.pool_plane = a->G.pool_plane;
    G.l0_scale = (float)std::ldexp(1.0, -(m - 1));
    G.l0_offset = (float)(-0.5 * (1.0 - std::ldexp(1.0, -(m - 1))));
    G.dx2 = (float)std::ldexp(1.0, -2 * m);
    L.nblocks = G.nbx * G.nby * G.nbz;
    L.pool_cells = a->pool_cells;
    plan_capacity(d.rb, d.max_slots, G.ncells, 3, &L.plan.cap_seam, &L.plan.cap_near, &L.plan.cap_ring);
    const size_t pool_bytes = (size_t)G.pool_plane * d.q * 4;
    int rc;
    if ((rc = dalloc(&L.pool[0], pool_bytes)) || (rc = dalloc(&L.pool[1], pool_bytes)) ||
        (rc = dalloc(&L.mac_pool, (size_t)G.pool_plane * 16)) || (rc = dalloc(&L.block_slot, (size_t)L.nblocks * 4)) ||
        (rc = dalloc(&L.slot_to_block, (size_t)d.max_slots * 4)) ||
        (rc = dalloc(&L.active, (size_t)(d.max_slots + 1) * 4)) ||
        (rc = dalloc(&L.pairs, (size_t)(2 * d.max_slots + 1) * 4)) || (rc = dalloc(&L.ccls, (size_t)G.ncells)) ||
        (rc = dalloc(&L.tile_table, (size_t)d.max_slots * kTileWords * 4)) || (rc = dalloc(&L.plan.count, 16)) ||
        (rc = dalloc(&L.plan.seam, (size_t)L.plan.cap_seam * 4)) ||
        (rc = dalloc(&L.plan.near, (size_t)L.plan.cap_near * 4)) ||
        (rc = dalloc(&L.plan.ring, (size_t)L.plan.cap_ring * 16)))
      return rc;
    G.block_slot = L.block_slot;
    G.slot_to_block = L.slot_to_block;
    G.ccls = L.ccls;
    G.tile_table = L.tile_table;
    if (d.manage_every > 0) {
      const std::vector<uint8_t> mask = level_mask(d, m, a->has_body && m == a->levels - 1);
      if ((rc = dalloc(&L.want, (size_t)L.nblocks)) || (rc = dalloc(&L.geom, (size_t)L.nblocks)) ||
          (rc = dalloc(&L.dying, (size_t)(2 * d.max_slots + 1) * 4)) ||
          (rc = dalloc(&L.fresh, (size_t)(2 * d.max_slots + 1) * 4)) ||
          (rc = dalloc(&L.free_list, (size_t)d.max_slots * 4)) || (rc = dalloc(&L.status, 16)))
        return rc;
      HIPCHK(hipMemcpy(L.geom, mask.data(), mask.size(), hipMemcpyHostToDevice));
      HIPCHK(hipMemsetAsync(L.want, 0, (size_t)L.nblocks, a->stream));
      HIPCHK(hipMemsetAsync(L.status, 0, 16, a->stream));
    }
    HIPCHK(hipMemsetAsync(L.block_slot, 0xff, (size_t)L.nblocks * 4, a->stream));
    HIPCHK(hipMemsetAsync(L.slot_to_block, 0xff, (size_t)d.max_slots * 4, a->stream));
    HIPCHK(hipMemsetAsync(L.active, 0, 4, a->stream));
    HIPCHK(hipMemsetAsync(L.mac_pool, 0, (size_t)G.pool_plane * 16, a->stream));
    HIPCHK(hipMemsetAsync(L.ccls, 0, (size_t)G.ncells, a->stream));
    const double tau = 2.0 * tau_p - 0.5;
    if (d.q == 19) fine_params<19>(L.fine, (float)(1.0 / tau), (float)(1.0 / tau), d.force);
    else fine_params<27>(L.fine, (float)(1.0 / tau), (float)(1.0 / tau), d.force);
    L.c2f = (float)(0.5 * (tau - 1.0) / (tau_p - 1.0));
    L.f2c = (float)(2.0 * (tau_p - 1.0) / (tau - 1.0));
    tau_p = tau;
  }
  // each level knows its child's block map (coveredByChild)
  for (int m = 1; m + 1 < a->levels; m++) {
    PoolGeom& G = m == 1 ? a->G : a->deep[m - 2].G;
    const DeepLevel& c = a->deep[m - 1];
    G.child_block_slot = c.block_slot;
    G.cnbx = c.G.nbx;
    G.cnby = c.G.nby;
  }
  if (a->has_body && a->levels > 2) {
    a->deep.back().G.body =
        PoolBody{1, d.body.centre[0], d.body.centre[1], d.body.centre[2], d.body.radius};
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_amr_create(const lbm_amr_desc* desc, lbm_amr** out) {
  if (!out) return fail(LBM_E_INVALID, "null output handle");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  lbm_amr* a = new lbm_amr();
  a->d = *desc;
  lbm_desc l0{};
  l0.abi_version = LBM_ABI_VERSION;
  l0.q = desc->q;
  l0.nx = desc->nx; l0.ny = desc->ny; l0.nz = desc->nz;
  l0.collision = desc->collision;
  l0.omega = desc->omega;
  l0.omega_minus = desc->omega_minus;
  for (int k = 0; k < 3; k++) l0.force[k] = desc->force[k];
  l0.storage = LBM_STORE_F32;
  l0.sponge_width = desc->sponge_width;
  for (int k = 0; k < 3; k++) {
    l0.sponge_axes[k] = desc->sponge_axes[k];
    l0.sponge_u[k] = desc->sponge_u[k];
  }
  if ((rc = lbm_solve_create(&l0, &a->l0))) {
    release(a);
    return rc;
  }
  solve_no_eager_graphs(a->l0);
  a->stream = solve_stream(a->l0);
  const int rb = desc->rb;
  a->fb = 2 * rb + 2 * kGhost;
  PoolGeom& G = a->G;
  G.nx = desc->nx; G.ny = desc->ny; G.nz = desc->nz;
  G.rb = rb; G.fb = a->fb;
  const int dim = desc->q == 9 ? 2 : 3;
  G.nbx = desc->nx / rb; G.nby = desc->ny / rb; G.nbz = dim == 2 ? 1 : desc->nz / rb;
  G.ncells = (uint32_t)((size_t)desc->nx * desc->ny * desc->nz);
  G.max_slots = desc->max_slots;
  // Planes are padded by 4 KiB: the plane size max_slots * fb^3 is a
  // multiple of 2^14 elements for every power-of-two slot count, and the
  // pool step's Q streams at one plane stride apart would share their low
  // address bits (the same DRAM channels and banks).
  a->pool_cells = (size_t)desc->max_slots * a->fb * a->fb * (dim == 3 ? a->fb : 1);
  G.pool_plane = (uint32_t)(a->pool_cells + 1024);
  a->nblocks = G.nbx * G.nby * G.nbz;
  plan_capacity(rb, desc->max_slots, G.ncells, dim, &a->plan.cap_seam, &a->plan.cap_near, &a->plan.cap_ring);
  const size_t pool_bytes = (size_t)G.pool_plane * desc->q * 4;
  if ((rc = dalloc(&a->pool[0], pool_bytes)) || (rc = dalloc(&a->pool[1], pool_bytes)) ||
      (rc = dalloc(&a->mac_pool, (size_t)G.pool_plane * 16)) || (rc = dalloc(&a->block_slot, (size_t)a->nblocks * 4)) ||
      (rc = dalloc(&a->slot_to_block, (size_t)desc->max_slots * 4)) || (rc = dalloc(&a->want, (size_t)a->nblocks)) ||
      (rc = dalloc(&a->crit, (size_t)a->nblocks * 4)) ||
      (rc = dalloc(&a->dying, (size_t)(2 * a->nblocks + 1) * 4)) ||
      (rc = dalloc(&a->fresh, (size_t)(2 * a->nblocks + 1) * 4)) ||
      (rc = dalloc(&a->active, (size_t)(desc->max_slots + 1) * 4)) ||
      (rc = dalloc(&a->free_list, (size_t)desc->max_slots * 4)) || (rc = dalloc(&a->status, 16)) ||
      (rc = dalloc(&a->pairs, (size_t)(2 * desc->max_slots + 1) * 4)) ||
      (rc = dalloc(&a->ccls, (size_t)G.ncells)) || (rc = dalloc(&a->plan.count, 16)) ||
      (rc = dalloc(&a->plan.seam, (size_t)a->plan.cap_seam * 4)) ||
      (rc = dalloc(&a->plan.near, (size_t)a->plan.cap_near * 4)) ||
      (rc = dalloc(&a->plan.ring, (size_t)a->plan.cap_ring * 16)) ||
      (rc = dalloc(&a->tile_table, (size_t)desc->max_slots * kTileWords * 4)) ||
      (rc = dalloc(&a->partials, 256 * 4 * 8)) || (rc = dalloc(&a->counter, 16)) || (rc = dalloc(&a->totals, 32))) {
    release(a);
    return rc;
  }
  a->levels = desc->levels == 0 ? 2 : desc->levels;
  G.l0_scale = 1.f;
  G.l0_offset = 0.f;
  G.dx2 = 0.25f;
  if (desc->body.shape == LBM_BODY_SPHERE) {
    a->has_body = true;
    // the body lives on the finest level alone (main-3d.js:1577-1597)
    if (a->levels == 2)
      G.body = PoolBody{1, desc->body.centre[0], desc->body.centre[1], desc->body.centre[2], desc->body.radius};
    if ((rc = dalloc(&a->force_partials, (size_t)pool_force_blocks(rb, desc->max_slots) * 48)) ||
        (rc = dalloc(&a->history, (size_t)kForceCap * 48)) || (rc = dalloc(&a->written, 8))) {
      release(a);
      return rc;
    }
  }
  if (desc->manage_every > 0) {
    // the body is wanted on the finest level alone
    const std::vector<uint8_t> m = level_mask(*desc, 1, a->has_body && a->levels == 2);
    hipError_t e = hipSuccess;
    if ((rc = dalloc(&a->geom, m.size())) ||
        (e = hipMemcpy(a->geom, m.data(), m.size(), hipMemcpyHostToDevice)) != hipSuccess) {
      release(a);
      return rc ? rc : hip_fail(e, "geometry mask");
    }
  }
  if (desc->reflux) {
    a->plan.cap_flux = G.ncells;
    if ((rc = dalloc(&a->plan.flux, (size_t)G.ncells * 4)) || (rc = dalloc(&a->flux_acc, (size_t)G.ncells * 16))) {
      release(a);
      return rc;
    }
  }
  G.block_slot = a->block_slot;
  G.slot_to_block = a->slot_to_block;
  G.ccls = a->ccls;
  G.tile_table = a->tile_table;
  hipError_t e = hipMemsetAsync(a->block_slot, 0xff, (size_t)a->nblocks * 4, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->slot_to_block, 0xff, (size_t)desc->max_slots * 4, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->want, 0, (size_t)a->nblocks, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->crit, 0, (size_t)a->nblocks * 4, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->active, 0, 4, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->status, 0, 16, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->counter, 0, 16, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->mac_pool, 0, (size_t)G.pool_plane * 16, a->stream);
  if (e == hipSuccess) e = hipMemsetAsync(a->ccls, 0, (size_t)G.ncells, a->stream);
  if (e == hipSuccess && a->written) e = hipMemsetAsync(a->written, 0, 8, a->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(a->stream);
  if (e != hipSuccess) {
    release(a);
    return hip_fail(e, "AMR init");
  }
  // per-level relaxation tau_f = 2 tau_c - 1/2 (main-3d.js:1188-1193).  The
  // fine step is BGK at OMEGA_FINE with the L0 body force unchanged, as the
  // pool step's overrides are set (main-3d.js:1387-1399,
  // common_d3_amr_step1.wgsl:347-386), whatever the L0 collision is.
  const double tc = 1.0 / desc->omega, tf = 2.0 * tc - 0.5;
  if (desc->q == 9) fine_params<9>(a->fine, (float)(1.0 / tf), (float)(1.0 / tf), desc->force);
  else if (desc->q == 19) fine_params<19>(a->fine, (float)(1.0 / tf), (float)(1.0 / tf), desc->force);
  else fine_params<27>(a->fine, (float)(1.0 / tf), (float)(1.0 / tf), desc->force);
  // post-collision non-equilibrium rescale (common_d3_pool.wgsl:172-187)
  a->c2f = (float)(0.5 * (tf - 1.0) / (tc - 1.0));
  a->f2c = (float)(2.0 * (tc - 1.0) / (tf - 1.0));
  if ((rc = create_deep_levels(a, tf))) {
    release(a);
    return rc;
  }
  *out = a;
  return LBM_OK;
}

int lbm_amr_destroy(lbm_amr* a) {
  if (a) (void)hipStreamSynchronize(a->stream);
  release(a);
  return LBM_OK;
}

int lbm_amr_upload_topology(lbm_amr* a, const int32_t* block_slot, const int32_t* slot_to_block) {
  if (!a || !block_slot || !slot_to_block) return fail(LBM_E_INVALID, "null argument");
  if (a->managed) {
    int rc0 = sync_host_maps(a);
    if (rc0) return rc0;
  }
  // the two maps must be inverse to each other
  for (int b = 0; b < a->nblocks; b++) {
    const int s = block_slot[b];
    if (s >= a->d.max_slots || (s >= 0 && slot_to_block[s] != b))
      return fail(LBM_E_INVALID, "blockSlot[%d] = %d is not inverted by slotToBlock", b, s);
  }
  for (int s = 0; s < a->d.max_slots; s++) {
    const int b = slot_to_block[s];
    if (b >= a->nblocks || (b >= 0 && block_slot[b] != s))
      return fail(LBM_E_INVALID, "slotToBlock[%d] = %d is not inverted by blockSlot", s, b);
  }
  int rc_geo = check_topology(a->d, block_slot);
  if (rc_geo) return rc_geo;
  // with finer levels, the finest one's upload is what places the body
  a->body_covered = a->levels == 2;
  a->h_block_slot1.assign(block_slot, block_slot + a->nblocks);
  if (a->seeded) {
    int rc0 = restrict_active(a);
    if (rc0) return rc0;
  }
  HIPCHK(hipMemcpyAsync(a->block_slot, block_slot, (size_t)a->nblocks * 4, hipMemcpyHostToDevice, a->stream));
  HIPCHK(hipMemcpyAsync(a->slot_to_block, slot_to_block, (size_t)a->d.max_slots * 4, hipMemcpyHostToDevice, a->stream));
  int rc = refresh_active(a);
  if (rc) return rc;
  // finer levels' parent cells live in these tiles: rebuild their lists
  for (DeepLevel& D : a->deep)
    if (!D.h_block_slot.empty() && (rc = refresh_level(a, D))) return rc;
  if (a->seeded) {
    // fill every tile of the new topology from the current L0 state
    std::vector<int32_t> pairs(1 + 2 * a->nactive);
    pairs[0] = a->nactive;
    int k = 0;
    for (int s = 0; s < a->d.max_slots; s++)
      if (slot_to_block[s] >= 0) { pairs[1 + 2 * k] = slot_to_block[s]; pairs[2 + 2 * k] = s; k++; }
    HIPCHK(hipMemcpyAsync(a->fresh, pairs.data(), pairs.size() * 4, hipMemcpyHostToDevice, a->stream));
    rc = fill_tiles(a, a->fresh, a->nactive);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(a->stream));
  }
  return LBM_OK;
}

int lbm_amr_upload_level_topology(lbm_amr* a, int32_t level, const int32_t* block_slot,
                                  const int32_t* slot_to_block) {
  if (!a || !block_slot || !slot_to_block) return fail(LBM_E_INVALID, "null argument");
  if (a->managed) {
    // the manager moved tiles since the host copies were taken
    int rc0 = sync_host_maps(a);
    if (rc0) return rc0;
  }
  if (level == 1) return lbm_amr_upload_topology(a, block_slot, slot_to_block);
  if (level < 1 || level >= a->levels) return fail(LBM_E_INVALID, "level %d outside 1 .. %d", level, a->levels - 1);
  DeepLevel& L = a->deep[level - 2];
  for (int b = 0; b < L.nblocks; b++) {
    const int s = block_slot[b];
    if (s >= a->d.max_slots || (s >= 0 && slot_to_block[s] != b))
      return fail(LBM_E_INVALID, "level %d blockSlot[%d] = %d is not inverted by slotToBlock", level, b, s);
  }
  for (int s = 0; s < a->d.max_slots; s++) {
    const int b = slot_to_block[s];
    if (b >= L.nblocks || (b >= 0 && block_slot[b] != s))
      return fail(LBM_E_INVALID, "level %d slotToBlock[%d] = %d is not inverted by blockSlot", level, s, b);
  }
  int rc = check_level_topology(a, level, block_slot);
  if (rc) return rc;
  L.h_block_slot.assign(block_slot, block_slot + L.nblocks);
  L.h_slot_to_block.assign(slot_to_block, slot_to_block + a->d.max_slots);
  HIPCHK(hipMemcpyAsync(L.block_slot, block_slot, (size_t)L.nblocks * 4, hipMemcpyHostToDevice, a->stream));
  HIPCHK(hipMemcpyAsync(L.slot_to_block, slot_to_block, (size_t)a->d.max_slots * 4, hipMemcpyHostToDevice,
                        a->stream));
  // this level's lists, and its children's (whose parent cells changed)
  for (int m = level; m < a->levels; m++) {
    DeepLevel& D = a->deep[m - 2];
    if (m > level && D.h_block_slot.empty()) continue;
    if ((rc = refresh_level(a, D))) return rc;
    if (a->seeded && (rc = fill_level(a, D))) return rc;
  }
  if (level == a->levels - 1) a->body_covered = true;
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_read_topology(lbm_amr* a, int32_t* block_slot, int32_t* slot_to_block) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (block_slot) HIPCHK(hipMemcpyAsync(block_slot, a->block_slot, (size_t)a->nblocks * 4, hipMemcpyDeviceToHost, a->stream));
  if (slot_to_block)
    HIPCHK(hipMemcpyAsync(slot_to_block, a->slot_to_block, (size_t)a->d.max_slots * 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_read_level_topology(lbm_amr* a, int32_t level, int32_t* block_slot, int32_t* slot_to_block) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (level == 1) return lbm_amr_read_topology(a, block_slot, slot_to_block);
  if (level < 1 || level >= a->levels) return fail(LBM_E_INVALID, "level %d outside 1 .. %d", level, a->levels - 1);
  const DeepLevel& L = a->deep[level - 2];
  if (block_slot) HIPCHK(hipMemcpyAsync(block_slot, L.block_slot, (size_t)L.nblocks * 4, hipMemcpyDeviceToHost, a->stream));
  if (slot_to_block)
    HIPCHK(hipMemcpyAsync(slot_to_block, L.slot_to_block, (size_t)a->d.max_slots * 4, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_upload_macro(lbm_amr* a, const float* mac4) {
  if (!a || !mac4) return fail(LBM_E_INVALID, "null argument");
  int rc = lbm_upload_macro(a->l0, mac4);
  if (rc) return rc;
  a->seeded = true;
  a->step = 0;
  if (a->written) HIPCHK(hipMemsetAsync(a->written, 0, 8, a->stream));
  // fill every tile of the current topology from the seeded L0 field
  std::vector<int32_t> s2b(a->d.max_slots);
  HIPCHK(hipMemcpy(s2b.data(), a->slot_to_block, s2b.size() * 4, hipMemcpyDeviceToHost));
  std::vector<int32_t> pairs(1);
  for (int s = 0; s < a->d.max_slots; s++)
    if (s2b[s] >= 0) { pairs.push_back(s2b[s]); pairs.push_back(s); }
  const int n = (int)(pairs.size() - 1) / 2;
  pairs[0] = n;
  HIPCHK(hipMemcpyAsync(a->fresh, pairs.data(), pairs.size() * 4, hipMemcpyHostToDevice, a->stream));
  rc = fill_tiles(a, a->fresh, n);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(a->stream));
  rc = refresh_active(a);
  if (rc) return rc;
  // each finer level from its just-filled parent
  for (DeepLevel& L : a->deep)
    if ((rc = fill_level(a, L))) return rc;
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_step(lbm_amr* a, int32_t n) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (!a->seeded) return fail(LBM_E_STATE, "step before upload_macro");
  if (a->has_body && !a->body_covered && a->d.manage_every <= 0)
    return fail(LBM_E_STATE, "step with a body before a topology that holds it");
  for (int s = 0; s < n; s++) {
    int rc = macro_step(a);
    if (rc) return rc;
  }
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_manage_now(lbm_amr* a) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (!a->seeded) return fail(LBM_E_STATE, "manage before upload_macro");
  int rc = manage(a);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(a->stream));
  return LBM_OK;
}

int lbm_amr_read_macro(lbm_amr* a, float* mac4) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  return lbm_read_macro(a->l0, mac4);
}

int lbm_amr_read_pool_macro(lbm_amr* a, float* mac_pool) {
  if (!a || !mac_pool) return fail(LBM_E_INVALID, "null argument");
  for (int k = 0; k < 4; k++)
    HIPCHK(hipMemcpy(mac_pool + (size_t)k * a->pool_cells, a->mac_pool + (size_t)k * a->G.pool_plane,
                     a->pool_cells * 4, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_amr_read_f(lbm_amr* a, float* f_l0, float* f_pool) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (f_l0) {
    int rc = lbm_read_f(a->l0, f_l0);
    if (rc) return rc;
  }
  if (f_pool) {
    const size_t n = a->pool_cells;
    for (int i = 0; i < a->d.q; i++)
      HIPCHK(hipMemcpy(f_pool + (size_t)i * n, a->pool[0] + (size_t)i * a->G.pool_plane, n * 4,
                       hipMemcpyDeviceToHost));
    for (int i = 0; i < a->d.q; i++) {
      const float w = lattice_w(a->d.q, i);
      float* p = f_pool + (size_t)i * n;
      for (size_t c = 0; c < n; c++) p[c] += w;
    }
  }
  return LBM_OK;
}

int lbm_amr_read_level_f(lbm_amr* a, int32_t level, float* f_pool) {
  if (!a || !f_pool) return fail(LBM_E_INVALID, "null argument");
  if (level < 1 || level >= a->levels) return fail(LBM_E_INVALID, "level %d outside 1 .. %d", level, a->levels - 1);
  if (level == 1) return lbm_amr_read_f(a, nullptr, f_pool);
  const DeepLevel& L = a->deep[level - 2];
  const size_t n = a->pool_cells;
  HIPCHK(hipStreamSynchronize(a->stream));
  for (int i = 0; i < a->d.q; i++) {
    HIPCHK(hipMemcpy(f_pool + (size_t)i * n, L.pool[0] + (size_t)i * L.G.pool_plane, n * 4, hipMemcpyDeviceToHost));
    const float w = lattice_w(a->d.q, i);
    float* p = f_pool + (size_t)i * n;
    for (size_t c = 0; c < n; c++) p[c] += w;
  }
  return LBM_OK;
}

int lbm_amr_read_level_macro(lbm_amr* a, int32_t level, float* mac_pool) {
  if (!a || !mac_pool) return fail(LBM_E_INVALID, "null argument");
  if (level < 1 || level >= a->levels) return fail(LBM_E_INVALID, "level %d outside 1 .. %d", level, a->levels - 1);
  if (level == 1) return lbm_amr_read_pool_macro(a, mac_pool);
  const DeepLevel& L = a->deep[level - 2];
  HIPCHK(hipStreamSynchronize(a->stream));
  for (int k = 0; k < 4; k++)
    HIPCHK(hipMemcpy(mac_pool + (size_t)k * a->pool_cells, L.mac_pool + (size_t)k * L.G.pool_plane,
                     a->pool_cells * 4, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_amr_read_criterion(lbm_amr* a, float* crit, uint8_t* want) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (crit) HIPCHK(hipMemcpy(crit, a->crit, (size_t)a->nblocks * 4, hipMemcpyDeviceToHost));
  if (want) HIPCHK(hipMemcpy(want, a->want, (size_t)a->nblocks, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_amr_read_active(lbm_amr* a, int32_t* list) {
  if (!a || !list) return fail(LBM_E_INVALID, "null argument");
  HIPCHK(hipMemcpy(list, a->active, (size_t)(a->nactive + 1) * 4, hipMemcpyDeviceToHost));
  return LBM_OK;
}

int lbm_amr_read_forces(lbm_amr* a, double* out, int32_t max_records, int32_t* count) {
  if (!a || !count || (max_records > 0 && !out)) return fail(LBM_E_INVALID, "null argument");
  *count = 0;
  if (!a->written) return LBM_OK;  // no body: no records
  unsigned long long written = 0;
  HIPCHK(hipMemcpyAsync(&written, a->written, 8, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  const unsigned long long avail = std::min<unsigned long long>(written, kForceCap);
  const unsigned long long take = std::min<unsigned long long>(avail, (unsigned long long)std::max(0, max_records));
  if (take) {
    std::vector<double> ring((size_t)kForceCap * 6);
    HIPCHK(hipMemcpyAsync(ring.data(), a->history, ring.size() * 8, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipStreamSynchronize(a->stream));
    for (unsigned long long k = 0; k < take; k++)
      memcpy(out + k * 6, ring.data() + ((written - take + k) % kForceCap) * 6, 48);
  }
  *count = (int32_t)take;
  return LBM_OK;
}

int lbm_amr_read_totals(lbm_amr* a, double* out4) {
  if (!a || !out4) return fail(LBM_E_INVALID, "null argument");
  HIPCHK(launch_pool_totals(a->d.q, solve_lattice(a->l0, 0), a->pool[0], a->G, a->active, a->partials, a->counter,
                            a->totals, a->stream));
  HIPCHK(hipMemcpyAsync(out4, a->totals, 32, hipMemcpyDeviceToHost, a->stream));
  HIPCHK(hipStreamSynchronize(a->stream));
  // finer levels in order, each weighted 8^-m, summed here in FP64
  for (const DeepLevel& L : a->deep) {
    if (L.nactive <= 0) continue;
    double v[4];
    HIPCHK(launch_level_totals(a->d.q, L.pool[0], L.G, L.active, std::ldexp(1.0, -3 * L.m), a->partials, a->counter,
                               a->totals, a->stream));
    HIPCHK(hipMemcpyAsync(v, a->totals, 32, hipMemcpyDeviceToHost, a->stream));
    HIPCHK(hipStreamSynchronize(a->stream));
    for (int k = 0; k < 4; k++) out4[k] += v[k];
  }
  return LBM_OK;
}

int lbm_amr_info(lbm_amr* a, int64_t* step, int32_t* active_tiles, int32_t* exhausted) {
  if (!a) return fail(LBM_E_INVALID, "null argument");
  if (step) *step = a->step;
  if (active_tiles) *active_tiles = a->nactive;
  if (exhausted) *exhausted = a->exhausted;
  return LBM_OK;
}

}  // extern "C"

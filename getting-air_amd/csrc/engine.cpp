// Host side of the dense engine: the lbm_solve object behind include/lbm.h.
//
// A solve owns two population lattices (ping-pong, shifted storage), the
// macroscopic field, the body classification, the per-tile force slots and
// one HIP stream.  Each step is one fused stream-collide launch (which also
// reduces the momentum-exchange force of the tiles touching the body into
// FP64 slots) followed, when a body is present, by a one-wave launch that
// sums the slots in a fixed order into the force history.  Only the last step
// of an lbm_step(n) call writes the macroscopic field, which is what the
// reference's readbacks observe (main-3d.js:3506-3539).  Chunks of steps are
// replayed from captured hipGraphs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/lbm.h"
#include "dense.h"
#include "rigid.h"
#include "engine.h"
#include "lattice_pairs.h"

namespace lbm {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(LBM_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

}  // namespace lbm

using namespace lbm;

#define HIPCHK(call)                                 \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

namespace {

// D2Q9 keeps the plainly rounded f32 weights (lattice-2d.mjs:119-125), whose
// sum is 1 + this; the shifted collision adds it back into rho - 1.
const float kD2Q9WeightSumMinus1 =
    (float)((double)0.44444444775581360f + 4.0 * (double)0.11111111193895340f + 4.0 * (double)0.027777777984738350f - 1.0);

// Tile shapes the autotuner times; block_x * block_y <= 256 (launch bound).
const int kShapes[][2] = {{64, 1}, {64, 2}, {64, 4}, {128, 1}, {128, 2}, {256, 1}, {32, 4}, {32, 8}};

struct TuneKey {
  int q, storage, general, nx, ny, nz, diffuse;
  bool operator<(const TuneKey& o) const {
    return std::tie(q, storage, general, nx, ny, nz, diffuse) <
           std::tie(o.q, o.storage, o.general, o.nx, o.ny, o.nz, o.diffuse);
  }
};
struct TuneVal { int bx, by, swz, nt, occ, lin; };
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune_cache;

const uint32_t kHistoryCap = 1u << 16;
const int kGraphChunk = 16;  // steps per captured graph (even: parity returns)

const int kEx9[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1}, kEy9[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
const int kEx3[27] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1};
const int kEy3[27] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1};
const int kEz3[27] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, -1, 1};

}  // namespace

struct lbm_solve {
  lbm_desc d{};
  int q = 0, dim = 3, storage = 0, elem = 4;
  uint32_t N = 0;
  bool general = false;
  bool body = false;
  void* f[2] = {nullptr, nullptr};
  int parity = 0;
  float* mac = nullptr;
  uint8_t* cls = nullptr;
  uint32_t* links = nullptr;        // bounce-back link masks of the cells next to the body
  std::vector<uint32_t> boundary;   // fluid cells with a solid streaming source
  int32_t* force_slot = nullptr;    // per tile of the current shape
  double* force_partials = nullptr;
  uint32_t nslots = 0;
  int slot_bx = 0, slot_by = 0, slot_lin = -1;
  size_t slot_bytes = 0, partial_bytes = 0;
  double* scratch = nullptr;        // totals partials
  unsigned* counters = nullptr;
  double* history = nullptr;
  unsigned long long* written = nullptr;
  double* totals = nullptr;
  uint32_t* salt = nullptr;         // dither step counters, one per lattice parity
  // the counter the next step reads is salt0 + (step - step_at_salt0): the
  // kernels advance the device words by one per step, and an upload
  // restarts them at 0 while a loaded state sets them
  uint32_t salt0 = 0;
  long long step_at_salt0 = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int bx = 64, by = 1, swz = 1, nt = 0;
  int occ = 0;                      // waves-per-SIMD cap of the step launches (0: none)
  int lin = 0;                      // linear tiles (blocks of consecutive cells)
  bool force_pending = false;       // last step's force partials not yet summed
  bool tuned = false;
  bool initialised = false;
  long long step = 0;
  size_t bytes = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  int graph_bx = 0, graph_by = 0, graph_swz = 0, graph_nt = 0, graph_occ = 0, graph_lin = 0;
  Arena* arena = nullptr;           // pool the buffers come from (serve loop), or null
  bool own_stream = true;
  // Both parity graphs and the autotuned tile are made when the state is set
  // (upload, load, tile change), so lbm_step only replays.  Off for serve
  // members and the AMR engine's L0 solve, which never run lbm_step chunks.
  bool eager_graphs = true;
  // moving rigid body (lbm_rigid_attach): device record, force partials,
  // launch arguments of its kernels
  lbm::RigidBody* rigid = nullptr;
  double* rigid_partials = nullptr;
  unsigned* rigid_counter = nullptr;
  lbm::RigidArgs rargs{};
};

namespace {

template <class T>
int dalloc(lbm_solve* s, T** p, size_t bytes) {
  if (s->arena) {
    *p = static_cast<T*>(s->arena->alloc(bytes));
    if (!*p) return fail(LBM_E_NOMEM, "pool has no room for %zu bytes", bytes);
  } else {
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess) {
      *p = nullptr;
      return fail(LBM_E_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    }
  }
  s->bytes += bytes;
  return LBM_OK;
}

void dfree(lbm_solve* s, void* p, size_t bytes) {
  if (!p) return;
  if (s->arena) s->arena->release(p);
  else (void)hipFree(p);
  s->bytes -= bytes;
}

int validate(const lbm_desc* d) {
  if (!d) return fail(LBM_E_INVALID, "null descriptor");
  if (d->abi_version != LBM_ABI_VERSION)
    return fail(LBM_E_INVALID, "descriptor abi_version %d, library %d", d->abi_version, LBM_ABI_VERSION);
  if (d->q != 9 && d->q != 19 && d->q != 27) return fail(LBM_E_INVALID, "q must be 9, 19 or 27 (got %d)", d->q);
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(LBM_E_INVALID, "grid %dx%dx%d", d->nx, d->ny, d->nz);
  if (d->q == 9 && d->nz != 1) return fail(LBM_E_INVALID, "D2Q9 needs nz = 1 (got %d)", d->nz);
  const double n = (double)d->nx * d->ny * d->nz;
  // one population plane must stay below 4 GiB (32-bit byte offsets in the kernels)
  if (n * 4.0 >= 4294967296.0) return fail(LBM_E_INVALID, "grid too large (%g cells; the limit is 2^30 per lattice)", n);
  if (!(d->omega > 0.f && d->omega < 2.f)) return fail(LBM_E_INVALID, "omega %g outside (0, 2)", d->omega);
  if (d->collision != LBM_BGK && d->collision != LBM_TRT) return fail(LBM_E_INVALID, "collision %d", d->collision);
  if (d->collision == LBM_TRT && !(d->omega_minus > 0.f && d->omega_minus < 2.f))
    return fail(LBM_E_INVALID, "omega_minus %g outside (0, 2)", d->omega_minus);
  if (d->storage < LBM_STORE_F32 || d->storage > LBM_STORE_BF16) return fail(LBM_E_INVALID, "storage %d", d->storage);
  const int sh = d->body.shape;
  if (sh != LBM_BODY_NONE) {
    if (d->q == 9 && sh != LBM_BODY_CIRCLE) return fail(LBM_E_INVALID, "D2Q9 bodies must be circles");
    if (d->q != 9 && sh != LBM_BODY_SPHERE) return fail(LBM_E_INVALID, "3D bodies must be spheres");
    if (!(d->body.radius > 0.f)) return fail(LBM_E_INVALID, "body radius %g", d->body.radius);
  }
  if (d->diffuse && sh == LBM_BODY_NONE) return fail(LBM_E_INVALID, "diffuse coupling without a body");
  if (d->diffuse != 0 && d->diffuse != 1) return fail(LBM_E_INVALID, "diffuse %d: 0 or 1", d->diffuse);
  if (!std::isfinite(d->chi_eps)) return fail(LBM_E_INVALID, "chi_eps %g", d->chi_eps);
  for (int a = 0; a < 3; a++) {
    if (!std::isfinite(d->force[a]) || !std::isfinite(d->sponge_u[a]))
      return fail(LBM_E_INVALID, "non-finite force or sponge velocity");
  }
  return LBM_OK;
}

// Per-pair collision constants of k_step, doubled: the kernel evaluates the
// pair form in doubled variables (P2 = 2P, M2 = 2M), and twice a binary32
// value is exact.
template <int Q>
void pair_constants(StepArgs& a, const lbm_desc& d) {
  using L = Lattice<Q>;
  const float om = a.omega, omm = a.omega_m;
  const float U[3] = {d.sponge_u[0], d.sponge_u[1], d.sponge_u[2]};
  const float uu = U[0] * U[0] + U[1] * U[1] + U[2] * U[2];
  for (int k = 0; k < Pairs<Q>::NP; k++) {
    const int i = Pairs<Q>::lo(k);
    const float w = L::w[i];
    const float eF = L::ex[i] * d.force[0] + L::ey[i] * d.force[1] + L::ez[i] * d.force[2];
    const float eU = L::ex[i] * U[0] + L::ey[i] * U[1] + L::ez[i] * U[2];
    a.c1[k] = 2.f * (9.f * (1.f - 0.5f * om) * w * eF);
    a.so[k] = 2.f * ((1.f - 0.5f * omm) * 3.f * w * eF);
    a.ae[k] = 2.f * (w * (4.5f * eU * eU - 1.5f * uu));
    a.bo[k] = 2.f * (3.f * w * eU);
  }
  a.ae[13] = L::w[0] * (-1.5f * uu);  // the rest population is not doubled
}

// CHI_EPS of the diffuse coupling (main-3d.js:1026 default 1.5).
float chi_eps(const lbm_desc& d) { return d.chi_eps > 0.f ? d.chi_eps : 1.5f; }

StepArgs step_args(lbm_solve* s, bool write_mac, bool with_force) {
  StepArgs a{};
  const lbm_desc& d = s->d;
  a.omega_m = d.collision == LBM_TRT ? d.omega_minus : d.omega;
  a.nx = d.nx; a.ny = d.ny; a.nz = d.nz;
  a.ncells = s->N;
  a.omega = d.omega;
  a.guo = 1.f - 0.5f * d.omega;
  a.wsum_m1 = s->q == 9 ? kD2Q9WeightSumMinus1 : 0.f;
  for (int k = 0; k < 3; k++) a.F[k] = d.force[k];
  a.wall_mask = (d.walls[0] ? 1 : 0) | (d.walls[1] ? 2 : 0) | (d.walls[2] && s->dim == 3 ? 4 : 0);
  a.wall_u0 = d.wall_u[0];
  a.wall_u1 = d.wall_u[1];
  a.sponge_w = d.sponge_width;
  a.sponge_mask = (d.sponge_axes[0] ? 1 : 0) | (d.sponge_axes[1] ? 2 : 0) | (d.sponge_axes[2] && s->dim == 3 ? 4 : 0);
  for (int k = 0; k < 3; k++) a.su[k] = d.sponge_u[k];
  a.cls = s->body || s->rigid ? s->cls : nullptr;
  a.links = s->body || s->rigid ? s->links : nullptr;
  a.rigid = s->rigid;
  a.solid_eq = d.solid_eq;
  a.solid_mac_rest = s->dim == 3;
  a.mac = write_mac ? s->mac : nullptr;
  a.salt_in = s->salt + s->parity;
  a.salt_out = s->salt + (1 - s->parity);
  if (with_force && s->body && s->nslots) {
    // partial buffers alternate with the lattice parity
    const size_t stride = (size_t)s->nslots * 6;
    a.force_slot = s->force_slot;
    a.force_partials = s->force_partials + stride * s->parity;
    a.prev_partials = s->force_partials + stride * (1 - s->parity);
    a.nslots = s->nslots;
    a.history = s->history;
    a.written = s->written;
    a.hist_cap = kHistoryCap;
  }
  a.body_c[0] = d.body.centre[0];
  a.body_c[1] = d.body.centre[1];
  a.body_c[2] = s->dim == 3 ? d.body.centre[2] : 0.0;
  a.periodic_arm = s->dim == 2;
  a.body_shape = d.body.shape == LBM_BODY_CIRCLE ? 2 : 1;
  a.body_r = d.body.radius;
  a.body_cf[0] = d.body.centre[0];
  a.body_cf[1] = d.body.centre[1];
  a.body_cf[2] = s->dim == 3 ? d.body.centre[2] : 0.f;
  a.chi_eps = chi_eps(d);
  if (s->lin) {
    const uint32_t t = (uint32_t)(s->bx * s->by);
    a.linear = 1;
    a.nblocks = (s->N + t - 1) / t;
    a.div_nx = make_invariant_div((uint32_t)d.nx);
    a.div_ny = make_invariant_div((uint32_t)d.ny);
  } else {
    a.ntx = (d.nx + s->bx - 1) / s->bx;
    a.nty = (d.ny + s->by - 1) / s->by;
    a.nblocks = (uint32_t)a.ntx * a.nty * d.nz;
  }
  a.swizzle = s->swz;
  if (s->q == 9) pair_constants<9>(a, d);
  else if (s->q == 19) pair_constants<19>(a, d);
  else pair_constants<27>(a, d);
  return a;
}

// Maps every tile of the current shape that holds a boundary cell to a slot,
// in ascending tile order (so the finalize sum order is fixed per shape).
int build_force_slots(lbm_solve* s) {
  if (!s->body) return LBM_OK;
  if (s->slot_bx == s->bx && s->slot_by == s->by && s->slot_lin == s->lin && s->force_slot) return LBM_OK;
  const lbm_desc& d = s->d;
  const int ntx = (d.nx + s->bx - 1) / s->bx, nty = (d.ny + s->by - 1) / s->by;
  const uint32_t t = (uint32_t)(s->bx * s->by);
  const size_t ntiles = s->lin ? (s->N + t - 1) / t : (size_t)ntx * nty * d.nz;
  std::vector<int32_t> slot(ntiles, -1);
  for (uint32_t cell : s->boundary) {
    const int x = cell % d.nx, y = (cell / d.nx) % d.ny, z = cell / (d.nx * d.ny);
    slot[s->lin ? cell / t : ((size_t)z * nty + y / s->by) * ntx + x / s->bx] = 0;
  }
  uint32_t n = 0;
  for (auto& v : slot)
    if (v == 0) v = (int32_t)n++;
  HIPCHK(hipStreamSynchronize(s->stream));
  dfree(s, s->force_slot, s->slot_bytes);
  dfree(s, s->force_partials, s->partial_bytes);
  s->force_slot = nullptr;
  s->force_partials = nullptr;
  s->slot_bytes = ntiles * 4;
  s->partial_bytes = 2 * std::max<size_t>(n, 1) * 6 * 8;  // one set per lattice parity
  int rc = dalloc(s, &s->force_slot, s->slot_bytes);
  if (rc) return rc;
  rc = dalloc(s, &s->force_partials, s->partial_bytes);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(s->force_slot, slot.data(), ntiles * 4, hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  s->nslots = n;
  s->slot_bx = s->bx;
  s->slot_by = s->by;
  s->slot_lin = s->lin;
  return LBM_OK;
}

// One step: fused stream-collide with this step's force partials.  With
// `finalize_prev` the same launch also appends the previous step's force
// record; flush_force appends the last one.
int enqueue_step(lbm_solve* s, bool write_mac, bool finalize_prev) {
  const void* fin = s->f[s->parity];
  void* fout = s->f[1 - s->parity];
  if (s->rigid) {
    // force at the current pose and the integrator, then the new pose's cells
    hipError_t e = launch_rigid_force(s->q, s->storage, fin, s->rargs, s->stream);
    if (e == hipSuccess) e = launch_rigid_classify(s->q, s->rargs, s->stream);
    if (e != hipSuccess) return hip_fail(e, "rigid body launch");
  }
  StepArgs a = step_args(s, write_mac, true);
  a.finalize_prev = finalize_prev && a.force_slot ? 1 : 0;
  hipError_t e = launch_step(s->q, s->storage, s->general, s->nt, s->occ, fin, fout, a, s->bx, s->by, s->stream,
                             s->d.diffuse != 0);
  if (e != hipSuccess) return hip_fail(e, "step kernel launch");
  s->force_pending = a.force_slot != nullptr;
  s->parity ^= 1;
  s->step++;
  return LBM_OK;
}

// Appends the force record of the last enqueued step.
int flush_force(lbm_solve* s) {
  if (!s->force_pending) return LBM_OK;
  ForceFinalizeArgs fa{s->force_partials + (size_t)s->nslots * 6 * (1 - s->parity), s->nslots, s->history,
                       s->written, kHistoryCap};
  hipError_t e = launch_force_finalize(fa, s->stream);
  if (e != hipSuccess) return hip_fail(e, "force finalize launch");
  s->force_pending = false;
  return LBM_OK;
}

}  // namespace

namespace lbm {

int solve_enqueue_covered_step(lbm_solve* s, const int32_t* cover, int rb, int nbx, int nby) {
  const void* fin = s->f[s->parity];
  void* fout = s->f[1 - s->parity];
  StepArgs a = step_args(s, true, false);
  a.cover = cover;
  a.cover_rb = rb;
  a.cover_nbx = nbx;
  a.cover_nby = nby;
  hipError_t e = launch_step(s->q, s->storage, true, s->nt, s->occ, fin, fout, a, s->bx, s->by, s->stream);
  if (e != hipSuccess) return hip_fail(e, "covered step launch");
  s->parity ^= 1;
  s->step++;
  return LBM_OK;
}

float* solve_lattice(lbm_solve* s, int which) {
  return static_cast<float*>(s->f[which ? 1 - s->parity : s->parity]);
}
float* solve_mac(lbm_solve* s) { return s->mac; }
void solve_no_eager_graphs(lbm_solve* s) { s->eager_graphs = false; }
hipStream_t solve_stream(lbm_solve* s) { return s->stream; }

}  // namespace lbm

namespace {

void drop_graphs(lbm_solve* s) {
  for (int p = 0; p < 2; p++) {
    if (s->graph[p]) (void)hipGraphExecDestroy(s->graph[p]);
    s->graph[p] = nullptr;
  }
}

// Captures kGraphChunk steps starting at the current parity.
int build_graph(lbm_solve* s) {
  const int p = s->parity;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  const long long step0 = s->step;
  const bool pending0 = s->force_pending;
  int rc = LBM_OK;
  for (int k = 0; k < kGraphChunk && rc == LBM_OK; k++) rc = enqueue_step(s, false, true);
  hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->step = step0;
  s->parity = p;
  s->force_pending = pending0;
  if (rc != LBM_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  e = hipGraphInstantiate(&s->graph[p], g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  s->graph_bx = s->bx; s->graph_by = s->by; s->graph_swz = s->swz; s->graph_nt = s->nt; s->graph_occ = s->occ;
  s->graph_lin = s->lin;
  return LBM_OK;
}

int autotune(lbm_solve* s);

// Autotunes (when the descriptor asks and no tile is set yet) and captures
// the graph chunks of both lattice parities, so that the first long
// lbm_step neither times tiles nor instantiates graphs.
int prepare_graphs(lbm_solve* s) {
  if (!s->eager_graphs || !s->initialised) return LBM_OK;
  if (s->d.autotune && !s->tuned) {
    int rc = autotune(s);
    if (rc) return rc;
  }
  if (s->graph_bx != s->bx || s->graph_by != s->by || s->graph_swz != s->swz || s->graph_nt != s->nt ||
      s->graph_occ != s->occ || s->graph_lin != s->lin)
    drop_graphs(s);
  const int p0 = s->parity;
  for (int p = 0; p < 2; p++) {
    if (s->graph[p]) continue;
    s->parity = p;
    int rc = build_graph(s);
    s->parity = p0;
    if (rc) return rc;
  }
  return LBM_OK;
}

// Times every candidate tile shape, XCD mapping, load/store flavour and
// occupancy cap on the live lattices (fin -> fout without swapping, so the
// state is unchanged), then times the four fastest again, interleaved, and
// keeps the fastest of those: one short timing per candidate is too noisy to
// separate shapes within a few percent.
int autotune(lbm_solve* s) {
  TuneKey key{s->q, s->storage, s->general ? 1 : 0, s->d.nx, s->d.ny, s->d.nz, s->d.diffuse};
  TuneVal bestv{s->bx, s->by, s->swz, s->nt, s->occ, s->lin};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      bestv = it->second;
      cached = true;
    }
  }
  if (!cached) {
    const int reps = s->N >= (1u << 20) ? 3 : 8;
    const void* fin = s->f[s->parity];
    void* fout = s->f[1 - s->parity];
    auto time_of = [&](const TuneVal& v, int n, float* ms) -> int {
      s->bx = v.bx; s->by = v.by; s->swz = v.swz; s->nt = v.nt; s->occ = v.occ; s->lin = v.lin;
      StepArgs a = step_args(s, false, false);
      const bool diff = s->d.diffuse != 0;
      HIPCHK(launch_step(s->q, s->storage, s->general, v.nt, v.occ, fin, fout, a, v.bx, v.by, s->stream, diff));
      HIPCHK(hipEventRecord(s->ev0, s->stream));
      for (int r = 0; r < n; r++)
        HIPCHK(launch_step(s->q, s->storage, s->general, v.nt, v.occ, fin, fout, a, v.bx, v.by, s->stream, diff));
      HIPCHK(hipEventRecord(s->ev1, s->stream));
      HIPCHK(hipEventSynchronize(s->ev1));
      HIPCHK(hipEventElapsedTime(ms, s->ev0, s->ev1));
      return LBM_OK;
    };
    std::vector<std::pair<float, TuneVal>> cand;
    for (int lin = 0; lin < 2; lin++)
      for (const auto& sh : kShapes) {
        // linear tiles only need a thread count: one shape per count
        if (lin && sh[1] != 1) continue;
        if (!lin && sh[0] > 64 && s->d.nx < sh[0] / 2) continue;
        for (int occ : {0, 4, 6}) {
          if (!occupancy_cap_ok(occ, sh[0] * sh[1])) continue;
          for (int swz = 0; swz < 2; swz++)
            for (int nt = 0; nt < 2; nt++) {
              const TuneVal v{sh[0], sh[1], swz, nt, occ, lin};
              float ms = 0.f;
              int rc = time_of(v, reps, &ms);
              if (rc) return rc;
              cand.push_back({ms, v});
            }
        }
      }
    std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    cand.resize(std::min<size_t>(cand.size(), 4));
    for (auto& c : cand) c.first = 1e30f;
    for (int round = 0; round < 3; round++)
      for (auto& c : cand) {
        float ms = 0.f;
        int rc = time_of(c.second, 2 * reps, &ms);
        if (rc) return rc;
        c.first = std::min(c.first, ms);
      }
    std::sort(cand.begin(), cand.end(), [](const auto& a, const auto& b) { return a.first < b.first; });
    if (!cand.empty()) bestv = cand.front().second;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = bestv;
  }
  s->bx = bestv.bx; s->by = bestv.by; s->swz = bestv.swz; s->nt = bestv.nt; s->occ = bestv.occ;
  s->lin = bestv.lin;
  s->tuned = true;
  return build_force_slots(s);
}

// GPU classification -> host list of fluid cells with a solid streaming
// source, ascending.
int build_boundary(lbm_solve* s) {
  const lbm_desc& d = s->d;
  ClassifyArgs c{};
  c.nx = d.nx; c.ny = d.ny; c.nz = d.nz;
  c.shape = d.body.shape == LBM_BODY_CIRCLE ? 2 : 1;
  c.radius = d.body.radius;
  c.cx = d.body.centre[0]; c.cy = d.body.centre[1]; c.cz = s->dim == 3 ? d.body.centre[2] : 0.f;
  // the diffuse coupling's force cells: chi >= 1e-6 needs phi / eps <
  // atanh(1 - 2e-6) = 6.9; the band is wider, the kernel tests chi itself
  c.band = d.diffuse ? 7.5f * chi_eps(d) : -1e30f;
  c.cls = s->cls;
  c.links = s->links;
  HIPCHK(launch_classify(c, s->stream));
  HIPCHK(launch_links(s->q, c, s->stream));
  std::vector<uint8_t> h(s->N);
  HIPCHK(hipMemcpyAsync(h.data(), s->cls, s->N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  s->boundary.clear();
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const uint32_t cell = (uint32_t)((z * ny + y) * nx + x);
        const uint8_t cc = h[cell];
        if (d.diffuse) {
          if (cc & kClsBand) s->boundary.push_back(cell);
          continue;
        }
        if (!(cc & 2) || (cc & 1)) continue;
        for (int i = 1; i < s->q; i++) {
          const int ex = s->q == 9 ? kEx9[i] : kEx3[i];
          const int ey = s->q == 9 ? kEy9[i] : kEy3[i];
          const int ez = s->q == 9 ? 0 : kEz3[i];
          const int sx = (x - ex + nx) % nx, sy = (y - ey + ny) % ny, sz = (z - ez + nz) % nz;
          if (h[(uint32_t)((sz * ny + sy) * nx + sx)] & 1) {
            s->boundary.push_back(cell);
            break;
          }
        }
      }
  return build_force_slots(s);
}

void release(lbm_solve* s) {
  if (!s) return;
  drop_graphs(s);
  void* ptrs[] = {s->f[0], s->f[1], s->mac, s->cls, s->links, s->force_slot, s->force_partials, s->scratch,
                  s->counters, s->history, s->written, s->totals, s->salt};
  for (void* p : ptrs) {
    if (!p) continue;
    if (s->arena) s->arena->release(p);
    else (void)hipFree(p);
  }
  for (void* p : {(void*)s->rigid, (void*)s->rigid_partials, (void*)s->rigid_counter})
    if (p) (void)hipFree(p);
  if (s->ev0) (void)hipEventDestroy(s->ev0);
  if (s->ev1) (void)hipEventDestroy(s->ev1);
  if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

int reset_history(lbm_solve* s) {
  s->force_pending = false;
  HIPCHK(hipMemsetAsync(s->written, 0, sizeof(unsigned long long), s->stream));
  return LBM_OK;
}

}  // namespace

namespace {

const int kTotalsBlocks = 1024;  // max totals partial blocks

int create_impl(const lbm_desc* desc, Arena* arena, hipStream_t stream, int bx, int by, int flags,
                lbm_solve** out) {
  if (!out) return fail(LBM_E_INVALID, "null output handle");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  lbm_solve* s = new lbm_solve();
  s->arena = arena;
  s->eager_graphs = arena == nullptr;
  s->d = *desc;
  s->q = desc->q;
  s->dim = desc->q == 9 ? 2 : 3;
  s->storage = desc->storage;
  s->elem = desc->storage == LBM_STORE_F32 ? 4 : 2;
  s->N = (uint32_t)((size_t)desc->nx * desc->ny * desc->nz);
  s->body = desc->body.shape != LBM_BODY_NONE;
  const bool walls = desc->walls[0] || desc->walls[1] || (desc->walls[2] && s->dim == 3);
  s->general = s->body || walls || desc->sponge_width > 0.f;
  {
    int flags = 0;
    default_tile(desc->storage, desc->nx, &s->bx, &s->by, &flags);
    s->swz = flags & 1;
    s->nt = (flags >> 1) & 1;
    s->lin = (flags >> 2) & 1;
    s->occ = (flags >> 4) & 15;
  }
  hipError_t e = hipSuccess;
  if (stream) {
    // a pool member: the pool's stream and the batched kernel's tile
    s->stream = stream;
    s->own_stream = false;
    s->bx = bx; s->by = by; s->swz = 0;
    s->lin = (flags >> 2) & 1;
    s->d.autotune = 0;
  } else {
    e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { release(s); return hip_fail(e, "hipStreamCreate"); }
  }
  if ((e = hipEventCreate(&s->ev0)) != hipSuccess || (e = hipEventCreate(&s->ev1)) != hipSuccess) {
    release(s);
    return hip_fail(e, "hipEventCreate");
  }
  const size_t lattice_bytes = lattice_elems(s->q, s->N) * s->elem;
  const int nred = kTotalsBlocks;
  if ((rc = dalloc(s, &s->f[0], lattice_bytes)) || (rc = dalloc(s, &s->f[1], lattice_bytes)) ||
      (rc = dalloc(s, &s->mac, (size_t)s->N * 16)) || (rc = dalloc(s, &s->scratch, (size_t)nred * 4 * 8)) ||
      (rc = dalloc(s, &s->counters, 64)) || (rc = dalloc(s, &s->history, (size_t)kHistoryCap * 6 * 8)) ||
      (rc = dalloc(s, &s->written, 8)) || (rc = dalloc(s, &s->totals, 64)) || (rc = dalloc(s, &s->salt, 8))) {
    release(s);
    return rc;
  }
  if ((e = hipMemsetAsync(s->counters, 0, 64, s->stream)) != hipSuccess ||
      (e = hipMemsetAsync(s->written, 0, 8, s->stream)) != hipSuccess ||
      (e = hipMemsetAsync(s->salt, 0, 8, s->stream)) != hipSuccess) {
    release(s);
    return hip_fail(e, "hipMemsetAsync");
  }
  if (s->body) {
    if ((rc = dalloc(s, &s->cls, s->N)) || (rc = dalloc(s, &s->links, (size_t)s->N * 4)) ||
        (rc = build_boundary(s))) {
      release(s);
      return rc;
    }
  }
  if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) {
    release(s);
    return hip_fail(e, "hipStreamSynchronize");
  }
  *out = s;
  return LBM_OK;
}

}  // namespace

namespace lbm {

static_assert(sizeof(lbm_state_header) == 64, "lbm_state_header is 64 bytes");

uint32_t solve_salt(const lbm_solve* s) { return s->salt0 + (uint32_t)(s->step - s->step_at_salt0); }

int state_header_check(const lbm_desc* d, const void* buf, size_t bytes, lbm_state_header* out) {
  if (!buf || bytes < sizeof(lbm_state_header)) return fail(LBM_E_INVALID, "state of %zu bytes has no header", bytes);
  lbm_state_header h;
  memcpy(&h, buf, sizeof h);
  if (memcmp(h.magic, "LBMSTATE", 8) || h.version != 1) return fail(LBM_E_INVALID, "not a version-1 solve state");
  const int nz = d->q == 9 ? 1 : d->nz;
  if (h.q != d->q || h.nx != d->nx || h.ny != d->ny || h.nz != nz || h.storage != d->storage)
    return fail(LBM_E_INVALID, "state of q%d %dx%dx%d storage %d does not match the solve (q%d %dx%dx%d storage %d)",
                h.q, h.nx, h.ny, h.nz, h.storage, d->q, d->nx, d->ny, nz, d->storage);
  const size_t elem = d->storage == LBM_STORE_F32 ? 4 : 2;
  const size_t payload = (size_t)d->nx * d->ny * nz * d->q * elem;
  if (h.payload_bytes != payload || bytes < sizeof h + payload + h.body_bytes)
    return fail(LBM_E_INVALID, "state payload of %llu bytes (buffer %zu), the solve needs %zu",
                (unsigned long long)h.payload_bytes, bytes, payload);
  if (h.body_bytes != 0 && h.body_bytes != sizeof(RigidBody))
    return fail(LBM_E_INVALID, "rigid body record of %u bytes, expected %zu", h.body_bytes, sizeof(RigidBody));
  if (h.step < 0) return fail(LBM_E_INVALID, "state step %lld", (long long)h.step);
  if (out) *out = h;
  return LBM_OK;
}

size_t solve_footprint(const lbm_desc* d) {
  const size_t n = (size_t)d->nx * d->ny * (d->q == 9 ? 1 : d->nz);
  const size_t elem = d->storage == LBM_STORE_F32 ? 4 : 2;
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  size_t total = 2 * al(lattice_elems(d->q, (uint32_t)n) * elem) + al(n * 16) + al((size_t)kTotalsBlocks * 4 * 8) + al(64) +
                 al((size_t)kHistoryCap * 6 * 8) + al(8) + al(64) + al(8);
  if (d->body.shape != LBM_BODY_NONE) {
    // class bytes, link masks, one slot index per tile (tiles of at least 64
    // cells) and two parity sets of 6 partials per slot (at most one per tile)
    const size_t tiles = (n + 63) / 64 + d->ny * (d->q == 9 ? 1 : d->nz);
    total += al(n) + al(n * 4) + al(tiles * 4) + al(2 * tiles * 6 * 8);
  }
  return total;
}

void default_tile(int storage, int nx, int* bx, int* by, int* flags) {
  // 256-cell tiles with non-temporal access (the streamed lattices stay out
  // of the caches).  Rows that are whole tiles take 2D 256x1 tiles (z and y
  // wave-uniform, fewer per-lane instructions); other widths take linear
  // tiles, which leave no lane idle.  Those two layouts number their tiles
  // alike when nx is a multiple of 256.  fp32 (256 B per wave load) runs
  // best with residency capped at 6 waves per SIMD (104.9 us against 106.8
  // at 4 and 108.4 uncapped), the 16-bit formats want every wave they can
  // get (D3Q19 sphere tile sweep on MI355X, block layout).
  *bx = 256;
  *by = 1;
  *flags = 0x2 | (nx > 0 && nx % 256 == 0 ? 0 : 0x4) | (storage == LBM_STORE_F32 ? 0x60 : 0);
}

int solve_create_in(const lbm_desc* d, Arena* arena, hipStream_t stream, int bx, int by, int flags,
                    lbm_solve** out) {
  if (!arena || !stream) return fail(LBM_E_INVALID, "pool solve needs an arena and a stream");
  return create_impl(d, arena, stream, bx, by, flags, out);
}

int solve_batch_entry(lbm_solve* s, BatchSolve* e, uint32_t* ntiles) {
  const int keep = s->parity;
  for (int p = 0; p < 2; p++) {
    s->parity = p;
    e->args[p] = step_args(s, false, s->body);
    e->f[p] = s->f[p];
  }
  s->parity = keep;
  e->mac = s->mac;
  *ntiles = e->args[0].nblocks;
  return LBM_OK;
}

int solve_advance(lbm_solve* s, int steps, bool finalize_force) {
  s->step += steps;
  s->parity ^= steps & 1;
  if (finalize_force && s->body && s->nslots) {
    // the partials of the last step live in the buffer of the parity it read
    ForceFinalizeArgs fa{};
    fa.partials = s->force_partials + (size_t)s->nslots * 6 * (1 - s->parity);
    fa.nslots = s->nslots;
    fa.history = s->history;
    fa.written = s->written;
    fa.capacity = kHistoryCap;
    HIPCHK(launch_force_finalize(fa, s->stream));
  }
  s->initialised = true;
  return LBM_OK;
}

uint32_t solve_cells(const lbm_solve* s) { return s->N; }

}  // namespace lbm

extern "C" {

const char* lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_abi_version(void) { return LBM_ABI_VERSION; }

int lbm_device_info(char* name, size_t name_len, double* hbm_bytes, int32_t* cu_count) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIPCHK(hipGetDeviceProperties(&p, dev));
  if (name && name_len) snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
  if (hbm_bytes) *hbm_bytes = (double)p.totalGlobalMem;
  if (cu_count) *cu_count = p.multiProcessorCount;
  return LBM_OK;
}


int lbm_solve_create(const lbm_desc* desc, lbm_solve** out) {
  return create_impl(desc, nullptr, nullptr, 0, 0, 0, out);
}

int lbm_solve_destroy(lbm_solve* s) {
  if (s) (void)hipStreamSynchronize(s->stream);
  release(s);
  return LBM_OK;
}

int lbm_upload_macro(lbm_solve* s, const float* mac4) {
  if (!s || !mac4) return fail(LBM_E_INVALID, "null argument");
  HIPCHK(hipMemcpyAsync(s->mac, mac4, (size_t)s->N * 16, hipMemcpyHostToDevice, s->stream));
  HIPCHK(launch_init_eq(s->q, s->storage, s->mac, s->f[0], s->N, s->stream));
  // Both lattices hold the equilibrium (main-3d.js:2994 seeds A and B).
  HIPCHK(hipMemcpyAsync(s->f[1], s->f[0], lattice_elems(s->q, s->N) * s->elem, hipMemcpyDeviceToDevice, s->stream));
  HIPCHK(hipMemsetAsync(s->salt, 0, 8, s->stream));
  s->parity = 0;
  s->step = 0;
  s->salt0 = 0;
  s->step_at_salt0 = 0;
  s->initialised = true;
  int rc = reset_history(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  return prepare_graphs(s);
}

int lbm_upload_f(lbm_solve* s, const float* f) {
  if (!s || !f) return fail(LBM_E_INVALID, "null argument");
  const size_t wire = (size_t)s->N * s->q * 4;
  float* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, wire));
  hipError_t e = hipMemcpyAsync(tmp, f, wire, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) e = launch_pack(s->q, s->storage, tmp, s->f[0], s->N, s->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(s->f[1], s->f[0], lattice_elems(s->q, s->N) * s->elem, hipMemcpyDeviceToDevice, s->stream);
  if (e == hipSuccess) e = hipMemsetAsync(s->salt, 0, 8, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "upload_f");
  s->parity = 0;
  s->step = 0;
  s->salt0 = 0;
  s->step_at_salt0 = 0;
  s->initialised = true;
  int rc = reset_history(s);
  if (rc) return rc;
  return prepare_graphs(s);
}

// Copies between the lattice's block layout and plane-major order:
// to_planes true reads `lat` and writes `planes`, false the reverse.
static void blocks_to_planes(char* lat, char* planes, int q, uint32_t n, int elem, bool to_planes) {
  for (uint32_t b0 = 0; b0 < n; b0 += kLatBlock) {
    const size_t run = (size_t)std::min<uint32_t>(kLatBlock, n - b0) * elem;
    for (int i = 0; i < q; i++) {
      char* l = lat + lat_elem(q, i, b0) * elem;
      char* p = planes + ((size_t)i * n + b0) * elem;
      if (to_planes) memcpy(p, l, run);
      else memcpy(l, p, run);
    }
  }
}

int lbm_state_size(lbm_solve* s, size_t* bytes) {
  if (!s || !bytes) return fail(LBM_E_INVALID, "null argument");
  *bytes = sizeof(lbm_state_header) + (size_t)s->N * s->q * s->elem + (s->rigid ? sizeof(RigidBody) : 0);
  return LBM_OK;
}

int lbm_save_state(lbm_solve* s, void* buf, size_t bytes) {
  if (!s || !buf) return fail(LBM_E_INVALID, "null argument");
  if (!s->initialised) return fail(LBM_E_STATE, "save_state before upload_macro / upload_f / load_state");
  const size_t payload = (size_t)s->N * s->q * s->elem;
  const size_t body = s->rigid ? sizeof(RigidBody) : 0;
  if (bytes < sizeof(lbm_state_header) + payload + body)
    return fail(LBM_E_INVALID, "state buffer of %zu bytes, need %zu", bytes,
                sizeof(lbm_state_header) + payload + body);
  // The payload holds plane-major populations ([Q][cells]); the lattice's
  // blocks are reordered on the host.
  std::vector<char> lat(lattice_elems(s->q, s->N) * s->elem);
  HIPCHK(hipMemcpyAsync(lat.data(), s->f[s->parity], lat.size(), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  blocks_to_planes(lat.data(), static_cast<char*>(buf) + sizeof(lbm_state_header), s->q, s->N, s->elem, true);
  if (body) {
    // the body's record after the populations (its pose matches the lattice)
    HIPCHK(hipMemcpyAsync(static_cast<char*>(buf) + sizeof(lbm_state_header) + payload, s->rigid, body,
                          hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
  }
  lbm_state_header h{};
  memcpy(h.magic, "LBMSTATE", 8);
  h.version = 1;
  h.q = s->q;
  h.nx = s->d.nx; h.ny = s->d.ny; h.nz = s->d.nz;
  h.storage = s->storage;
  h.salt = solve_salt(s);
  h.step = s->step;
  h.payload_bytes = payload;
  h.body_bytes = (uint32_t)body;
  memcpy(buf, &h, sizeof h);
  return LBM_OK;
}

int lbm_load_state(lbm_solve* s, const void* buf, size_t bytes) {
  if (!s || !buf) return fail(LBM_E_INVALID, "null argument");
  lbm_state_header h{};
  int rc = state_header_check(&s->d, buf, bytes, &h);
  if (rc) return rc;
  if ((h.body_bytes != 0) != (s->rigid != nullptr))
    return fail(LBM_E_INVALID, h.body_bytes ? "the state carries a rigid body; attach one before loading it"
                                            : "the solve has a rigid body and the state carries none");
  std::vector<char> lat(lattice_elems(s->q, s->N) * s->elem, 0);
  blocks_to_planes(lat.data(), const_cast<char*>(static_cast<const char*>(buf)) + sizeof(lbm_state_header), s->q, s->N,
                   s->elem, false);
  HIPCHK(hipMemcpyAsync(s->f[0], lat.data(), lat.size(), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipMemcpyAsync(s->f[1], s->f[0], lat.size(), hipMemcpyDeviceToDevice, s->stream));
  const uint32_t words[2] = {h.salt, h.salt};
  HIPCHK(hipMemcpyAsync(s->salt, words, 8, hipMemcpyHostToDevice, s->stream));
  if (s->rigid) {
    // the saved record, then the cells of its pose (everything else cleared)
    HIPCHK(hipMemcpyAsync(s->rigid, static_cast<const char*>(buf) + sizeof(lbm_state_header) + h.payload_bytes,
                          sizeof(RigidBody), hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemsetAsync(s->cls, 0, s->N, s->stream));
    HIPCHK(hipMemsetAsync(s->links, 0, (size_t)s->N * 4, s->stream));
    HIPCHK(launch_rigid_classify(s->q, s->rargs, s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  s->parity = 0;
  s->step = h.step;
  s->salt0 = h.salt;
  s->step_at_salt0 = h.step;
  s->initialised = true;
  s->force_pending = false;
  rc = reset_history(s);
  if (rc) return rc;
  return prepare_graphs(s);
}

int lbm_read_f(lbm_solve* s, float* f) {
  if (!s || !f) return fail(LBM_E_INVALID, "null argument");
  const size_t wire = (size_t)s->N * s->q * 4;
  float* tmp = nullptr;
  HIPCHK(hipMalloc((void**)&tmp, wire));
  hipError_t e = launch_unpack(s->q, s->storage, s->f[s->parity], tmp, s->N, s->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(f, tmp, wire, hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(tmp);
  if (e != hipSuccess) return hip_fail(e, "read_f");
  return LBM_OK;
}

int lbm_step(lbm_solve* s, int32_t n) {
  if (!s) return fail(LBM_E_INVALID, "null solve");
  if (n < 0) return fail(LBM_E_INVALID, "negative step count");
  if (!s->initialised) return fail(LBM_E_STATE, "step before upload_macro / upload_f");
  if (n == 0) return LBM_OK;
  if (s->d.autotune && !s->tuned) {
    int rc = autotune(s);
    if (rc) return rc;
  }
  if (s->graph_bx != s->bx || s->graph_by != s->by || s->graph_swz != s->swz || s->graph_nt != s->nt ||
      s->graph_occ != s->occ || s->graph_lin != s->lin)
    drop_graphs(s);
  // First step eagerly (nothing earlier in this call to finalize), then
  // captured chunks and single steps that each finalize their predecessor's
  // force, the last step writing the macroscopic field, then the last force.
  int rc = enqueue_step(s, n == 1, false);
  if (rc) return rc;
  int left = n - 2;  // steps between the first and the last
  while (left >= kGraphChunk) {
    if (!s->graph[s->parity]) {
      rc = build_graph(s);
      if (rc) return rc;
    }
    HIPCHK(hipGraphLaunch(s->graph[s->parity], s->stream));
    // kGraphChunk is even, so the parity after the chunk is unchanged
    s->step += kGraphChunk;
    s->force_pending = s->body && s->nslots;
    left -= kGraphChunk;
  }
  while (left-- > 0) {
    rc = enqueue_step(s, false, true);
    if (rc) return rc;
  }
  if (n >= 2) {
    rc = enqueue_step(s, true, true);
    if (rc) return rc;
  }
  return flush_force(s);
}

int lbm_step_timed(lbm_solve* s, int32_t n, double* ms) {
  if (!s || !ms) return fail(LBM_E_INVALID, "null argument");
  if (s->d.autotune && !s->tuned) {
    int rc = autotune(s);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(s->ev0, s->stream));
  int rc = lbm_step(s, n);
  if (rc) return rc;
  HIPCHK(hipEventRecord(s->ev1, s->stream));
  HIPCHK(hipEventSynchronize(s->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms = t;
  return LBM_OK;
}

int lbm_sync(lbm_solve* s) {
  if (!s) return fail(LBM_E_INVALID, "null solve");
  HIPCHK(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_macro(lbm_solve* s, float* mac4) {
  if (!s || !mac4) return fail(LBM_E_INVALID, "null argument");
  HIPCHK(hipMemcpyAsync(mac4, s->mac, (size_t)s->N * 16, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_forces(lbm_solve* s, double* out, int32_t max_records, int32_t* count) {
  if (!s || !count || (max_records > 0 && !out)) return fail(LBM_E_INVALID, "null argument");
  unsigned long long written = 0;
  HIPCHK(hipMemcpyAsync(&written, s->written, 8, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  const unsigned long long avail = std::min<unsigned long long>(written, kHistoryCap);
  const unsigned long long take = std::min<unsigned long long>(avail, (unsigned long long)std::max(0, max_records));
  // the `take` most recent records, oldest first
  if (take) {
    std::vector<double> ring((size_t)kHistoryCap * 6);
    HIPCHK(hipMemcpyAsync(ring.data(), s->history, ring.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (unsigned long long k = 0; k < take; k++) {
      const unsigned long long rec = written - take + k;
      memcpy(out + k * 6, ring.data() + (rec % kHistoryCap) * 6, 48);
    }
  }
  *count = (int32_t)take;
  return LBM_OK;
}

int lbm_read_totals(lbm_solve* s, double* out4) {
  if (!s || !out4) return fail(LBM_E_INVALID, "null argument");
  TotalsArgs a{};
  a.ncells = s->N;
  a.partials = s->scratch;
  a.counter = s->counters + 1;
  a.out = s->totals;
  const int nblocks = (int)std::min<uint32_t>(1024u, (s->N + 255) / 256);
  HIPCHK(launch_totals(s->q, s->storage, s->f[s->parity], a, nblocks, s->stream));
  HIPCHK(hipMemcpyAsync(out4, s->totals, 32, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_read_classes(lbm_solve* s, uint8_t* out) {
  if (!s || !out) return fail(LBM_E_INVALID, "null argument");
  if (!s->cls) {
    memset(out, 0, s->N);
    return LBM_OK;
  }
  HIPCHK(hipMemcpyAsync(out, s->cls, s->N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return LBM_OK;
}

int lbm_solve_info(lbm_solve* s, int64_t* step, int32_t* block_x, int32_t* block_y, int32_t* flags,
                   int64_t* device_bytes) {
  if (!s) return fail(LBM_E_INVALID, "null solve");
  if (step) *step = s->step;
  if (block_x) *block_x = s->bx;
  if (block_y) *block_y = s->by;
  if (flags) *flags = s->swz | (s->nt << 1) | (s->lin << 2) | (s->occ << 4);
  if (device_bytes) *device_bytes = (int64_t)s->bytes;
  return LBM_OK;
}

int lbm_time_step_kernel(lbm_solve* s, int32_t n, double* ms_per_launch) {
  if (!s || !ms_per_launch || n < 1) return fail(LBM_E_INVALID, "bad argument");
  if (!s->initialised) return fail(LBM_E_STATE, "time_step_kernel before upload");
  if (s->d.autotune && !s->tuned) {
    int rc = autotune(s);
    if (rc) return rc;
  }
  const void* fin = s->f[s->parity];
  void* fout = s->f[1 - s->parity];
  StepArgs a = step_args(s, false, true);
  HIPCHK(launch_step(s->q, s->storage, s->general, s->nt, s->occ, fin, fout, a, s->bx, s->by, s->stream,
                     s->d.diffuse != 0));
  HIPCHK(hipEventRecord(s->ev0, s->stream));
  for (int r = 0; r < n; r++)
    HIPCHK(launch_step(s->q, s->storage, s->general, s->nt, s->occ, fin, fout, a, s->bx, s->by, s->stream,
                     s->d.diffuse != 0));
  HIPCHK(hipEventRecord(s->ev1, s->stream));
  HIPCHK(hipEventSynchronize(s->ev1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, s->ev0, s->ev1));
  *ms_per_launch = (double)t / n;
  return LBM_OK;
}

int lbm_set_tile(lbm_solve* s, int32_t block_x, int32_t block_y, int32_t flags) {
  if (!s) return fail(LBM_E_INVALID, "null solve");
  if (block_x < 1 || block_y < 1 || block_x * block_y > 256 || (block_x * block_y) % 64)
    return fail(LBM_E_INVALID, "tile %dx%d: need a multiple of 64 threads, at most 256", block_x, block_y);
  const int occ = (flags >> 4) & 15;
  if (!occupancy_cap_ok(occ, block_x * block_y))
    return fail(LBM_E_INVALID, "occupancy cap %d waves per SIMD cannot be set for %d-thread blocks", occ,
                block_x * block_y);
  s->bx = block_x;
  s->by = block_y;
  s->swz = flags & 1;
  s->nt = (flags >> 1) & 1;
  s->lin = (flags >> 2) & 1;
  s->occ = occ;
  s->tuned = true;
  int rc = build_force_slots(s);
  if (rc) return rc;
  return prepare_graphs(s);
}

int lbm_probe_copy_bandwidth(size_t bytes, int32_t reps, double* gbs, double* ms_per_copy) {
  if (bytes < 16 || reps < 1 || !gbs) return fail(LBM_E_INVALID, "bad argument");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, st);
  if (e == hipSuccess) e = launch_copy_probe(a, b, bytes, st);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int r = 0; r < reps && e == hipSuccess; r++) e = launch_copy_probe(a, b, bytes, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  if (e1) (void)hipEventDestroy(e1);
  if (e0) (void)hipEventDestroy(e0);
  if (st) (void)hipStreamDestroy(st);
  if (b) (void)hipFree(b);
  if (a) (void)hipFree(a);
  if (e != hipSuccess) return hip_fail(e, "copy probe");
  const double ms = (double)t / reps;
  *gbs = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
  if (ms_per_copy) *ms_per_copy = ms;
  return LBM_OK;
}

// --- moving rigid body --------------------------------------------------

namespace {

// Quaternion helpers in the FP64 operation order of the host restatement
// (d3-body.mjs:161-218), for the attach-time angular momentum.
void q_mul_h(const double p[4], const double q[4], double o[4]) {
  o[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
  o[1] = p[0] * q[1] + p[1] * q[0] + p[2] * q[3] - p[3] * q[2];
  o[2] = p[0] * q[2] - p[1] * q[3] + p[2] * q[0] + p[3] * q[1];
  o[3] = p[0] * q[3] + p[1] * q[2] - p[2] * q[1] + p[3] * q[0];
}
void q_rotate_h(const double q[4], const double v[3], bool inverse, double o[3]) {
  const double qc[4] = {q[0], -q[1], -q[2], -q[3]};
  const double vq[4] = {0.0, v[0], v[1], v[2]};
  double t[4], u[4];
  if (inverse) { q_mul_h(qc, vq, t); q_mul_h(t, q, u); }
  else { q_mul_h(q, vq, t); q_mul_h(t, qc, u); }
  o[0] = u[1]; o[1] = u[2]; o[2] = u[3];
}

}  // namespace

int lbm_rigid_attach(lbm_solve* s, const lbm_rigid_body* b) {
  if (!s || !b) return fail(LBM_E_INVALID, "null argument");
  if (s->dim != 3) return fail(LBM_E_INVALID, "a rigid body needs a 3D solve (q 19 or 27)");
  if (s->body) return fail(LBM_E_INVALID, "the solve already has a pinned body in its descriptor");
  if (s->arena) return fail(LBM_E_INVALID, "rigid bodies attach to standalone solves only");
  const double a = b->a, bb = b->shape == LBM_SHAPE_ROUNDBOX ? b->b : b->a,
               c = b->shape == LBM_SHAPE_SPHERE ? b->a : b->c, r = b->shape == LBM_SHAPE_ROUNDBOX ? b->r : 0.0;
  if (b->shape < LBM_SHAPE_SPHERE || b->shape > LBM_SHAPE_ROUNDBOX) return fail(LBM_E_INVALID, "shape %d", b->shape);
  if (!(a > 0.0) || !(bb > 0.0) || !(c > 0.0) || !(r >= 0.0) || r > std::min(a, std::min(bb, c)))
    return fail(LBM_E_INVALID, "body dimensions a %g b %g c %g r %g", a, bb, c, r);
  if (!(b->density > 0.0)) return fail(LBM_E_INVALID, "density %g", b->density);
  if (!(b->v_max >= 0.0) || !(b->o_max >= 0.0)) return fail(LBM_E_INVALID, "clamps v_max %g o_max %g", b->v_max, b->o_max);
  const double qn = std::sqrt(b->q[0] * b->q[0] + b->q[1] * b->q[1] + b->q[2] * b->q[2] + b->q[3] * b->q[3]);
  if (!(std::fabs(qn - 1.0) < 1e-9)) return fail(LBM_E_INVALID, "orientation quaternion of norm %.17g", qn);
  const double n[3] = {(double)s->d.nx, (double)s->d.ny, (double)s->d.nz};
  for (int k = 0; k < 3; k++)
    if (!std::isfinite(b->x[k]) || !std::isfinite(b->v[k]) || !std::isfinite(b->omega[k]))
      return fail(LBM_E_INVALID, "non-finite body state");

  RigidBody h{};
  h.shape = b->shape;
  h.a = a; h.b = bb; h.cz = c; h.r = r;
  const double pi = 3.14159265358979323846;
  double vol;
  if (b->shape == LBM_SHAPE_SPHERE) vol = (4.0 / 3.0) * pi * (a * a * a);
  else if (b->shape == LBM_SHAPE_SPHEROID) vol = (4.0 / 3.0) * pi * a * a * c;
  else vol = 8.0 * a * bb * c;
  h.mass = b->density * vol;
  const double m = h.mass;
  if (b->shape == LBM_SHAPE_SPHERE) { h.I[0] = h.I[1] = h.I[2] = 0.4 * m * a * a; }
  else if (b->shape == LBM_SHAPE_SPHEROID) { h.I[0] = h.I[1] = m * (a * a + c * c) / 5.0; h.I[2] = 2.0 * m * a * a / 5.0; }
  else {
    h.I[0] = m * (bb * bb + c * c) / 3.0;
    h.I[1] = m * (a * a + c * c) / 3.0;
    h.I[2] = m * (a * a + bb * bb) / 3.0;
  }
  h.rc = b->shape == LBM_SHAPE_SPHERE ? a : (b->shape == LBM_SHAPE_SPHEROID ? std::max(a, c) : std::sqrt(a * a + bb * bb + c * c));
  for (int k = 0; k < 4; k++) h.q[k] = b->q[k];
  double Lb[3], Lbody[3];
  q_rotate_h(h.q, b->omega, true, Lb);
  for (int k = 0; k < 3; k++) Lbody[k] = Lb[k] * h.I[k];
  q_rotate_h(h.q, Lbody, false, h.L);
  for (int k = 0; k < 3; k++) {
    h.c[k] = b->x[k];
    h.v[k] = b->v[k];
    h.w[k] = b->omega[k];
    h.g[k] = b->gravity[k];
    h.win_n[k] = b->window[k] ? n[k] : 0.0;
    h.win_a[k] = b->window_anchor[k];
  }
  h.v_max = b->v_max;
  h.o_max = b->o_max;
  h.pinned = b->pinned ? 1 : 0;
  h.planar = b->planar ? 1 : 0;
  h.no_fluid = b->no_fluid_force ? 1 : 0;
  h.swept = b->swept_force ? 1 : 0;
  h.rho_w_local = b->rho_w_local ? 1 : 0;

  RigidArgs& ra = s->rargs;
  ra = RigidArgs{};
  ra.nx = s->d.nx; ra.ny = s->d.ny; ra.nz = s->d.nz;
  ra.half = (int)std::ceil(h.rc) + 3;
  const int dims[3] = {s->d.nx, s->d.ny, s->d.nz};
  const int move = (int)std::ceil(std::min(h.v_max, (double)std::max(dims[0], std::max(dims[1], dims[2])))) + 1;
  for (int k = 0; k < 3; k++) {
    ra.fbox[k] = std::min(2 * ra.half + 2, dims[k]);
    ra.cbox[k] = std::min(2 * ra.half + 2 + move, dims[k]);
  }
  int rc = LBM_OK;
  if (!s->cls && (rc = dalloc(s, &s->cls, s->N))) return rc;
  if (!s->links && (rc = dalloc(s, &s->links, (size_t)s->N * 4))) return rc;
  if (!s->rigid) {
    HIPCHK(hipMalloc((void**)&s->rigid, sizeof(RigidBody)));
    HIPCHK(hipMalloc((void**)&s->rigid_counter, 16));
  }
  if (s->rigid_partials) {
    (void)hipFree(s->rigid_partials);
    s->rigid_partials = nullptr;
  }
  ra.body = s->rigid;
  ra.cls = s->cls;
  ra.links = s->links;
  ra.counter = s->rigid_counter;
  ra.history = s->history;
  ra.written = s->written;
  ra.hist_cap = kHistoryCap;
  HIPCHK(hipMalloc((void**)&s->rigid_partials, (size_t)rigid_force_blocks(ra) * 6 * 8));
  ra.partials = s->rigid_partials;
  HIPCHK(hipMemsetAsync(s->cls, 0, s->N, s->stream));
  HIPCHK(hipMemsetAsync(s->links, 0, (size_t)s->N * 4, s->stream));
  HIPCHK(hipMemsetAsync(s->rigid_counter, 0, 16, s->stream));
  HIPCHK(hipMemcpyAsync(s->rigid, &h, sizeof h, hipMemcpyHostToDevice, s->stream));
  HIPCHK(launch_rigid_classify(s->q, ra, s->stream));
  rc = reset_history(s);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(s->stream));
  if (!s->general) {
    s->general = true;
    s->tuned = false;
  }
  drop_graphs(s);
  return prepare_graphs(s);
}

int lbm_rigid_read(lbm_solve* s, lbm_rigid_state* out) {
  if (!s || !out) return fail(LBM_E_INVALID, "null argument");
  if (!s->rigid) return fail(LBM_E_STATE, "no rigid body attached");
  RigidBody h{};
  HIPCHK(hipMemcpyAsync(&h, s->rigid, sizeof h, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int k = 0; k < 3; k++) {
    out->x[k] = h.c[k]; out->v[k] = h.v[k]; out->L[k] = h.L[k]; out->omega[k] = h.w[k]; out->d[k] = h.d[k];
    out->force[k] = h.force[k]; out->torque[k] = h.torque[k]; out->inertia[k] = h.I[k];
  }
  for (int k = 0; k < 4; k++) out->q[k] = h.q[k];
  out->mass = h.mass;
  return LBM_OK;
}

}  // extern "C"

// Batched serving: the lbm_serve_* entry points of include/lbm.h.
//
// Many independent dense solves live in one packed device pool (one
// allocation of the byte budget, first-fit with coalescing) and advance
// together.  A host worker thread owns all device work:
//
//   loop:  launch the chunk of the resident set on the compute stream:
//            per lattice class (Q, storage, general) a device solve table,
//            a tile table and a captured graph of K launches, launch j
//            stepping every member with j < remaining (k_step_batch)
//          while it runs, on the setup stream: copy back the members that
//            finished in the previous chunk (last force record, macroscopic
//            field, force history), free their pool memory, and admit
//            queued requests (FIFO) that fit the byte budget, until the
//            resident cells reach the target that fills the device
//          wait for the chunk; members that reached their step count are
//            set aside for copy-back during the next chunk
//
// Requests that arrive while others run join at the next chunk boundary;
// host set-up and copy-back overlap device work.
// A member's arithmetic depends only on its own data, its class's tile shape
// and its own step count (the 16-bit dither is seeded by it), so its result
// is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lbm.h"
#include "dense.h"
#include "engine.h"

using namespace lbm;

#define HIPCHK_RET(call)                             \
  do {                                               \
    hipError_t e_ = (call);                          \
    if (e_ != hipSuccess) return hip_fail(e_, #call); \
  } while (0)

namespace {

// First-fit pool over one device allocation.
class Pool : public Arena {
 public:
  int init(size_t bytes) {
    hipError_t e = hipMalloc((void**)&base_, bytes);
    if (e != hipSuccess) return fail(LBM_E_NOMEM, "pool hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    size_ = bytes;
    free_[0] = bytes;
    return LBM_OK;
  }
  ~Pool() override {
    if (base_) (void)hipFree(base_);
  }
  void* alloc(size_t bytes) override {
    std::lock_guard<std::mutex> g(mu_);
    bytes = (bytes + 255) & ~(size_t)255;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      used_[off] = bytes;
      in_use_ += bytes;
      peak_ = std::max(peak_, in_use_);
      return base_ + off;
    }
    return nullptr;
  }
  void release(void* p) override {
    std::lock_guard<std::mutex> g(mu_);
    const size_t off = static_cast<char*>(p) - base_;
    auto u = used_.find(off);
    if (u == used_.end()) return;
    size_t start = off, len = u->second;
    in_use_ -= len;
    used_.erase(u);
    auto next = free_.lower_bound(start);
    if (next != free_.end() && next->first == start + len) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == start) {
        start = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[start] = len;
  }
  size_t size() const { return size_; }
  size_t in_use() const { return in_use_; }
  size_t peak() const { return peak_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0, in_use_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;
  std::unordered_map<size_t, size_t> used_;
  std::mutex mu_;
};

struct Request {
  int64_t ticket = 0;
  lbm_desc d{};
  std::vector<float> mac_in;
  int steps = 0;
  size_t footprint = 0;
  uint32_t cells = 0;
  double t_submit = 0, t_admit = 0, t_finish = 0;
  int status = LBM_OK;
  std::string error;
  bool done = false;
  std::vector<float> mac_out;
  std::vector<double> forces;  // 6 per recorded step
  std::vector<uint8_t> state_in;   // start from this saved state instead of mac_in
  bool keep_state = false;
  std::vector<uint8_t> state_out;  // the final state (keep_state)
};

struct Member {
  std::shared_ptr<Request> r;
  lbm_solve* s = nullptr;
  BatchSolve entry{};
  uint32_t ntiles = 0;
  int done_steps = 0;
  bool pending = false;   // force partials of the last step not yet summed
  bool in_chunk = false;  // part of the chunk in flight (members admitted
                          // while a chunk runs join the next one)
};

struct ClassKey {
  int q, storage, general;
  int rows;   // nx is a multiple of the layout block (row-mode kernels)
  int force;  // a nonzero body force (the kernels with Guo terms)
  bool operator<(const ClassKey& o) const {
    return std::tie(q, storage, general, rows, force) < std::tie(o.q, o.storage, o.general, o.rows, o.force);
  }
};

struct Class {
  ClassKey key{};
  int bx = 64, by = 1, nt = 0, occ = 0, lin = 0;
  std::vector<Member*> members;
  bool dirty = true;  // membership changed since the tables were built
  BatchSolve* d_solves = nullptr;
  BatchTile* d_tiles = nullptr;
  size_t cap_solves = 0, cap_tiles = 0;
  uint32_t ntiles = 0;
  PersistBarrier* d_bar = nullptr;  // the persistent chunk's grid barrier words
  // host sources of the tables' async uploads: they must outlive the copies,
  // so they live here until the next chunk (each chunk ends with a sync)
  std::vector<BatchSolve> h_solves;
  std::vector<BatchTile> h_tiles;
};

double now_s(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct lbm_serve {
  lbm_serve_desc d{};
  int device = 0;
  std::chrono::steady_clock::time_point t0;
  Pool pool;
  hipStream_t stream = nullptr;
  hipStream_t setup = nullptr;       // admissions and copy-back
  hipEvent_t setup_done = nullptr;   // recorded after each admission round
  hipEvent_t chunk_ev[2] = {nullptr, nullptr};  // around each chunk's graph launch
  double chunk_bytes_pending = 0;    // algorithmic bytes of the chunk in flight
  std::vector<std::unique_ptr<Member>> finishing;  // done, results not yet copied back
  std::vector<int> finishing_adv;
  double resident_cells = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> requests;
  std::vector<std::unique_ptr<Member>> resident;
  std::map<ClassKey, std::unique_ptr<Class>> classes;
  int64_t next_ticket = 1;
  bool stop = false;
  std::thread worker;
  // graph of the current chunk (all classes); re-captured when any class
  // changes, then updated in place or, when its shape changed, instantiated
  hipGraphExec_t graph = nullptr;
  int64_t graph_instantiations = 0, graph_updates = 0;
  // device time of the chunk graphs (HIP events on the compute stream) and
  // the algorithmic bytes they moved: 2 Q s per cell update
  double device_seconds = 0, device_bytes = 0;
  // stats
  int64_t submitted = 0, completed = 0, failed = 0;
  double cell_updates = 0, first_submit = -1, last_finish = 0;
  std::vector<double> latencies;
  int peak_resident = 0, chunks = 0;
  size_t resident_bytes = 0, peak_resident_bytes = 0;
};

namespace {

void finish(lbm_serve* sv, const std::shared_ptr<Request>& r, int status, const std::string& err) {
  std::lock_guard<std::mutex> g(sv->mu);
  r->status = status;
  r->error = err;
  r->done = true;
  r->t_finish = now_s(sv->t0);
  if (status == LBM_OK) {
    sv->completed++;
    sv->cell_updates += (double)r->cells * r->steps;
    sv->latencies.push_back(r->t_finish - r->t_submit);
    sv->last_finish = std::max(sv->last_finish, r->t_finish);
  } else {
    sv->failed++;
  }
  sv->cv_done.notify_all();
}

// Tile shape of a lattice class: the dense autotuner run once on the class's
// first request (its own grid) when that grid is large enough for its
// timings to speak for the class's big members (2^20 cells; a small grid is
// launch-bound and would pick by noise), else the plain solve's default
// tile.  The class's members have other widths, so its tiles are linear
// (blocks of consecutive cells, no idle lanes at any width) with the tuned
// thread count, load flavour and occupancy cap.
void choose_tile(lbm_serve* sv, Class* c, const Request& r) {
  if (!sv->d.autotune || r.cells < (1u << 20)) return;
  lbm_desc d = r.d;
  d.autotune = 1;
  lbm_solve* t = nullptr;
  if (lbm_solve_create(&d, &t)) return;
  int64_t step = 0, bytes = 0;
  int32_t bx = 64, by = 1, flags = 0;
  if (!lbm_upload_macro(t, r.mac_in.data()) && !lbm_step(t, 1) &&
      !lbm_solve_info(t, &step, &bx, &by, &flags, &bytes)) {
    c->bx = bx * by;
    c->by = 1;
    c->nt = (flags >> 1) & 1;
    c->occ = (flags >> 4) & 15;
    c->lin = 1;
  }
  lbm_solve_destroy(t);
}

// Admits queued requests that fit, in arrival order.
void admit(lbm_serve* sv) {
  for (;;) {
    std::shared_ptr<Request> r;
    {
      std::lock_guard<std::mutex> g(sv->mu);
      if (sv->queue.empty()) return;
      if (sv->d.max_resident > 0 && (int)sv->resident.size() >= sv->d.max_resident) return;
      if (sv->d.target_cells > 0 && sv->resident_cells >= sv->d.target_cells) return;
      if (sv->resident_bytes + sv->queue.front()->footprint > sv->pool.size()) return;
      r = sv->queue.front();
      sv->queue.pop_front();
    }
    const ClassKey key{r->d.q, r->d.storage,
                       (r->d.body.shape != LBM_BODY_NONE || r->d.walls[0] || r->d.walls[1] ||
                        (r->d.walls[2] && r->d.q != 9) || r->d.sponge_width > 0.f)
                           ? 1
                           : 0,
                       r->d.nx % (int)kLatBlock == 0 ? 1 : 0,
                       (r->d.force[0] != 0.f || r->d.force[1] != 0.f || r->d.force[2] != 0.f) ? 1 : 0};
    auto& cp = sv->classes[key];
    if (!cp) {
      cp = std::make_unique<Class>();
      cp->key = key;
      int flags = 0;
      default_tile(key.storage, 0, &cp->bx, &cp->by, &flags);
      cp->nt = (flags >> 1) & 1;
      cp->lin = (flags >> 2) & 1;
      cp->occ = (flags >> 4) & 15;
      choose_tile(sv, cp.get(), *r);
    }
    auto m = std::make_unique<Member>();
    m->r = r;
    int rc = solve_create_in(&r->d, &sv->pool, sv->setup, cp->bx, cp->by, cp->lin << 2, &m->s);
    if (rc == LBM_E_NOMEM) {
      // fragmentation: wait for a retirement and try again
      std::lock_guard<std::mutex> g(sv->mu);
      sv->queue.push_front(r);
      return;
    }
    if (!rc)
      rc = r->state_in.empty() ? lbm_upload_macro(m->s, r->mac_in.data())
                               : lbm_load_state(m->s, r->state_in.data(), r->state_in.size());
    if (!rc) rc = solve_batch_entry(m->s, &m->entry, &m->ntiles);
    if (rc) {
      const std::string err = lbm_last_error();
      if (m->s) lbm_solve_destroy(m->s);
      finish(sv, r, rc, err);
      continue;
    }
    r->mac_in.clear();
    r->mac_in.shrink_to_fit();
    r->state_in.clear();
    r->state_in.shrink_to_fit();
    std::lock_guard<std::mutex> g(sv->mu);
    r->t_admit = now_s(sv->t0);
    sv->resident_bytes += r->footprint;
    sv->resident_cells += r->cells;
    sv->peak_resident_bytes = std::max(sv->peak_resident_bytes, sv->resident_bytes);
    cp->members.push_back(m.get());
    cp->dirty = true;
    sv->resident.push_back(std::move(m));
    sv->peak_resident = std::max(sv->peak_resident, (int)sv->resident.size());
  }
}

template <class T>
int grow(T** p, size_t* cap, size_t n) {
  if (n <= *cap) return LBM_OK;
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  const size_t c = std::max(n, *cap * 2);
  hipError_t e = hipMalloc((void**)p, c * sizeof(T));
  if (e != hipSuccess) return hip_fail(e, "serve table");
  *cap = c;
  return LBM_OK;
}

// Uploads every class's solve table for this chunk and, when membership
// changed, its tile table.  Returns whether the chunk graph must be
// re-captured.  The executable graph holds the tables' device addresses; it
// is not launched again before launch_chunk has updated it with the new
// capture, so tables may be reallocated under it (the stream is idle here:
// complete_chunk synchronised it).
int prepare(lbm_serve* sv, bool* rebuild) {
  *rebuild = sv->graph == nullptr;
  for (auto& kv : sv->classes) {
    Class* c = kv.second.get();
    if (c->dirty) {
      *rebuild = true;
      std::vector<BatchTile>& tiles = c->h_tiles;
      tiles.clear();
      for (uint32_t k = 0; k < c->members.size(); k++)
        for (uint32_t t = 0; t < c->members[k]->ntiles; t++) tiles.push_back(BatchTile{k, t});
      c->ntiles = (uint32_t)tiles.size();
      int rc = grow(&c->d_tiles, &c->cap_tiles, std::max<size_t>(tiles.size(), 1));
      if (rc) return rc;
      if (!tiles.empty())
        HIPCHK_RET(hipMemcpyAsync(c->d_tiles, tiles.data(), tiles.size() * sizeof(BatchTile), hipMemcpyHostToDevice,
                                  sv->stream));
      rc = grow(&c->d_solves, &c->cap_solves, std::max<size_t>(c->members.size(), 1));
      if (rc) return rc;
      c->dirty = false;
    }
    c->h_solves.resize(c->members.size());
    for (size_t k = 0; k < c->members.size(); k++) {
      Member* m = c->members[k];
      BatchSolve e = m->entry;
      e.parity = m->done_steps & 1;
      e.remaining = m->r->steps - m->done_steps;
      e.pending = m->pending ? 1 : 0;
      e.step0 = solve_salt(m->s);  // the dither counter of the solve's next step
      c->h_solves[k] = e;
    }
    if (!c->h_solves.empty())
      HIPCHK_RET(hipMemcpyAsync(c->d_solves, c->h_solves.data(), c->h_solves.size() * sizeof(BatchSolve),
                                hipMemcpyHostToDevice, sv->stream));
  }
  return LBM_OK;
}

int enqueue_chunk(lbm_serve* sv, int K) {
  for (int j = 0; j < K; j++) {
    for (auto& kv : sv->classes) {
      Class* c = kv.second.get();
      if (c->members.empty()) continue;
      BatchArgs b{c->d_solves, c->d_tiles, j};
      const bool rows = c->key.rows && step_row_mode((int)kLatBlock, c->lin != 0, c->bx, c->by);
      HIPCHK_RET(launch_step_batch(c->key.q, c->key.storage, c->key.general != 0, c->nt != 0, c->occ, rows,
                                   c->key.force != 0, b, c->ntiles, c->bx, c->by, sv->stream));
    }
  }
  return LBM_OK;
}

// Algorithmic bytes of the chunk about to run: every member steps
// min(remaining, K) times.
void set_chunk_bytes(lbm_serve* sv, int K) {
  std::lock_guard<std::mutex> g(sv->mu);
  double bytes = 0;
  for (auto& m : sv->resident) {
    const int n = std::min(m->r->steps - m->done_steps, K);
    const double elem = m->r->d.storage == LBM_STORE_F32 ? 4.0 : 2.0;
    bytes += (double)n * m->r->cells * 2.0 * m->r->d.q * elem;
  }
  sv->chunk_bytes_pending = bytes;
}

// Launches the chunk of the current resident set on the compute stream
// (after the set-up stream's last admissions) and returns without waiting.
int launch_chunk(lbm_serve* sv) {
  const int K = std::max(1, sv->d.chunk_steps);
  {
    std::lock_guard<std::mutex> g(sv->mu);
    for (auto& m : sv->resident) m->in_chunk = true;
  }
  HIPCHK_RET(hipStreamWaitEvent(sv->stream, sv->setup_done, 0));
  bool rebuild = false;
  int rc = prepare(sv, &rebuild);
  if (rc) return rc;
  if (sv->d.persistent) {
    // one cooperative launch per lattice class runs the whole chunk
    set_chunk_bytes(sv, K);
    HIPCHK_RET(hipEventRecord(sv->chunk_ev[0], sv->stream));
    for (auto& kv : sv->classes) {
      Class* c = kv.second.get();
      if (c->members.empty()) continue;
      BatchArgs b{c->d_solves, c->d_tiles, 0};
      const bool rows = c->key.rows && step_row_mode((int)kLatBlock, c->lin != 0, c->bx, c->by);
      if (!c->d_bar) HIPCHK_RET(hipMalloc((void**)&c->d_bar, sizeof(PersistBarrier)));
      HIPCHK_RET(launch_step_persistent(c->key.q, c->key.storage, c->key.general != 0, c->nt != 0, c->occ, rows,
                                        c->key.force != 0, b, c->ntiles, c->bx, c->by, K, c->d_bar, sv->stream));
    }
    HIPCHK_RET(hipEventRecord(sv->chunk_ev[1], sv->stream));
    return LBM_OK;
  }
  if (rebuild) {
    // Re-capture the chunk; when only launch shapes and table addresses
    // changed (the same classes, so the same kernels in the same order), the
    // executable graph is updated in place, which costs far less host time
    // than instantiating a new one.
    hipGraph_t g = nullptr;
    HIPCHK_RET(hipStreamBeginCapture(sv->stream, hipStreamCaptureModeThreadLocal));
    rc = enqueue_chunk(sv, K);
    hipError_t e = hipStreamEndCapture(sv->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "chunk capture");
    bool updated = false;
    if (sv->graph) {
      hipGraphNode_t bad = nullptr;
      hipGraphExecUpdateResult res = hipGraphExecUpdateSuccess;
      updated = hipGraphExecUpdate(sv->graph, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
      (void)hipGetLastError();
      if (!updated) {
        (void)hipGraphExecDestroy(sv->graph);
        sv->graph = nullptr;
      }
    }
    if (!updated) {
      e = hipGraphInstantiate(&sv->graph, g, nullptr, nullptr, 0);
      if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return hip_fail(e, "chunk instantiate");
      }
      sv->graph_instantiations++;
    } else {
      sv->graph_updates++;
    }
    (void)hipGraphDestroy(g);
  }
  set_chunk_bytes(sv, K);
  HIPCHK_RET(hipEventRecord(sv->chunk_ev[0], sv->stream));
  HIPCHK_RET(hipGraphLaunch(sv->graph, sv->stream));
  HIPCHK_RET(hipEventRecord(sv->chunk_ev[1], sv->stream));
  return LBM_OK;
}

// After the chunk: advance every member that ran in it by its steps (those
// admitted during the chunk have not moved); members that reached their step
// count leave the resident set for copy-back.
int complete_chunk(lbm_serve* sv) {
  const int K = std::max(1, sv->d.chunk_steps);
  HIPCHK_RET(hipStreamSynchronize(sv->stream));
  if (sv->d.persistent) {
    // a persistent launch whose grid barrier ran out stopped early
    for (auto& kv : sv->classes) {
      Class* c = kv.second.get();
      if (c->members.empty() || !c->d_bar) continue;
      uint32_t arrived = 0;
      HIPCHK_RET(hipMemcpy(&arrived, &c->d_bar->arrived, sizeof arrived, hipMemcpyDeviceToHost));
      if (arrived >> 31)
        return fail(LBM_E_DEVICE, "persistent chunk: a grid barrier wait ran out (blocks not co-resident)");
    }
  }
  float ms = 0.f;
  HIPCHK_RET(hipEventElapsedTime(&ms, sv->chunk_ev[0], sv->chunk_ev[1]));
  std::lock_guard<std::mutex> g(sv->mu);
  sv->chunks++;
  sv->device_seconds += ms * 1e-3;
  sv->device_bytes += sv->chunk_bytes_pending;
  std::vector<std::unique_ptr<Member>> keep;
  for (auto& m : sv->resident) {
    if (!m->in_chunk) {
      keep.push_back(std::move(m));
      continue;
    }
    const int adv = std::min(m->r->steps - m->done_steps, K);
    m->done_steps += adv;
    if (m->done_steps >= m->r->steps) {
      sv->resident_cells -= m->r->cells;
      sv->finishing_adv.push_back(adv);
      sv->finishing.push_back(std::move(m));
    } else {
      solve_advance(m->s, adv, false);
      m->pending = m->entry.args[0].force_slot != nullptr;
      keep.push_back(std::move(m));
    }
  }
  sv->resident.swap(keep);
  for (auto& kv : sv->classes) {
    Class* c = kv.second.get();
    std::vector<Member*> mem;
    for (Member* m : c->members)
      if (m->done_steps < m->r->steps) mem.push_back(m);
    if (mem.size() != c->members.size()) {
      c->members.swap(mem);
      c->dirty = true;
    }
  }
  return LBM_OK;
}

// Copies back the results of the members set aside by complete_chunk (on the
// set-up stream) and frees their pool memory.
void retire(lbm_serve* sv) {
  std::vector<std::unique_ptr<Member>> done;
  std::vector<int> adv;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    done.swap(sv->finishing);
    adv.swap(sv->finishing_adv);
  }
  for (size_t k = 0; k < done.size(); k++) {
    Member& m = *done[k];
    Request& r = *m.r;
    int rc = solve_advance(m.s, adv[k], true);
    if (!rc) {
      r.mac_out.resize((size_t)r.cells * 4);
      rc = lbm_read_macro(m.s, r.mac_out.data());
    }
    if (!rc && r.keep_state) {
      size_t n = 0;
      rc = lbm_state_size(m.s, &n);
      if (!rc) {
        r.state_out.resize(n);
        rc = lbm_save_state(m.s, r.state_out.data(), n);
      }
    }
    if (!rc && m.entry.args[0].force_slot) {
      const int maxr = std::min(r.steps, 1 << 16);
      r.forces.resize((size_t)maxr * 6);
      int32_t cnt = 0;
      rc = lbm_read_forces(m.s, r.forces.data(), maxr, &cnt);
      r.forces.resize((size_t)cnt * 6);
    }
    const std::string err = rc ? lbm_last_error() : "";
    lbm_solve_destroy(m.s);
    {
      std::lock_guard<std::mutex> g(sv->mu);
      sv->resident_bytes -= r.footprint;
    }
    finish(sv, m.r, rc, err);
  }
}

void fail_resident(lbm_serve* sv, int rc) {
  const std::string err = lbm_last_error();
  std::vector<std::unique_ptr<Member>> all;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    all.swap(sv->resident);
    for (auto& m : sv->finishing) all.push_back(std::move(m));
    sv->finishing.clear();
    sv->finishing_adv.clear();
    for (auto& kv : sv->classes) {
      kv.second->members.clear();
      kv.second->dirty = true;
    }
    sv->resident_bytes = 0;
    sv->resident_cells = 0;
  }
  for (auto& m : all) {
    lbm_solve_destroy(m->s);
    finish(sv, m->r, rc, err);
  }
}

void worker(lbm_serve* sv) {
  (void)hipSetDevice(sv->device);
  for (;;) {
    bool running = false;
    {
      std::lock_guard<std::mutex> g(sv->mu);
      running = !sv->resident.empty();
    }
    int rc = LBM_OK;
    if (running) rc = launch_chunk(sv);
    // host work and copies overlap the chunk
    retire(sv);
    admit(sv);
    (void)hipEventRecord(sv->setup_done, sv->setup);
    if (!rc && running) rc = complete_chunk(sv);
    if (rc) {
      fail_resident(sv, rc);
      continue;
    }
    if (!running) {
      std::unique_lock<std::mutex> lk(sv->mu);
      if (!sv->resident.empty() || !sv->finishing.empty()) continue;
      if (sv->stop) break;
      if (sv->queue.empty()) sv->cv_work.wait(lk, [&] { return sv->stop || !sv->queue.empty(); });
    }
  }
  // requests still queued when stopping
  std::deque<std::shared_ptr<Request>> left;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    left.swap(sv->queue);
  }
  for (auto& r : left) finish(sv, r, LBM_E_STATE, "server stopped before the request ran");
}

std::shared_ptr<Request> find(lbm_serve* sv, int64_t ticket) {
  auto it = sv->requests.find(ticket);
  return it == sv->requests.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

int lbm_serve_create(const lbm_serve_desc* desc, lbm_serve** out) {
  if (!out || !desc) return fail(LBM_E_INVALID, "null argument");
  *out = nullptr;
  if (desc->abi_version != LBM_ABI_VERSION) return fail(LBM_E_INVALID, "descriptor abi_version %d", desc->abi_version);
  if (!(desc->budget_bytes >= 1 << 20)) return fail(LBM_E_INVALID, "budget_bytes %g below 1 MiB", desc->budget_bytes);
  if (desc->max_resident < 0 || desc->chunk_steps < 0 || desc->target_cells < 0)
    return fail(LBM_E_INVALID, "negative max_resident / chunk_steps / target_cells");
  if (desc->persistent != 0 && desc->persistent != 1) return fail(LBM_E_INVALID, "persistent %d: 0 or 1", desc->persistent);
  if (desc->persistent) {
    int dev = 0, coop = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev);
    if (e != hipSuccess) return hip_fail(e, "cooperative launch attribute");
    if (!coop) return fail(LBM_E_DEVICE, "persistent: the device has no cooperative launch");
  }
  auto sv = std::make_unique<lbm_serve>();
  sv->d = *desc;
  if (sv->d.chunk_steps == 0) sv->d.chunk_steps = 32;
  sv->t0 = std::chrono::steady_clock::now();
  hipError_t e = hipGetDevice(&sv->device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  int rc = sv->pool.init((size_t)desc->budget_bytes);
  if (rc) return rc;
  e = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&sv->setup, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&sv->setup_done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(sv->setup_done, sv->setup);
  if (e == hipSuccess) e = hipEventCreate(&sv->chunk_ev[0]);
  if (e == hipSuccess) e = hipEventCreate(&sv->chunk_ev[1]);
  if (e != hipSuccess) return hip_fail(e, "serve streams");
  lbm_serve* p = sv.release();
  p->worker = std::thread(worker, p);
  *out = p;
  return LBM_OK;
}

int lbm_serve_destroy(lbm_serve* sv) {
  if (!sv) return LBM_OK;
  {
    std::lock_guard<std::mutex> g(sv->mu);
    sv->stop = true;
    sv->cv_work.notify_all();
  }
  if (sv->worker.joinable()) sv->worker.join();
  if (sv->graph) (void)hipGraphExecDestroy(sv->graph);
  for (auto& kv : sv->classes) {
    if (kv.second->d_solves) (void)hipFree(kv.second->d_solves);
    if (kv.second->d_tiles) (void)hipFree(kv.second->d_tiles);
    if (kv.second->d_bar) (void)hipFree(kv.second->d_bar);
  }
  if (sv->stream) (void)hipStreamDestroy(sv->stream);
  if (sv->setup) (void)hipStreamDestroy(sv->setup);
  if (sv->setup_done) (void)hipEventDestroy(sv->setup_done);
  for (hipEvent_t ev : sv->chunk_ev)
    if (ev) (void)hipEventDestroy(ev);
  delete sv;
  return LBM_OK;
}

int lbm_serve_submit(lbm_serve* sv, const lbm_desc* desc, const float* mac4, int32_t steps, int64_t* ticket) {
  if (!mac4) return fail(LBM_E_INVALID, "null argument");
  return lbm_serve_submit_state(sv, desc, mac4, nullptr, 0, steps, 0, ticket);
}

int lbm_serve_submit_state(lbm_serve* sv, const lbm_desc* desc, const float* mac4, const void* state,
                           size_t state_bytes, int32_t steps, int32_t keep_state, int64_t* ticket) {
  if (!sv || !desc || !ticket) return fail(LBM_E_INVALID, "null argument");
  if ((mac4 == nullptr) == (state == nullptr)) return fail(LBM_E_INVALID, "give exactly one of mac4 and state");
  if (steps < 1) return fail(LBM_E_INVALID, "steps %d", steps);
  // validate the descriptor now (lbm_solve_create would on the worker)
  lbm_desc d = *desc;
  d.autotune = 0;
  if (d.abi_version != LBM_ABI_VERSION) return fail(LBM_E_INVALID, "descriptor abi_version %d", d.abi_version);
  if (d.q != 9 && d.q != 19 && d.q != 27) return fail(LBM_E_INVALID, "q must be 9, 19 or 27 (got %d)", d.q);
  if (d.nx < 1 || d.ny < 1 || d.nz < 1) return fail(LBM_E_INVALID, "grid %dx%dx%d", d.nx, d.ny, d.nz);
  if (d.diffuse) return fail(LBM_E_INVALID, "the serve loop runs the bounce-back coupling (diffuse = 0)");
  auto r = std::make_shared<Request>();
  r->d = d;
  r->steps = steps;
  r->cells = (uint32_t)((size_t)d.nx * d.ny * (d.q == 9 ? 1 : d.nz));
  r->footprint = solve_footprint(&d);
  if (r->footprint > (size_t)sv->d.budget_bytes)
    return fail(LBM_E_NOMEM, "request needs %zu bytes, more than the whole budget (%.0f)", r->footprint,
                sv->d.budget_bytes);
  if (state) {
    lbm_state_header h{};
    int rc = state_header_check(&d, state, state_bytes, &h);
    if (rc) return rc;
    if (h.body_bytes) return fail(LBM_E_INVALID, "a serve request cannot carry a rigid body");
    const uint8_t* b = static_cast<const uint8_t*>(state);
    r->state_in.assign(b, b + state_bytes);
  } else {
    r->mac_in.assign(mac4, mac4 + (size_t)r->cells * 4);
  }
  r->keep_state = keep_state != 0;
  std::lock_guard<std::mutex> g(sv->mu);
  r->ticket = sv->next_ticket++;
  r->t_submit = now_s(sv->t0);
  if (sv->first_submit < 0) sv->first_submit = r->t_submit;
  sv->submitted++;
  sv->requests[r->ticket] = r;
  sv->queue.push_back(r);
  sv->cv_work.notify_all();
  *ticket = r->ticket;
  return LBM_OK;
}

int lbm_serve_wait(lbm_serve* sv, int64_t ticket, double timeout_s, lbm_serve_result* out) {
  if (!sv) return fail(LBM_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto r = find(sv, ticket);
  if (!r) return fail(LBM_E_INVALID, "unknown ticket %lld", (long long)ticket);
  auto ready = [&] { return r->done; };
  if (timeout_s < 0) sv->cv_done.wait(lk, ready);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return fail(LBM_E_STATE, "request %lld not finished within %g s", (long long)ticket, timeout_s);
  if (out) {
    memset(out, 0, sizeof *out);
    out->ticket = r->ticket;
    out->status = r->status;
    out->steps = r->steps;
    out->cells = r->cells;
    out->t_submit = r->t_submit;
    out->t_admit = r->t_admit;
    out->t_finish = r->t_finish;
    if (r->forces.size() >= 6)
      for (int k = 0; k < 6; k++) out->force[k] = r->forces[r->forces.size() - 6 + k];
  }
  if (r->status) return fail(r->status, "request %lld: %s", (long long)ticket, r->error.c_str());
  return LBM_OK;
}

int lbm_serve_read_macro(lbm_serve* sv, int64_t ticket, float* mac4) {
  if (!sv || !mac4) return fail(LBM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto r = find(sv, ticket);
  if (!r) return fail(LBM_E_INVALID, "unknown ticket %lld", (long long)ticket);
  if (!r->done || r->status) return fail(LBM_E_STATE, "request %lld has no result", (long long)ticket);
  std::copy(r->mac_out.begin(), r->mac_out.end(), mac4);
  return LBM_OK;
}

int lbm_serve_read_forces(lbm_serve* sv, int64_t ticket, double* out6, int32_t max_records, int32_t* count) {
  if (!sv || !count || (max_records > 0 && !out6)) return fail(LBM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto r = find(sv, ticket);
  if (!r) return fail(LBM_E_INVALID, "unknown ticket %lld", (long long)ticket);
  if (!r->done || r->status) return fail(LBM_E_STATE, "request %lld has no result", (long long)ticket);
  const int32_t n = std::min<int32_t>((int32_t)(r->forces.size() / 6), std::max(max_records, 0));
  std::copy(r->forces.end() - (size_t)n * 6, r->forces.end(), out6);
  *count = n;
  return LBM_OK;
}

int lbm_serve_read_state(lbm_serve* sv, int64_t ticket, void* buf, size_t* size) {
  if (!sv || !size) return fail(LBM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto r = find(sv, ticket);
  if (!r) return fail(LBM_E_INVALID, "unknown ticket %lld", (long long)ticket);
  if (!r->done || r->status) return fail(LBM_E_STATE, "request %lld has no result", (long long)ticket);
  if (!r->keep_state) return fail(LBM_E_STATE, "request %lld was not submitted with keep_state", (long long)ticket);
  if (buf) {
    if (*size < r->state_out.size())
      return fail(LBM_E_INVALID, "state buffer of %zu bytes, need %zu", *size, r->state_out.size());
    std::copy(r->state_out.begin(), r->state_out.end(), static_cast<uint8_t*>(buf));
  }
  *size = r->state_out.size();
  return LBM_OK;
}

int lbm_serve_release(lbm_serve* sv, int64_t ticket) {
  if (!sv) return fail(LBM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  auto r = find(sv, ticket);
  if (!r) return fail(LBM_E_INVALID, "unknown ticket %lld", (long long)ticket);
  if (!r->done) return fail(LBM_E_STATE, "request %lld still running", (long long)ticket);
  sv->requests.erase(ticket);
  return LBM_OK;
}

int lbm_serve_drain(lbm_serve* sv, double timeout_s) {
  if (!sv) return fail(LBM_E_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(sv->mu);
  auto idle = [&] { return sv->completed + sv->failed == sv->submitted; };
  if (timeout_s < 0) sv->cv_done.wait(lk, idle);
  else if (!sv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), idle))
    return fail(LBM_E_STATE, "%lld requests still running after %g s",
                (long long)(sv->submitted - sv->completed - sv->failed), timeout_s);
  return LBM_OK;
}

int lbm_serve_read_stats(lbm_serve* sv, lbm_serve_stats* out) {
  if (!sv || !out) return fail(LBM_E_INVALID, "null argument");
  std::lock_guard<std::mutex> g(sv->mu);
  memset(out, 0, sizeof *out);
  out->submitted = sv->submitted;
  out->completed = sv->completed;
  out->failed = sv->failed;
  out->resident = (int64_t)sv->resident.size();
  out->queued = (int64_t)sv->queue.size();
  out->cell_updates = sv->cell_updates;
  out->first_submit = sv->first_submit;
  out->last_finish = sv->last_finish;
  if (!sv->latencies.empty()) {
    std::vector<double> l = sv->latencies;
    std::sort(l.begin(), l.end());
    auto pct = [&](double p) { return l[std::min(l.size() - 1, (size_t)(p * (l.size() - 1) + 0.5))]; };
    out->p50_latency = pct(0.50);
    out->p99_latency = pct(0.99);
  }
  const double span = sv->last_finish - sv->first_submit;
  out->mlups = span > 0 ? sv->cell_updates / span / 1e6 : 0;
  out->pool_bytes = (double)sv->pool.size();
  out->peak_resident_bytes = (double)sv->pool.peak();
  out->peak_resident = sv->peak_resident;
  out->chunks = sv->chunks;
  out->device_seconds = sv->device_seconds;
  out->device_bytes = sv->device_bytes;
  out->graph_instantiations = sv->graph_instantiations;
  out->graph_updates = sv->graph_updates;
  return LBM_OK;
}

double lbm_serve_clock(lbm_serve* sv) { return sv ? now_s(sv->t0) : 0.0; }

}  // extern "C"

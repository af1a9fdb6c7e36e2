// Error reporting shared by the engine's host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "../../include/lbm.h"
#include "dense.h"

namespace lbm {
extern thread_local std::string g_last_error;
// Records a formatted message for lbm_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);

// Hooks the AMR engine uses on its dense L0 solve.
// One step that skips the cells of refined blocks and writes the
// macroscopic field of the others.
int solve_enqueue_covered_step(lbm_solve* s, const int32_t* cover, int rb, int nbx, int nby);
// which = 0: the lattice holding the current state; 1: the other one.
float* solve_lattice(lbm_solve* s, int which);
float* solve_mac(lbm_solve* s);
hipStream_t solve_stream(lbm_solve* s);
// Uploads stop capturing graph chunks ahead of time (a solve stepped only
// through the hooks above).
void solve_no_eager_graphs(lbm_solve* s);

// Hooks the serve loop uses.  A solve can live in a larger device pool: its
// buffers then come from `Arena` instead of hipMalloc, and it runs on the
// pool's stream with a fixed tile shape (the batched kernel's).
struct Arena {
  virtual ~Arena() = default;
  virtual void* alloc(size_t bytes) = 0;  // null when the pool has no room
  virtual void release(void* p) = 0;
};
// Device bytes a solve of this descriptor takes (what solve_create_in asks
// the arena for, alignment included).
size_t solve_footprint(const lbm_desc* d);
// Tile of a solve that is not autotuned, as lbm_set_tile takes it; nx <= 0
// (a serve class, whose members differ in width) always gives linear tiles.
void default_tile(int storage, int nx, int* bx, int* by, int* flags);
// `flags` as lbm_set_tile takes them; a member uses its tile layout (bit 2,
// linear tiles), the batched launch the rest.
int solve_create_in(const lbm_desc* d, Arena* arena, hipStream_t stream, int bx, int by, int flags,
                    lbm_solve** out);
// The solve's batch table entry (step arguments by lattice parity, lattice
// and field pointers; the progress fields are the caller's) and its tiles.
int solve_batch_entry(lbm_solve* s, BatchSolve* e, uint32_t* ntiles);
// Host bookkeeping after `steps` batched steps: step count, lattice parity,
// and the force history of the last step summed (enqueued on the stream).
int solve_advance(lbm_solve* s, int steps, bool finalize_force);
uint32_t solve_cells(const lbm_solve* s);
// Dither counter of the solve's next step.
uint32_t solve_salt(const lbm_solve* s);
// Validates a saved state against a descriptor (magic, version, q, grid,
// storage, payload size) and returns its header.
int state_header_check(const lbm_desc* d, const void* buf, size_t bytes, lbm_state_header* out);
}  // namespace lbm

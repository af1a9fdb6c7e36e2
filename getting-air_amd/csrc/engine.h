// Error reporting shared by the engine's host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

#include "../../include/lbm.h"

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
}  // namespace lbm

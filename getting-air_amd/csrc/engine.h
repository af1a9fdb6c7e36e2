// Error reporting shared by the engine's host translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <string>

namespace lbm {
extern thread_local std::string g_last_error;
// Records a formatted message for lbm_last_error() and returns `code`.
int fail(int code, const char* fmt, ...);
int hip_fail(hipError_t e, const char* what);
}  // namespace lbm

// lbm_totals_noalloc: the FP64 population totals of lbm_read_totals for a
// caller-held device lattice, with no device allocation anywhere on the path.
//
// The totals kernel (k_totals) writes one FP64 partial per block and a
// completion counter, and the block that arrives last sums the partials in
// block order.  Here those partials and the counter live in `scratch`, a
// fixed array in the library's code object: each GPU the library runs on
// holds its own copy until the process ends.  Host code never passes the
// array's name to a kernel (in host code it names a host-side stand-in):
// the first call on a device looks up that device's copy with
// hipGetSymbolAddress, and every launch takes the pointer found there.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/lbm.h"
#include "dense.h"
#include "engine.h"

namespace {

// Partials of at most kMaxBlocks blocks (4 doubles each), then the counter.
constexpr int kMaxBlocks = 1024;
constexpr size_t kPartialBytes = (size_t)kMaxBlocks * 4 * sizeof(double);
constexpr size_t kScratchBytes = kPartialBytes + 256;
static_assert(kScratchBytes == LBM_NOALLOC_SCRATCH_BYTES, "lbm.h states the scratch size");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_ptr[kMaxDevices] = {};

// Address of this device's `scratch`, looked up once per device with that
// device current.
int scratch_ptr(unsigned char** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return lbm::hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return lbm::fail(LBM_E_DEVICE, "device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return lbm::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return lbm::fail(LBM_E_DEVICE, "hipGetSymbolAddress(scratch) gave a null address");
    g_ptr[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_ptr[dev]);
  return LBM_OK;
}

// The counter must start at zero.  k_totals' last block resets it, but the
// array keeps whatever the last call left, so each call clears it first with
// a kernel (not a memset node, so a caller may capture the call in a graph).
__global__ void k_clear_counter(unsigned* counter) { *counter = 0u; }

}  // namespace

extern "C" int lbm_totals_noalloc(int32_t q, int32_t storage, int64_t ncells, const void* lattice, double* out4,
                                  void* stream) {
  using namespace lbm;
  if (q != 9 && q != 19 && q != 27) return fail(LBM_E_INVALID, "q must be 9, 19 or 27 (got %d)", q);
  if (storage < LBM_STORE_F32 || storage > LBM_STORE_BF16) return fail(LBM_E_INVALID, "bad storage %d", storage);
  if (ncells < 1 || ncells > LBM_NOALLOC_MAX_CELLS)
    return fail(LBM_E_INVALID, "ncells %lld outside 1..%lld (the no-allocation totals' limit)", (long long)ncells,
                (long long)LBM_NOALLOC_MAX_CELLS);
  if (!lattice || !out4) return fail(LBM_E_INVALID, "null lattice or output");
  unsigned char* base = nullptr;
  int rc = scratch_ptr(&base);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  TotalsArgs a{};
  a.ncells = (uint32_t)ncells;
  a.partials = reinterpret_cast<double*>(base);
  a.counter = reinterpret_cast<unsigned*>(base + kPartialBytes);
  a.out = out4;
  // the grid of lbm_read_totals, so the fixed-order sum is the same
  const int nblocks = (int)std::min<int64_t>(kMaxBlocks, (ncells + 255) / 256);
  hipLaunchKernelGGL(k_clear_counter, dim3(1), dim3(1), 0, s, a.counter);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "k_clear_counter");
  e = launch_totals(q, storage, lattice, a, nblocks, s);
  if (e != hipSuccess) return hip_fail(e, "k_totals");
  return LBM_OK;
}

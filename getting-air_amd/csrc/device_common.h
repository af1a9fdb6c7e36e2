// Device-side helpers shared by the engine's HIP translation units:
// population storage formats, FP64 wave/block reductions and the fixed-order
// "last block sums the partials" pattern.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace lbm {

// ---------------------------------------------------------------------------
// Population storage.  Every format holds the SHIFTED population
// g_i = f_i - w_i (the reference's f-pack mode 2, common_fpack.wgsl:105-121,
// applied to all widths), and the kernels collide in g directly.  g is one to
// two orders of magnitude smaller than f, so the same number of bits carries
// that much less absolute round-off; the arithmetic is FP32 throughout.
// load/store move g; the 16-bit formats round to nearest even.
// ---------------------------------------------------------------------------
// Element i of a plane as a 32-bit byte offset from the plane base: the base
// is wave-uniform, so loads and stores can use the scalar-base + 32-bit
// vector-offset addressing form instead of 64-bit per-lane address math.
// Every plane is below 4 GiB (the engine caps cells per plane).
template <class T>
__device__ __forceinline__ const T* at(const T* p, uint32_t i) {
  return reinterpret_cast<const T*>(reinterpret_cast<const char*>(p) + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ T* at(T* p, uint32_t i) {
  return reinterpret_cast<T*>(reinterpret_cast<char*>(p) + i * (uint32_t)sizeof(T));
}

struct StoreF32 {
  using T = float;
  static constexpr int kBytes = 4;
  __device__ __forceinline__ static float load(const T* p, uint32_t i) { return *at(p, i); }
  __device__ __forceinline__ static void store(T* p, uint32_t i, float g) { *at(p, i) = g; }
  __device__ __forceinline__ static float load_nt(const T* p, uint32_t i) { return __builtin_nontemporal_load(at(p, i)); }
  __device__ __forceinline__ static void store_nt(T* p, uint32_t i, float g) { __builtin_nontemporal_store(g, at(p, i)); }
  __device__ __forceinline__ static uint32_t dither(uint32_t, int) { return 0; }
  __device__ __forceinline__ static T round_sr(float g, uint32_t) { return g; }
};

__device__ __forceinline__ float half_bits_to_float(uint16_t h) {
  return static_cast<float>(__builtin_bit_cast(_Float16, h));
}
__device__ __forceinline__ uint16_t float_to_half_bits(float v) {
  // fptrunc lowers to v_cvt_f16_f32, round-to-nearest-even with f16
  // subnormals (the default gfx950 mode).
  return __builtin_bit_cast(uint16_t, static_cast<_Float16>(v));
}
__device__ __forceinline__ float bf16_bits_to_float(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}
__device__ __forceinline__ uint16_t float_to_bf16_bits(float v) {
  uint32_t u = __float_as_uint(v);
  if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<uint16_t>((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                                      // ties to even
  return static_cast<uint16_t>(u >> 16);
}

// Stochastic rounding of the 16-bit formats: r, uniform over the bits that
// the format drops, is added to the binary32 magnitude bits, and the sum is
// truncated toward zero.  The expected stored value is then g itself, so the
// rounding error has no systematic part even where the same g is stored step
// after step (a steady flow), which round-to-nearest-even cannot avoid.
struct StoreF16 {
  using T = uint16_t;
  static constexpr int kBytes = 2;
  // 13 dropped bits for binary16 normals (f16 subnormals truncate a few more)
  __device__ __forceinline__ static uint32_t dither(uint32_t h, int off) { return __builtin_amdgcn_ubfe(h, off, 13); }
  __device__ __forceinline__ static T round_sr(float g, uint32_t r) {
    const float d = __uint_as_float(__float_as_uint(g) + r);
    return __builtin_bit_cast(uint16_t, __builtin_amdgcn_cvt_pkrtz(d, 0.f)[0]);
  }
  __device__ __forceinline__ static float load(const T* p, uint32_t i) { return half_bits_to_float(*at(p, i)); }
  __device__ __forceinline__ static void store(T* p, uint32_t i, float g) { *at(p, i) = float_to_half_bits(g); }
  __device__ __forceinline__ static float load_nt(const T* p, uint32_t i) {
    return half_bits_to_float(__builtin_nontemporal_load(at(p, i)));
  }
  __device__ __forceinline__ static void store_nt(T* p, uint32_t i, float g) {
    __builtin_nontemporal_store(float_to_half_bits(g), at(p, i));
  }
};

struct StoreBF16 {
  using T = uint16_t;
  static constexpr int kBytes = 2;
  __device__ __forceinline__ static uint32_t dither(uint32_t h, int off) { return __builtin_amdgcn_ubfe(h, off, 16); }
  __device__ __forceinline__ static T round_sr(float g, uint32_t r) {
    return static_cast<uint16_t>((__float_as_uint(g) + r) >> 16);
  }
  __device__ __forceinline__ static float load(const T* p, uint32_t i) { return bf16_bits_to_float(*at(p, i)); }
  __device__ __forceinline__ static void store(T* p, uint32_t i, float g) { *at(p, i) = float_to_bf16_bits(g); }
  __device__ __forceinline__ static float load_nt(const T* p, uint32_t i) {
    return bf16_bits_to_float(__builtin_nontemporal_load(at(p, i)));
  }
  __device__ __forceinline__ static void store_nt(T* p, uint32_t i, float g) {
    __builtin_nontemporal_store(float_to_bf16_bits(g), at(p, i));
  }
};

// Chooses the plain or non-temporal access per kernel instantiation.
template <class ST, bool NT>
__device__ __forceinline__ float ld(const typename ST::T* p, uint32_t i) {
  if constexpr (NT) return ST::load_nt(p, i);
  else return ST::load(p, i);
}
template <class ST, bool NT>
__device__ __forceinline__ void st(typename ST::T* p, uint32_t i, float g) {
  if constexpr (NT) ST::store_nt(p, i, g);
  else ST::store(p, i, g);
}
// Store with stochastic rounding (16-bit formats); r from ST::dither.
template <class ST, bool NT>
__device__ __forceinline__ void st_sr(typename ST::T* p, uint32_t i, float g, uint32_t r) {
  const typename ST::T v = ST::round_sr(g, r);
  if constexpr (NT) __builtin_nontemporal_store(v, at(p, i));
  else *at(p, i) = v;
}

// Two stochastically rounded stores of one cell (planes pi and pj).  binary16
// converts both values with one packed round-toward-zero conversion, the
// same bits as two single conversions.  The high half is taken from the
// packed word by a shift: a store of the vector's element 1 itself was
// compiled into a store of the low half.
template <class ST, bool NT>
__device__ __forceinline__ void st_sr_pair(typename ST::T* pi, typename ST::T* pj, uint32_t i, float gi, float gj,
                                           uint32_t ri, uint32_t rj) {
  if constexpr (std::is_same_v<ST, StoreF16>) {
    const uint32_t h = __builtin_bit_cast(uint32_t, __builtin_amdgcn_cvt_pkrtz(__uint_as_float(__float_as_uint(gi) + ri),
                                                                               __uint_as_float(__float_as_uint(gj) + rj)));
    const uint16_t hi = (uint16_t)h, hj = (uint16_t)(h >> 16);
    if constexpr (NT) {
      __builtin_nontemporal_store(hi, at(pi, i));
      __builtin_nontemporal_store(hj, at(pj, i));
    } else {
      *at(pi, i) = hi;
      *at(pj, i) = hj;
    }
  } else {
    st_sr<ST, NT>(pi, i, gi, ri);
    st_sr<ST, NT>(pj, i, gj, rj);
  }
}

// Per-cell, per-step 32-bit hash for the dither (a murmur3-style finalizer).
__device__ __forceinline__ uint32_t cell_hash(uint32_t cell, uint32_t salt) {
  uint32_t h = cell * 0x9E3779B1u + salt * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return h;
}

// ---------------------------------------------------------------------------
// Correctly rounded binary32 arithmetic for integer decisions (solid class
// bytes, refinement criterion) that a host restatement must reproduce bit for
// bit.  Each operation is evaluated in binary64 and rounded once to binary32.
// A product of two binary32 values is exact in binary64; for +, -, / and
// sqrt, rounding to binary64 first and then to binary32 gives the correctly
// rounded binary32 result because 53 >= 2 * 24 + 2 (Figueroa, "When is
// double rounding innocuous?", SIGNUM 1995).
// ---------------------------------------------------------------------------
// contract(off) keeps the operations free of the contraction flag, so that
// no later pass may fuse one of them with another into an FMA.
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
  return (float)((double)a + (double)b);
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
  return (float)((double)a - (double)b);
}
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return (float)((double)a * (double)b);
}
__device__ __forceinline__ float div_rn(float a, float b) {
#pragma clang fp contract(off)
  return (float)((double)a / (double)b);
}
__device__ __forceinline__ float sqrt_rn(float a) {
#pragma clang fp contract(off)
  return (float)__builtin_sqrt((double)a);
}

// ---------------------------------------------------------------------------
// FP64 reductions.  Wave64 butterfly with __shfl_xor, then one slot per wave
// in LDS, then wave 0 folds the slots.  The order of every addition is fixed
// by the launch shape, so results repeat bit for bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums K doubles across a block of up to 1024 threads; lane 0 of wave 0 ends
// with the totals in v[].  `lds` holds at least 16*K doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double* lds) {
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const int nthreads = blockDim.x * blockDim.y;
  const int lane = tid & 63, wave = tid >> 6, nwaves = (nthreads + 63) >> 6;
#pragma unroll
  for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) lds[wave * K + k] = v[k];
  }
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = lane < nwaves ? lds[lane * K + k] : 0.0;
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = wave_sum(v[k]);
  }
}

// Writes this block's K partials, then the last block to arrive sums all
// blocks' partials in block order and hands the totals to `finish`.  No block
// waits on another, so the kernel always drains.  `counter` must be zero on
// entry; the last block resets it.
template <int K, class Finish>
__device__ __forceinline__ void partials_then_fixed_order_sum(const double (&v)[K], double* partials,
                                                              unsigned* counter, Finish finish) {
  __shared__ bool last;
  const int tid = threadIdx.x + threadIdx.y * blockDim.x;
  const unsigned nblocks = gridDim.x * gridDim.y * gridDim.z;
  const unsigned b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  if (tid == 0) {
#pragma unroll
    for (int k = 0; k < K; k++) partials[(size_t)b * K + k] = v[k];
    __threadfence();
    last = atomicAdd(counter, 1u) == nblocks - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  if (tid < K) {
    double s = 0.0;
    for (unsigned j = 0; j < nblocks; j++)
      s += __hip_atomic_load(&partials[(size_t)j * K + tid], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    finish(tid, s);
  }
  if (tid == 0) *counter = 0u;
}

}  // namespace lbm

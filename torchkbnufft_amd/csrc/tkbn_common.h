// Shared definitions for the table-mode Kaiser-Bessel NUFFT kernels (gfx950).
//
// Conventions of every entry point (declared in include/tkbn_hip.h):
//   * pointers are device pointers to contiguous arrays; complex data is
//     interleaved (re, im) float2 / double2, trajectories are real;
//   * kernels are enqueued on the caller's stream; nothing synchronises the
//     host and nothing allocates device memory;
//   * the return value is a status code (TKBN_OK == 0); tkbn_last_error()
//     returns a thread-local message for the last failure on this thread.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tkbn_hip.h"

namespace tkbn {

void set_error(const char* fmt, ...);
// Records the name of the interpolation kernel a call has just enqueued (thread-local;
// read back with tkbn_last_interp_kernel).
void note_kernel(const char* name);

#define TKBN_REQUIRE(cond, code, ...)     \
  do {                                    \
    if (!(cond)) {                        \
      ::tkbn::set_error(__VA_ARGS__);     \
      return (code);                      \
    }                                     \
  } while (0)

#define TKBN_HIP_CHECK(expr)                                                        \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess) {                                                         \
      ::tkbn::set_error("%s failed: %s", #expr, hipGetErrorString(_e));             \
      return TKBN_ERR_HIP;                                                          \
    }                                                                               \
  } while (0)

// Checks the launch that was just enqueued (configuration errors only; no sync).
#define TKBN_LAUNCH_CHECK(name)                                                     \
  do {                                                                              \
    hipError_t _e = hipGetLastError();                                              \
    if (_e != hipSuccess) {                                                         \
      ::tkbn::set_error("launch of %s failed: %s", name, hipGetErrorString(_e));    \
      return TKBN_ERR_HIP;                                                          \
    }                                                                               \
  } while (0)

template <typename T>
struct cplx_t;
template <>
struct cplx_t<float> {
  using type = float2;
};
template <>
struct cplx_t<double> {
  using type = double2;
};
template <typename T>
using cplx = typename cplx_t<T>::type;

template <typename T>
__host__ __device__ inline cplx<T> make_c(T re, T im) {
  cplx<T> r;
  r.x = re;
  r.y = im;
  return r;
}

template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}

// a * conj(b)
template <typename C>
__device__ __forceinline__ C cmul_conj(C a, C b) {
  C r;
  r.x = a.x * b.x + a.y * b.y;
  r.y = a.y * b.x - a.x * b.y;
  return r;
}

// acc += a * b
template <typename C>
__device__ __forceinline__ void cmac(C& acc, C a, C b) {
  acc.x += a.x * b.x - a.y * b.y;
  acc.y += a.x * b.y + a.y * b.x;
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Division by an invariant divisor d >= 1 as a multiply-high and shift (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994):
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1, n / d = (mulhi(n, m) + n) >> s.
// Exact for 0 <= n < 2^31 (the sum then fits in 32 bits).
struct FastDiv {
  uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {  // d >= 1
  uint32_t s = 0;
  while (s < 31 && (1ull << s) < d) ++s;
  const uint64_t m = ((((uint64_t)1 << 32) * (((uint64_t)1 << s) - d)) / d) + 1;
  return FastDiv{d, (uint32_t)m, s};
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) { return (__umulhi(n, f.m) + n) >> f.s; }
__device__ __forceinline__ uint32_t fmodd(uint32_t n, const FastDiv& f) { return n - fdiv(n, f) * f.d; }

// Zero-fill n elements (a kernel, not hipMemsetAsync, so callers stay safe to
// capture in a HIP graph).
template <typename E>
__global__ void __launch_bounds__(256) zero_fill_kernel(E* __restrict__ p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = E{};
}

template <typename E>
inline int launch_zero_fill(E* p, int64_t n, hipStream_t s) {
  if (n == 0) return TKBN_OK;
  const int64_t blocks = ceil_div(n, 256) < 8192 ? ceil_div(n, 256) : 8192;
  zero_fill_kernel<E><<<dim3((unsigned)blocks), dim3(256), 0, s>>>(p, n);
  TKBN_LAUNCH_CHECK("zero_fill_kernel");
  return TKBN_OK;
}


}  // namespace tkbn

// Shared helpers of the op2hip library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <type_traits>

#include "op2hip.h"

namespace op2hip {

void set_error(const char *fmt, ...);

// Device address of the no-allocation entry point's scratch array on
// `device` (the current device); defined beside the array (p2.hip).
int noalloc_scratch_address(int device, void **out);

#define OP2HIP_CHECK(expr)                                                     \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::op2hip::set_error("%s failed: %s (%s:%d)", #expr,                      \
                          hipGetErrorString(e_), __FILE__, __LINE__);          \
      return OP2HIP_ERR_HIP;                                                   \
    }                                                                          \
  } while (0)

#define OP2HIP_LAUNCHED()                                                      \
  do {                                                                         \
    hipError_t e_ = hipGetLastError();                                         \
    if (e_ != hipSuccess) {                                                    \
      ::op2hip::set_error("kernel launch failed: %s (%s:%d)",                  \
                          hipGetErrorString(e_), __FILE__, __LINE__);          \
      return OP2HIP_ERR_HIP;                                                   \
    }                                                                          \
  } while (0)

#define OP2HIP_REQUIRE(cond, ...)                                              \
  do {                                                                         \
    if (!(cond)) {                                                             \
      ::op2hip::set_error(__VA_ARGS__);                                        \
      return OP2HIP_ERR_ARG;                                                   \
    }                                                                          \
  } while (0)

inline hipStream_t stream_of(void *s) { return reinterpret_cast<hipStream_t>(s); }

inline uint32_t ceil_div(int64_t a, int64_t b) {
  return static_cast<uint32_t>((a + b - 1) / b);
}

constexpr int kWave = 64;

// Element index of thread slot i of a launch over [start, end): the identity,
// or an indirection through a subset / colour list.
__device__ __forceinline__ int32_t element_of(int32_t i, const int32_t *list) {
  return list ? list[i] : i;
}

// Device-scope atomic add.  The library is built with -munsafe-fp-atomics so
// float/double adds map to the hardware global (and LDS) atomics.
template <typename T>
__device__ __forceinline__ void atomic_add(T *p, T v) {
  atomicAdd(p, v);
}

// Plain read-modify-write add (no concurrent writer by construction).
template <typename T>
__device__ __forceinline__ void plain_add(T *p, T v) {
  *p += v;
}

// Binary search of `col` in the sorted colidx[lo, hi); returns the slot or -1.
__device__ __forceinline__ int32_t find_slot(const int32_t *__restrict__ colidx,
                                             int32_t lo, int32_t hi, int32_t col) {
  int32_t l = lo, h = hi - 1;
  while (l <= h) {
    const int32_t mid = l + ((h - l) >> 1);
    const int32_t c = colidx[mid];
    if (c == col) return mid;
    if (c < col) l = mid + 1; else h = mid - 1;
  }
  return -1;
}

}  // namespace op2hip

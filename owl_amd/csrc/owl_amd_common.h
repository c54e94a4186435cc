// Shared definitions for the owl_amd HIP library (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/owl_amd.h"

typedef __bf16 bf16_t;
typedef _Float16 f16_t;
// OCP FP8 e4m3 (e4m3fn: bias 7, max 448, 0x7f = NaN), stored as a byte.
enum class fp8e4m3_t : unsigned char {};

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

#define OWL_CHECK_HIP(expr)                                   \
  do {                                                        \
    hipError_t _e = (expr);                                   \
    if (_e != hipSuccess) {                                   \
      owl_amd_set_error("%s:%d %s: %s", __FILE__, __LINE__,   \
                        #expr, hipGetErrorString(_e));        \
      return OWL_AMD_ERR_HIP;                                 \
    }                                                         \
  } while (0)

extern "C" void owl_amd_set_error(const char *fmt, ...);

// Conversions between storage types and the FP32 accumulator.
template <typename T> __device__ __forceinline__ float to_f32(T v) { return (float)v; }
template <typename T> __device__ __forceinline__ T from_f32(float v) { return (T)v; }
template <> __device__ __forceinline__ double from_f32<double>(float v) { return (double)v; }

namespace owl_amd {
// Hardware dispatches workgroup i of a 1-D grid to XCD i % 8, and each XCD
// has its own L2.  Map workgroup i to a logical tile so that XCD x runs a
// contiguous range of logical tiles (the first nblocks % 8 XCDs one more
// than the others): tiles that share operands then share an L2.  A bijection
// on [0, nblocks) for every nblocks.
__device__ __forceinline__ int xcd_swizzle(int bid, int nblocks) {
  const int q = nblocks >> 3, r = nblocks & 7, x = bid & 7;
  return x * q + (x < r ? x : r) + (bid >> 3);
}
}  // namespace owl_amd

// FP8 (OCP e4m3) forward support: per-tensor amax scaling and quantisation.
//
// quantize(src): amax = max |src|;  scale = amax / 448 (448 = largest finite
// e4m3 value);  q = e4m3(src * 448 / amax) rounded to nearest even and
// saturated at +-448.  The forward conv then runs the FP8 MFMA on q and
// multiplies the FP32 accumulator by scale_x * scale_w in its epilogue.
#include "conv_kernels.h"
#include "owl_amd_common.h"

namespace owl_amd {

// float -> e4m3 (bias 7, 3 mantissa bits, subnormals below 2^-6, no
// infinities, max 448), round to nearest even, saturating.
__device__ __forceinline__ unsigned char f32_to_e4m3(float f) {
  unsigned sign = (__float_as_uint(f) >> 24) & 0x80u;
  float a = fabsf(f);
  if (a != a) return (unsigned char)(sign | 0x7Fu);
  if (a >= 448.0f) return (unsigned char)(sign | 0x7Eu);
  if (a < 0.015625f) {  // subnormal range: multiples of 2^-9
    float m = __builtin_rintf(a * 512.0f);
    return (unsigned char)(sign | (unsigned)m);  // m == 8 encodes 2^-6 exactly
  }
  int e = ((int)((__float_as_uint(a) >> 23) & 0xFFu)) - 127;  // -6 .. 8
  float frac = __uint_as_float((__float_as_uint(a) & 0x007FFFFFu) | 0x3F800000u);  // [1, 2)
  float q = __builtin_rintf((frac - 1.0f) * 8.0f);
  int qi = (int)q;
  if (qi == 8) {
    qi = 0;
    ++e;
  }
  unsigned code = ((unsigned)(e + 7) << 3) | (unsigned)qi;
  if (code > 0x7Eu) code = 0x7Eu;
  return (unsigned char)(sign | code);
}

__global__ void amax_reset_kernel(unsigned *amax) { *amax = 0u; }

template <typename T>
__global__ __launch_bounds__(256) void amax_kernel(const T *__restrict__ src, int64_t n, unsigned *amax) {
  float m = 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf((float)src[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) atomicMax(amax, __float_as_uint(m));  // |x| bits order like floats
}

template <typename T>
__global__ __launch_bounds__(256) void quant_kernel(const T *__restrict__ src, int64_t n,
                                                    const unsigned *__restrict__ amax,
                                                    unsigned char *__restrict__ dst) {
  const float a = __uint_as_float(*amax);
  const float inv = a > 0.0f ? 448.0f / a : 1.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = f32_to_e4m3((float)src[i] * inv);
}

__global__ void scale_finalize_kernel(unsigned *amax_then_scale) {
  float a = __uint_as_float(*amax_then_scale);
  float s = a > 0.0f ? a / 448.0f : 1.0f;
  *(float *)amax_then_scale = s;
}

template <typename T>
void run_fp8_quantize(const T *src, int64_t n, unsigned char *dst, float *scale, hipStream_t st) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  unsigned *a = (unsigned *)scale;
  launch_k(amax_reset_kernel, dim3(1), dim3(1), st, a);
  launch_k((amax_kernel<T>), dim3((unsigned)blocks), dim3(256), st, src, n, a);
  launch_k((quant_kernel<T>), dim3((unsigned)blocks), dim3(256), st, src, n, a, dst);
  launch_k(scale_finalize_kernel, dim3(1), dim3(1), st, a);
}

// Device-side e4m3 encoder exposed for the conversion test.
__global__ void e4m3_encode_kernel(const float *src, int64_t n, unsigned char *dst) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = f32_to_e4m3(src[i]);
}

void run_e4m3_encode(const float *src, int64_t n, unsigned char *dst, hipStream_t st) {
  launch_k(e4m3_encode_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), st, src, n, dst);
}

// FP8 kernel [tap][ci][co] -> [tap][co][ci] (the rows of ci the halo forward
// reads with 16-byte loads); one thread per output byte, writes coalesced.
__global__ __launch_bounds__(256) void fp8_pack_krows_kernel(const unsigned char *__restrict__ w,
                                                             unsigned char *__restrict__ wp, int64_t n, int Ci,
                                                             int Co) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256) {
    const int ci = (int)(e % Ci);
    const int64_t r = e / Ci;
    const int co = (int)(r % Co);
    const int64_t t = r / Co;
    wp[e] = w[(t * Ci + ci) * Co + co];
  }
}

void run_fp8_pack_krows(const unsigned char *w, unsigned char *wp, int64_t T, int Ci, int Co, hipStream_t st) {
  const int64_t n = T * Ci * Co;
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  launch_k(fp8_pack_krows_kernel, dim3((unsigned)blocks), dim3(256), st, w, wp, n, Ci, Co);
}

template void run_fp8_quantize<float>(const float *, int64_t, unsigned char *, float *, hipStream_t);
template void run_fp8_quantize<bf16_t>(const bf16_t *, int64_t, unsigned char *, float *, hipStream_t);

}  // namespace owl_amd

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
// infinities, max 448), round to nearest even, saturating; NaN -> 0x7F with
// the sign.  Normal range [2^-6, 448): rebias the FP32 exponent by 127 - 7
// and round the 23-bit mantissa to its top 3 bits in integer arithmetic
// (add 2^19 - 1 plus the kept lsb, shift by 20; a carry moves into the
// exponent, 448 and above saturate).  The subnormal range takes multiples of
// 2^-9 by rounding a * 512.  This gives the same code as decomposing the
// value into exponent and fraction and rounding (frac - 1) * 8 with rintf,
// for every one of the 2^32 FP32 inputs (checked exhaustively on the host),
// in a quarter of the instructions.
__device__ __forceinline__ unsigned char f32_to_e4m3(float f) {
  const unsigned u = __float_as_uint(f);
  const unsigned sign = (u >> 24) & 0x80u;
  const unsigned a = u & 0x7FFFFFFFu;
  if (a > 0x7F800000u) return (unsigned char)(sign | 0x7Fu);  // NaN
  if (a >= 0x43E00000u) return (unsigned char)(sign | 0x7Eu);  // >= 448 (and inf): saturate
  if (a < 0x3C800000u) {  // below 2^-6: multiples of 2^-9 (m == 8 encodes 2^-6 exactly)
    const float m = __builtin_rintf(__uint_as_float(a) * 512.0f);
    return (unsigned char)(sign | (unsigned)m);
  }
  const unsigned r = a - (120u << 23);
  const unsigned code = (r + 0x7FFFFu + ((r >> 20) & 1u)) >> 20;
  return (unsigned char)(sign | (code > 0x7Eu ? 0x7Eu : code));
}

// Per-workgroup maxima of the amax pass, read back by every workgroup of the
// quantisation pass: no atomics (2048 device-scope atomicMax on one address
// took ~25 us of a 30 us amax pass) and no reset or finalize kernels.  One
// array per GPU (a device variable, referenced by name only in device code);
// quantisations that overlap on one GPU race, like the other per-device
// library buffers.
constexpr int AMAX_PARTS = 2048;
__device__ float g_amax_part[AMAX_PARTS];

// Both passes move whole 16-byte chunks of src (8 BF16 or 4 FP32 values),
// four in flight per thread and step, when src is 16-byte aligned and dst 8-
// (BF16) or 4-byte (FP32) aligned, the rest element by element: one element
// per thread with byte stores ran the quantisation of a 51 MB BF16 tensor at
// 0.23 TB/s.
template <typename T>
__global__ __launch_bounds__(256) void amax_kernel(const T *__restrict__ src, int64_t n, bool vec) {
  constexpr int V = 16 / (int)sizeof(T);
  float m = 0.0f;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i0 = 0;
  if (vec) {
    const int64_t nv = n / V;
    const uint4 *s4 = (const uint4 *)src;
    auto take = [&](const uint4 &c) __attribute__((always_inline)) {
      T e[V];
      __builtin_memcpy(e, &c, 16);
#pragma unroll
      for (int k = 0; k < V; ++k) m = fmaxf(m, fabsf((float)e[k]));
    };
    int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; v + 3 * stride < nv; v += 4 * stride) {
      const uint4 c0 = s4[v], c1 = s4[v + stride], c2 = s4[v + 2 * stride], c3 = s4[v + 3 * stride];
      take(c0);
      take(c1);
      take(c2);
      take(c3);
    }
    for (; v < nv; v += stride) take(s4[v]);
    i0 = nv * V;
  }
  for (int64_t i = i0 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride) m = fmaxf(m, fabsf((float)src[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) g_amax_part[blockIdx.x] = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
}

// amax from the parts of the amax pass (parts entries), in every workgroup
__device__ __forceinline__ float amax_of_parts(int parts) {
  float m = 0.0f;
  for (int k = threadIdx.x; k < parts; k += 256) m = fmaxf(m, g_amax_part[k]);
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  __shared__ float wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  return fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
}

template <typename T>
__global__ __launch_bounds__(256) void quant_kernel(const T *__restrict__ src, int64_t n, int parts,
                                                    unsigned char *__restrict__ dst, float *__restrict__ scale,
                                                    bool vec) {
  constexpr int V = 16 / (int)sizeof(T);
  const float a = amax_of_parts(parts);
  const float inv = a > 0.0f ? 448.0f / a : 1.0f;
  if (blockIdx.x == 0 && threadIdx.x == 0) *scale = a > 0.0f ? a / 448.0f : 1.0f;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i0 = 0;
  if (vec) {
    const int64_t nv = n / V;
    const uint4 *s4 = (const uint4 *)src;
    auto put = [&](const uint4 &c, int64_t v) __attribute__((always_inline)) {
      T e[V];
      __builtin_memcpy(e, &c, 16);
      unsigned char q[V];
#pragma unroll
      for (int k = 0; k < V; ++k) q[k] = f32_to_e4m3((float)e[k] * inv);
      if constexpr (V == 8) {
        uint2 o;
        __builtin_memcpy(&o, q, 8);
        ((uint2 *)dst)[v] = o;
      } else {
        unsigned o;
        __builtin_memcpy(&o, q, 4);
        ((unsigned *)dst)[v] = o;
      }
    };
    int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (; v + 3 * stride < nv; v += 4 * stride) {
      const uint4 c0 = s4[v], c1 = s4[v + stride], c2 = s4[v + 2 * stride], c3 = s4[v + 3 * stride];
      put(c0, v);
      put(c1, v + stride);
      put(c2, v + 2 * stride);
      put(c3, v + 3 * stride);
    }
    for (; v < nv; v += stride) put(s4[v], v);
    i0 = nv * V;
  }
  for (int64_t i = i0 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += stride)
    dst[i] = f32_to_e4m3((float)src[i] * inv);
}

template <typename T>
void run_fp8_quantize(const T *src, int64_t n, unsigned char *dst, float *scale, hipStream_t st) {
  constexpr int V = 16 / (int)sizeof(T);
  const bool vec = ((uintptr_t)src % 16 == 0) && ((uintptr_t)dst % V == 0);
  const int64_t units = vec ? n / V + n % V : n;
  int64_t blocks = (units + 1023) / 1024;  // about four units per thread
  if (blocks > AMAX_PARTS) blocks = AMAX_PARTS;
  if (blocks < 1) blocks = 1;
  launch_k((amax_kernel<T>), dim3((unsigned)blocks), dim3(256), st, src, n, vec);
  launch_k((quant_kernel<T>), dim3((unsigned)blocks), dim3(256), st, src, n, (int)blocks, dst, scale, vec);
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

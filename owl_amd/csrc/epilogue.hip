// Adjoint of the conv layer's fused epilogue y = act(z + bias), z = conv(x, w):
//   dz = dy * (signum(pre) + 1) / 2   (relu; owl_algodiff_ops.ml:774 with
//        signum(x) = x > 0 ? 1 : x < 0 ? -1 : x, owl_ndarray_maths_stub.c:2756,
//        so the factor is 1 above zero, 1/2 at +-0, 0 below, NaN for NaN)
//   dz = dy                           (no activation)
//   db[c] = sum over pixels of dz[., c]
// pre = z + bias is [npix][C] with C fastest.  The bias gradient is summed in
// a fixed order -- within a thread in pixel order, then over the lanes of a
// block in lane order, then over pixel chunks (strided sums in chunk order,
// then a fixed tree) -- so every launch gives the same bits.
#include "owl_amd_common.h"
#include "conv_kernels.h"

namespace owl_amd {

// relu factor of Owl's adjoint, in the arithmetic type A
template <typename A> __device__ __forceinline__ A relu_factor(A z) {
  const A s = z > (A)0 ? (A)1 : (z < (A)0 ? (A)-1 : z);  // signum, keeping +-0 and NaN
  return (s + (A)1) / (A)2;
}

// g * factor, the IEEE product, for a result stored as T.  With FP16 results
// the compiler selects the FP32 product and its rounding to FP16 as one
// mixed-precision FMA with a +0 addend (v_fma_mixlo_f16 ..., 0), which gives
// +0 where the product is -0 (a negative gradient times the factor 0, as
// Owl's dy * (signum z + 1) / 2 keeps it); so for FP16 the product's sign
// bit is set from the operands' explicitly, which the compiler cannot fold.
template <typename T, typename A> __device__ __forceinline__ A adjoint_product(A g, A factor) {
  A p = g * factor;
  if constexpr (__is_same(T, f16_t)) {
    const unsigned sign = (__builtin_bit_cast(unsigned, g) ^ __builtin_bit_cast(unsigned, factor)) & 0x80000000u;
    p = __builtin_bit_cast(A, (__builtin_bit_cast(unsigned, p) & 0x7fffffffu) | sign);
  }
  return p;
}

// Block: 256 threads = PL pixel lanes x CW channel vectors of VEC channels.
// Grid: (pixel chunks of PCH pixels, channel-vector blocks of CW).
template <typename T, typename A, int VEC>
__global__ __launch_bounds__(256) void bias_act_bwd_kernel(int act, int64_t npix, int C, int CW, int PCH,
                                                           const T *__restrict__ pre, const T *dy, T *dz,
                                                           A *__restrict__ partial) {
  __shared__ A red[256 * VEC];
  const int t = threadIdx.x;
  const int cw = t % CW, pl = t / CW, PL = 256 / CW;
  const int cv = blockIdx.y * CW + cw;  // channel vector
  const int nv = C / VEC;
  const int64_t p0 = (int64_t)blockIdx.x * PCH;
  int64_t p1 = p0 + PCH;
  if (p1 > npix) p1 = npix;
  A s[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) s[v] = (A)0;
  if (pl < PL && cv < nv) {
    for (int64_t p = p0 + pl; p < p1; p += PL) {
      const int64_t e = p * C + (int64_t)cv * VEC;
      T g[VEC], z[VEC];
      __builtin_memcpy(g, dy + e, sizeof(g));
      if (act == OWL_AMD_ACT_RELU) __builtin_memcpy(z, pre + e, sizeof(z));
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        A d = (A)g[v];
        if (act == OWL_AMD_ACT_RELU) d = adjoint_product<T, A>(d, relu_factor<A>((A)z[v]));
        g[v] = (T)d;
        s[v] += (A)g[v];  // the stored (rounded) gradient is what the bias sums
      }
      __builtin_memcpy(dz + e, g, sizeof(g));
    }
  }
  if (!partial) return;
#pragma unroll
  for (int v = 0; v < VEC; ++v) red[t * VEC + v] = s[v];
  __syncthreads();
  if (pl == 0 && cv < nv) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      A a = (A)0;
      for (int l = 0; l < PL; ++l) a += red[(l * CW + cw) * VEC + v];
      partial[(int64_t)blockIdx.x * C + (int64_t)cv * VEC + v] = a;
    }
  }
}

// db[c] = sum over chunks of partial[chunk][c]: one workgroup per channel,
// thread t adds chunks t, t + 256, ... in order, then a fixed binary tree
// combines the 256 thread sums.  (One thread per channel adding every chunk
// in turn was a chain of dependent loads: the whole adjoint took 220 us on a
// 64x56x56x128 BF16 layer, 784 chunks; now 31 us.)
template <typename A>
__global__ __launch_bounds__(256) void bias_grad_sum_kernel(const A *__restrict__ partial, int nchunks, int C,
                                                           A *__restrict__ db) {
  __shared__ A red[256];
  const int c = blockIdx.x;
  A a = (A)0;
  for (int k = threadIdx.x; k < nchunks; k += 256) a += partial[(int64_t)k * C + c];
  red[threadIdx.x] = a;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[c] = red[0];
}

// Number of pixel chunks the launch below uses (the partial buffer holds
// chunks x C values of A).
int bias_act_chunks(int64_t npix) {
  const int64_t PCH = 256;
  return (int)((npix + PCH - 1) / PCH);
}

template <typename T, typename A>
void run_bias_act_bwd(int act, int64_t npix, int C, const T *pre, const T *dy, T *dz, A *partial, A *db,
                      hipStream_t st) {
  constexpr int VEC = 16 / sizeof(T);
  const int PCH = 256;
  const int nchunks = bias_act_chunks(npix);
  const bool vec_ok = C % VEC == 0 && ((((uintptr_t)pre | (uintptr_t)dy | (uintptr_t)dz) & 15) == 0);
  const int vw = vec_ok ? VEC : 1;
  const int nv = C / vw;
  int CW = nv < 256 ? nv : 256;
  while (256 % CW != 0) --CW;  // lanes tile the block exactly
  dim3 grid((unsigned)nchunks, (unsigned)((nv + CW - 1) / CW));
  if (vec_ok)
    launch_k((bias_act_bwd_kernel<T, A, VEC>), grid, dim3(256), st, act, npix, C, CW, PCH, pre, dy, dz,
                       db ? partial : (A *)nullptr);
  else
    launch_k((bias_act_bwd_kernel<T, A, 1>), grid, dim3(256), st, act, npix, C, CW, PCH, pre, dy, dz,
                       db ? partial : (A *)nullptr);
  if (db)
    launch_k((bias_grad_sum_kernel<A>), dim3((unsigned)C), dim3(256), st, partial, nchunks, C, db);
}

// Adjoint of pooled = max_pool2d 2x2 stride 2 VALID (relu(pre)) for even H
// and W: a thread per (window, 16-byte channel vector) reads the window's
// four y and pre vectors and its dp vector, sends dp to the window's first
// strict maximum of y in Owl's order ((h, w), (h, w + 1), (h + 1, w),
// (h + 1, w + 1); dx starts at 0, so the cell receives 0 + dp), multiplies by
// Owl's relu factor and stores the four dz vectors; y = relu(pre) >= 0, so
// every window has a maximum.  Bias sums: per thread in window order (the
// four cells in that order), over lanes in lane order, then over chunks of
// WCH windows as in the plain adjoint.
template <typename T, typename A, int VEC>
__global__ __launch_bounds__(256) void maxpool_relu_bwd_kernel(int64_t nwin, int Wp, int C, int CW, int WCH,
                                                               const T *__restrict__ y, const T *__restrict__ pre,
                                                               const T *__restrict__ dp, T *__restrict__ dz,
                                                               A *__restrict__ partial) {
  __shared__ A red[256 * VEC];
  const int t = threadIdx.x;
  const int cw = t % CW, pl = t / CW, PL = 256 / CW;
  const int cv = blockIdx.y * CW + cw;
  const int nv = C / VEC;
  const int64_t w0 = (int64_t)blockIdx.x * WCH;
  int64_t w1 = w0 + WCH;
  if (w1 > nwin) w1 = nwin;
  const int W = 2 * Wp;
  A s[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) s[v] = (A)0;
  if (pl < PL && cv < nv) {
    const int64_t c0 = (int64_t)cv * VEC;
    for (int64_t wi = w0 + pl; wi < w1; wi += PL) {
      const int64_t bi = wi / Wp;  // b * Hp + i
      const int j = (int)(wi - bi * Wp);
      const int64_t q0 = (2 * bi) * W + 2 * j;  // (b, 2i, 2j): H = 2 Hp
      const int64_t cell[4] = {q0, q0 + 1, q0 + W, q0 + W + 1};
      T yv[4][VEC], zv[4][VEC], d[VEC];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        __builtin_memcpy(yv[k], y + cell[k] * C + c0, sizeof(yv[k]));
        __builtin_memcpy(zv[k], pre + cell[k] * C + c0, sizeof(zv[k]));
      }
      __builtin_memcpy(d, dp + wi * C + c0, sizeof(d));
      int arg[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        A best = -__builtin_inf();
        arg[v] = -1;
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (best < (A)yv[k][v]) {
            best = (A)yv[k][v];
            arg[v] = k;
          }
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        T o[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const A g = arg[v] == k ? (A)(T)((A)0 + (A)d[v]) : (A)0;
          o[v] = (T)adjoint_product<T, A>(g, relu_factor<A>((A)zv[k][v]));
          s[v] += (A)o[v];
        }
        __builtin_memcpy(dz + cell[k] * C + c0, o, sizeof(o));
      }
    }
  }
  if (!partial) return;
#pragma unroll
  for (int v = 0; v < VEC; ++v) red[t * VEC + v] = s[v];
  __syncthreads();
  if (pl == 0 && cv < nv) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      A a = (A)0;
      for (int l = 0; l < PL; ++l) a += red[(l * CW + cw) * VEC + v];
      partial[(int64_t)blockIdx.x * C + (int64_t)cv * VEC + v] = a;
    }
  }
}

int maxpool_relu_bwd_chunks(int64_t nwin) { return (int)((nwin + 63) / 64); }

template <typename T, typename A>
void run_maxpool_relu_bwd(int64_t B, int H, int W, int C, const T *y, const T *pre, const T *dp, T *dz, A *partial,
                          A *db, hipStream_t st) {
  constexpr int VEC = 16 / sizeof(T);
  const int WCH = 64;
  const int64_t nwin = B * (int64_t)(H / 2) * (W / 2);
  const int nchunks = maxpool_relu_bwd_chunks(nwin);
  const bool vec_ok =
      C % VEC == 0 && ((((uintptr_t)pre | (uintptr_t)y | (uintptr_t)dp | (uintptr_t)dz) & 15) == 0);
  const int nv = C / (vec_ok ? VEC : 1);
  int CW = nv < 256 ? nv : 256;
  while (256 % CW != 0) --CW;
  dim3 grid((unsigned)nchunks, (unsigned)((nv + CW - 1) / CW));
  if (vec_ok)
    launch_k((maxpool_relu_bwd_kernel<T, A, VEC>), grid, dim3(256), st, nwin, W / 2, C, CW, WCH, y, pre, dp, dz,
             db ? partial : (A *)nullptr);
  else
    launch_k((maxpool_relu_bwd_kernel<T, A, 1>), grid, dim3(256), st, nwin, W / 2, C, CW, WCH, y, pre, dp, dz,
             db ? partial : (A *)nullptr);
  if (db) launch_k((bias_grad_sum_kernel<A>), dim3((unsigned)C), dim3(256), st, partial, nchunks, C, db);
}

template void run_maxpool_relu_bwd<float, float>(int64_t, int, int, int, const float *, const float *,
                                                 const float *, float *, float *, float *, hipStream_t);
template void run_maxpool_relu_bwd<double, double>(int64_t, int, int, int, const double *, const double *,
                                                   const double *, double *, double *, double *, hipStream_t);
template void run_maxpool_relu_bwd<bf16_t, float>(int64_t, int, int, int, const bf16_t *, const bf16_t *,
                                                  const bf16_t *, bf16_t *, float *, float *, hipStream_t);
template void run_maxpool_relu_bwd<f16_t, float>(int64_t, int, int, int, const f16_t *, const f16_t *,
                                                 const f16_t *, f16_t *, float *, float *, hipStream_t);

template void run_bias_act_bwd<float, float>(int, int64_t, int, const float *, const float *, float *, float *,
                                             float *, hipStream_t);
template void run_bias_act_bwd<double, double>(int, int64_t, int, const double *, const double *, double *,
                                               double *, double *, hipStream_t);
template void run_bias_act_bwd<bf16_t, float>(int, int64_t, int, const bf16_t *, const bf16_t *, bf16_t *,
                                              float *, float *, hipStream_t);
template void run_bias_act_bwd<f16_t, float>(int, int64_t, int, const f16_t *, const f16_t *, f16_t *, float *,
                                             float *, hipStream_t);

}  // namespace owl_amd

namespace owl_amd {
// Zero fill of [p, p + bytes): 16-byte stores from the first 16-byte boundary,
// single bytes before it and after the last whole chunk.  Degenerate-shape
// results are cleared with it instead of hipMemsetAsync, so a HIP graph that
// captures such a call records a kernel node, never a memset node.
__global__ void zero_fill_kernel(unsigned char *p, uint64_t head, uint64_t chunks, uint64_t bytes) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t c = i; c < chunks; c += n) *(uint4 *)(p + head + 16 * c) = make_uint4(0, 0, 0, 0);
  // the head bytes and the tail bytes (fewer than 16 each)
  const uint64_t tail0 = head + 16 * chunks;
  if (i < head) p[i] = 0;
  if (i < bytes - tail0) p[tail0 + i] = 0;
}

void run_zero_fill(void *p, size_t bytes, hipStream_t st) {
  if (bytes == 0) return;
  unsigned char *b = (unsigned char *)p;
  uint64_t head = (16 - ((uintptr_t)b & 15)) & 15;
  if (head > bytes) head = bytes;
  const uint64_t chunks = (bytes - head) / 16;
  uint64_t blocks = (chunks + 255) / 256;
  if (blocks < 1) blocks = 1;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(zero_fill_kernel, dim3((unsigned)blocks), dim3(256), 0, st, b, head, chunks, (uint64_t)bytes);
}
}  // namespace owl_amd

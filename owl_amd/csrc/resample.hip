// 2-D nearest-neighbour upsampling of NHWC tensors by integer scales (sh, sw).
//
// Forward: y[b, i, j, c] = x[b, i / sh, j / sw, c] -- Owl computes it as
// repeat x [|1; sh; sw; 1|] (owl_dense_ndarray_generic.ml:8223-8234).
//
// Backward (owl_ndarray_upsampling_impl.h:9-60): dx[b, hi, wi, c] is the sum of
// dy over the output cells that map to (hi, wi), where output row h maps to
// min(h / sh, H - 1) and column w to min(w / sw, W - 1).  Owl scatters in the
// order h ascending, then w ascending, into a zeroed dx; each thread here
// gathers its input cell's cells in that same order starting from zero, so the
// sums have Owl's rounding exactly (in the element type; 16-bit types add in
// FP32 and round once).
#include "owl_amd_common.h"
#include "conv_kernels.h"

namespace owl_amd {

template <typename T> struct AccOf { typedef T type; };
template <> struct AccOf<bf16_t> { typedef float type; };
template <> struct AccOf<f16_t> { typedef float type; };

// one thread per VEC channels of one output cell
template <typename T, int VEC>
__global__ __launch_bounds__(256) void upsample_fwd_kernel(const T *__restrict__ x, T *__restrict__ y, int64_t B,
                                                          int H, int W, int C, int Ho, int Wo, int sh, int sw) {
  const int cv = C / VEC;
  const int64_t total = B * Ho * Wo * cv;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int c = (int)(t % cv) * VEC;
    int64_t r = t / cv;
    const int j = (int)(r % Wo);
    r /= Wo;
    const int i = (int)(r % Ho);
    const int64_t b = r / Ho;
    int hi = i / sh, wi = j / sw;
    if (hi > H - 1) hi = H - 1;
    if (wi > W - 1) wi = W - 1;
    T v[VEC];
    __builtin_memcpy(v, x + ((b * H + hi) * W + wi) * C + c, sizeof(v));
    __builtin_memcpy(y + ((b * Ho + i) * Wo + j) * C + c, v, sizeof(v));
  }
}

template <typename T, int VEC>
__global__ __launch_bounds__(256) void upsample_bwd_kernel(const T *__restrict__ dy, T *__restrict__ dx, int64_t B,
                                                          int H, int W, int C, int Ho, int Wo, int sh, int sw) {
  typedef typename AccOf<T>::type A;
  const int cv = C / VEC;
  const int64_t total = B * H * W * cv;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (int64_t)gridDim.x * 256) {
    const int c = (int)(t % cv) * VEC;
    int64_t r = t / cv;
    const int wi = (int)(r % W);
    r /= W;
    const int hi = (int)(r % H);
    const int64_t b = r / H;
    // output rows / columns mapping to this cell; the last row and column
    // also take every output beyond the scaled extent (Owl's clamp)
    const int h0 = hi * sh, w0 = wi * sw;
    int h1 = hi == H - 1 ? Ho : h0 + sh, w1 = wi == W - 1 ? Wo : w0 + sw;
    if (h1 > Ho) h1 = Ho;
    if (w1 > Wo) w1 = Wo;
    A s[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) s[v] = (A)0;
    for (int h = h0; h < h1; ++h)
      for (int w = w0; w < w1; ++w) {
        T g[VEC];
        __builtin_memcpy(g, dy + ((b * Ho + h) * Wo + w) * C + c, sizeof(g));
#pragma unroll
        for (int v = 0; v < VEC; ++v) s[v] += (A)g[v];
      }
    T o[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) o[v] = (T)s[v];
    __builtin_memcpy(dx + ((b * H + hi) * W + wi) * C + c, o, sizeof(o));
  }
}

template <typename T>
void run_upsample(bool backward, const T *src, T *dst, int64_t B, int H, int W, int C, int Ho, int Wo, int sh,
                  int sw, hipStream_t st) {
  constexpr int VEC = 16 / sizeof(T);
  const bool vec = C % VEC == 0 && ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0);
  const int64_t cells = backward ? B * H * W : B * Ho * Wo;
  const int64_t work = cells * (vec ? C / VEC : C);
  int64_t blocks = (work + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  if (blocks < 1) blocks = 1;
  dim3 grid((unsigned)blocks), blk(256);
  if (backward) {
    if (vec) launch_k((upsample_bwd_kernel<T, VEC>), grid, blk, st, src, dst, B, H, W, C, Ho, Wo, sh, sw);
    else launch_k((upsample_bwd_kernel<T, 1>), grid, blk, st, src, dst, B, H, W, C, Ho, Wo, sh, sw);
  } else {
    if (vec) launch_k((upsample_fwd_kernel<T, VEC>), grid, blk, st, src, dst, B, H, W, C, Ho, Wo, sh, sw);
    else launch_k((upsample_fwd_kernel<T, 1>), grid, blk, st, src, dst, B, H, W, C, Ho, Wo, sh, sw);
  }
}

template void run_upsample<float>(bool, const float *, float *, int64_t, int, int, int, int, int, int, int,
                                  hipStream_t);
template void run_upsample<double>(bool, const double *, double *, int64_t, int, int, int, int, int, int, int,
                                   hipStream_t);
template void run_upsample<bf16_t>(bool, const bf16_t *, bf16_t *, int64_t, int, int, int, int, int, int, int,
                                   hipStream_t);
template void run_upsample<f16_t>(bool, const f16_t *, f16_t *, int64_t, int, int, int, int, int, int, int,
                                  hipStream_t);

}  // namespace owl_amd

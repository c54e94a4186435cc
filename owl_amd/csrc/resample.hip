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
// Strided scatter of rows of row_bytes bytes: dst[b, h, w, d] = src[b, h / sh,
// w / sw, d / sd] where h, w and d are multiples of the strides (and the
// source cell exists), zero elsewhere.  This is the backward input of a 1x1
// convolution with strides and no padding, whose dx is dy . w^T on the
// strided grid and zero off it, built from the stride-1 backward input of
// the subsampled grid.  A workgroup writes one dst row (b, h) of W x D
// pixels, each lane whole U-sized units: with upr >= 64 units per pixel the
// lanes of a wave stride one pixel, else each wave covers 64 / upr pixels
// per step.  The row's grid test is made once; the (w, d) test once per
// pixel and lane (a per-unit index decomposition by division cost more time
// than the memory traffic).
template <typename U>
__global__ __launch_bounds__(256) void stride_scatter_kernel(const U *__restrict__ src, U *__restrict__ dst, int H,
                                                             int W, int D, int Hs, int Ws, int Ds, int sh, int sw,
                                                             int sd, int upr) {
  const int bh = blockIdx.x;  // b * H + h
  const int b = bh / H, h = bh - b * H;
  const int hs = h / sh;
  const bool hon = h == hs * sh && hs < Hs;
  const int npix = W * D;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ppw = upr >= 64 ? 1 : 64 / upr;  // pixels per wave step
  const int sp = upr >= 64 ? 0 : lane / upr;
  const int u0 = upr >= 64 ? lane : lane - sp * upr;
  const bool lane_ok = upr >= 64 || sp < ppw;
  U *drow = dst + (int64_t)bh * npix * upr;
  const U *srow = src + ((int64_t)b * Hs + (hon ? hs : 0)) * Ws * Ds * upr;
  for (int p0 = wv * ppw; p0 < npix; p0 += 4 * ppw) {
    const int pix = p0 + sp;
    if (!lane_ok || pix >= npix) continue;
    int w = pix, d = 0;
    if (D != 1) {
      w = pix / D;
      d = pix - w * D;
    }
    const int ws = w / sw, ds = sd == 1 ? d : d / sd;
    const bool on = hon && w == ws * sw && d == ds * sd && ws < Ws && ds < Ds;
    for (int u = u0; u < upr; u += 64) {
      U v;
      __builtin_memset(&v, 0, sizeof(U));
      if (on) v = srow[((int64_t)ws * Ds + ds) * upr + u];
      drow[(int64_t)pix * upr + u] = v;
    }
  }
}

template <typename U>
static void launch_scatter(const void *src, void *dst, int64_t B, int Hs, int Ws, int Ds, int H, int W, int D, int sh,
                           int sw, int sd, int upr, hipStream_t st) {
  if (B * H == 0) return;
  launch_k((stride_scatter_kernel<U>), dim3((unsigned)(B * H)), dim3(256), st, (const U *)src, (U *)dst, H, W, D, Hs,
           Ws, Ds, sh, sw, sd, upr);
}

// Phase interleave for the backward input of a strided convolution: dx is
// the union of s_h s_w s_d phase grids, phase (c_h, c_w, c_d) holding the
// pixels (s_h m + c_h, s_w n + c_w, s_d l + c_d) as a dense [B][Mh][Mw][Md]
// grid of rows of row_bytes bytes (at byte offset off in src; off < 0: the
// phase has no kernel taps and its pixels are zero).  A workgroup copies one
// source row (phase, b, m); the phase is found from the workgroup index by a
// uniform scan of the phases' first workgroups, so every lookup is scalar.
template <typename U>
__global__ __launch_bounds__(256) void phase_interleave_kernel(const char *__restrict__ src, U *__restrict__ dst, int H,
                                                               int W, int D, int sh, int sw, int sd, int upr,
                                                               PhaseMap pm) {
  const int64_t bid = blockIdx.x;
  int p = 0;
  while (p + 1 < pm.n && bid >= pm.first[p + 1]) ++p;
  const int Mh = pm.mh[p], Mw = pm.mw[p], Md = pm.md[p];
  const int64_t r = bid - pm.first[p];  // b * Mh + m
  const int64_t b = r / Mh;
  const int m = (int)(r - b * Mh);
  const int h = sh * m + pm.ch[p];
  const int npix = Mw * Md;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ppw = upr >= 64 ? 1 : 64 / upr;
  const int sp = upr >= 64 ? 0 : lane / upr;
  const int u0 = upr >= 64 ? lane : lane - sp * upr;
  const bool lane_ok = upr >= 64 || sp < ppw;
  const bool zero = pm.off[p] < 0;
  const U *srow = (const U *)(src + (zero ? 0 : pm.off[p])) + r * npix * upr;
  U *drow = dst + ((int64_t)b * H + h) * W * D * upr;
  for (int p0 = wv * ppw; p0 < npix; p0 += 4 * ppw) {
    const int pix = p0 + sp;
    if (!lane_ok || pix >= npix) continue;
    int n = pix, l = 0;
    if (Md != 1) {
      n = pix / Md;
      l = pix - n * Md;
    }
    const int64_t dpix = (int64_t)(sw * n + pm.cw[p]) * D + sd * l + pm.cd[p];
    for (int u = u0; u < upr; u += 64) {
      U v;
      __builtin_memset(&v, 0, sizeof(U));
      if (!zero) v = srow[(int64_t)pix * upr + u];
      drow[dpix * upr + u] = v;
    }
  }
}

template <typename U>
static void launch_interleave(const void *src, void *dst, int H, int W, int D, int sh, int sw, int sd, int upr,
                              const PhaseMap &pm, hipStream_t st) {
  const int64_t blocks = pm.first[pm.n];
  if (blocks <= 0) return;
  launch_k((phase_interleave_kernel<U>), dim3((unsigned)blocks), dim3(256), st, (const char *)src, (U *)dst, H, W, D,
           sh, sw, sd, upr, pm);
}

void run_phase_interleave(const void *src, void *dst, int H, int W, int D, int sh, int sw, int sd, int64_t row_bytes,
                          const PhaseMap &pm, hipStream_t st) {
  uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  for (int q = 0; q < pm.n; ++q)
    if (pm.off[q] > 0) al |= (uintptr_t)pm.off[q];
  for (int ub : {16, 8, 4}) {
    if (row_bytes % ub == 0 && al % ub == 0) {
      const int upr = (int)(row_bytes / ub);
      if (ub == 16) launch_interleave<uint4>(src, dst, H, W, D, sh, sw, sd, upr, pm, st);
      else if (ub == 8) launch_interleave<uint2>(src, dst, H, W, D, sh, sw, sd, upr, pm, st);
      else launch_interleave<uint32_t>(src, dst, H, W, D, sh, sw, sd, upr, pm, st);
      return;
    }
  }
  launch_interleave<uint16_t>(src, dst, H, W, D, sh, sw, sd, (int)(row_bytes / 2), pm, st);
}

// Stride-phase gather: dst[b, m, n, l] = src[b, sh m + ch, sw n + cw, sd l + cd]
// (rows of row_bytes bytes; dst is the dense [B][Mh][Mw][Md] grid of one
// phase).  A workgroup copies one dst row (b, m), lanes over U-sized units as
// in the scatter above.
template <typename U>
__global__ __launch_bounds__(256) void stride_gather_kernel(const U *__restrict__ src, U *__restrict__ dst, int H,
                                                            int W, int D, int Mh, int Mw, int Md, int sh, int sw,
                                                            int sd, int ch, int cw, int cd, int upr) {
  const int64_t bm = blockIdx.x;  // b * Mh + m
  const int64_t b = bm / Mh;
  const int m = (int)(bm - b * Mh);
  const int h = sh * m + ch;
  const int npix = Mw * Md;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ppw = upr >= 64 ? 1 : 64 / upr;
  const int sp = upr >= 64 ? 0 : lane / upr;
  const int u0 = upr >= 64 ? lane : lane - sp * upr;
  const bool lane_ok = upr >= 64 || sp < ppw;
  const U *srow = src + ((int64_t)b * H + h) * W * D * upr;
  U *drow = dst + bm * npix * upr;
  for (int p0 = wv * ppw; p0 < npix; p0 += 4 * ppw) {
    const int pix = p0 + sp;
    if (!lane_ok || pix >= npix) continue;
    int n = pix, l = 0;
    if (Md != 1) {
      n = pix / Md;
      l = pix - n * Md;
    }
    const int64_t spix = (int64_t)(sw * n + cw) * D + sd * l + cd;
    for (int u = u0; u < upr; u += 64) drow[(int64_t)pix * upr + u] = srow[spix * upr + u];
  }
}

template <typename U>
static void launch_gather(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                          int sw, int sd, int ch, int cw, int cd, int upr, hipStream_t st) {
  if (B * Mh == 0) return;
  launch_k((stride_gather_kernel<U>), dim3((unsigned)(B * Mh)), dim3(256), st, (const U *)src, (U *)dst, H, W, D, Mh,
           Mw, Md, sh, sw, sd, ch, cw, cd, upr);
}

void run_stride_gather(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                       int sw, int sd, int ch, int cw, int cd, int64_t row_bytes, hipStream_t st) {
  const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  for (int ub : {16, 8, 4}) {
    if (row_bytes % ub == 0 && al % ub == 0) {
      const int upr = (int)(row_bytes / ub);
      if (ub == 16) launch_gather<uint4>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, ch, cw, cd, upr, st);
      else if (ub == 8) launch_gather<uint2>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, ch, cw, cd, upr, st);
      else launch_gather<uint32_t>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, ch, cw, cd, upr, st);
      return;
    }
  }
  launch_gather<uint16_t>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, ch, cw, cd, (int)(row_bytes / 2), st);
}

// Space-to-depth: dst[b, m, n, l, ((c_h sw + c_w) sd + c_d) C + c] =
// src[b, sh m + c_h, sw n + c_w, sd l + c_d, c] (zero past the image), dst
// the [B][Mh][Mw][Md][sh sw sd C] grid; rows of row_bytes = C element bytes
// per phase.  A workgroup writes one dst row (b, m); each wave step covers
// one (pixel, phase) chunk run.
template <typename U>
__global__ __launch_bounds__(256) void space_to_depth_kernel(const U *__restrict__ src, U *__restrict__ dst, int H,
                                                             int W, int D, int Mh, int Mw, int Md, int sh, int sw,
                                                             int sd, int upr) {
  const int64_t bm = blockIdx.x;  // b * Mh + m
  const int64_t b = bm / Mh;
  const int m = (int)(bm - b * Mh);
  const int nph = sh * sw * sd;
  const int nitems = Mw * Md * nph;  // (pixel, phase) pairs of the row
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int ppw = upr >= 64 ? 1 : 64 / upr;
  const int sp = upr >= 64 ? 0 : lane / upr;
  const int u0 = upr >= 64 ? lane : lane - sp * upr;
  const bool lane_ok = upr >= 64 || sp < ppw;
  U *drow = dst + bm * (int64_t)nitems * upr;
  for (int p0 = wv * ppw; p0 < nitems; p0 += 4 * ppw) {
    const int it = p0 + sp;
    if (!lane_ok || it >= nitems) continue;
    const int pix = it / nph, ph = it - pix * nph;
    int n = pix, l = 0;
    if (Md != 1) {
      n = pix / Md;
      l = pix - n * Md;
    }
    int ch = ph, cw = 0, cd = 0;
    if (nph != sh) {
      cd = ph % sd;
      const int r = ph / sd;
      cw = r % sw;
      ch = r / sw;
    }
    const int h = sh * m + ch, w = sw * n + cw, d = sd * l + cd;
    const bool in = h < H && w < W && d < D;
    const U *srow = src + (((int64_t)b * H + h) * W + w) * (int64_t)D * upr + (int64_t)d * upr;
    for (int u = u0; u < upr; u += 64) {
      U v;
      __builtin_memset(&v, 0, sizeof(U));
      if (in) v = srow[u];
      drow[(int64_t)it * upr + u] = v;
    }
  }
}

template <typename U>
static void launch_s2d(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                       int sw, int sd, int upr, hipStream_t st) {
  if (B * Mh == 0) return;
  launch_k((space_to_depth_kernel<U>), dim3((unsigned)(B * Mh)), dim3(256), st, (const U *)src, (U *)dst, H, W, D, Mh,
           Mw, Md, sh, sw, sd, upr);
}

void run_space_to_depth(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                        int sw, int sd, int64_t row_bytes, hipStream_t st) {
  const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  for (int ub : {16, 8, 4}) {
    if (row_bytes % ub == 0 && al % ub == 0) {
      const int upr = (int)(row_bytes / ub);
      if (ub == 16) launch_s2d<uint4>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, upr, st);
      else if (ub == 8) launch_s2d<uint2>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, upr, st);
      else launch_s2d<uint32_t>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, upr, st);
      return;
    }
  }
  launch_s2d<uint16_t>(src, dst, B, H, W, D, Mh, Mw, Md, sh, sw, sd, (int)(row_bytes / 2), st);
}

// Kernel for the space-to-depth convolution: dst[t][ph][ci][co] = src tap
// tap_of[t * nph + ph] ([ci][co] block), or zeros where it is -1.
__global__ __launch_bounds__(256) void s2d_kernel_relayout(const uint16_t *__restrict__ src, uint16_t *__restrict__ dst,
                                                           int upb, int wg_per_block, S2dTaps tm) {
  const int j = blockIdx.x / wg_per_block;
  const int64_t u = (int64_t)(blockIdx.x - j * wg_per_block) * 256 + threadIdx.x;
  if (u >= upb) return;
  const int t = tm.tap[j];
  dst[(int64_t)j * upb + u] = t >= 0 ? src[(int64_t)t * upb + u] : (uint16_t)0;
}

void run_s2d_kernel(const void *src, void *dst, int64_t block_bytes, const S2dTaps &tm, hipStream_t st) {
  const int upb = (int)(block_bytes / 2);
  const int wpb = (upb + 255) / 256;
  if (tm.n <= 0 || wpb <= 0) return;
  launch_k(s2d_kernel_relayout, dim3((unsigned)(tm.n * wpb)), dim3(256), st, (const uint16_t *)src, (uint16_t *)dst,
           upb, wpb, tm);
}

// Tap-block gather: block j of dst (block_bytes bytes) = block src_tap[j] of
// src; a workgroup copies 256 16-byte chunks (or 4-byte words) of one block.
template <typename U>
__global__ __launch_bounds__(256) void tap_gather_kernel(const U *__restrict__ src, U *__restrict__ dst, int upb,
                                                         int wg_per_block, TapMap tm) {
  const int j = blockIdx.x / wg_per_block;
  const int64_t u = (int64_t)(blockIdx.x - j * wg_per_block) * 256 + threadIdx.x;
  if (u < upb) dst[(int64_t)j * upb + u] = src[(int64_t)tm.src[j] * upb + u];
}

void run_tap_gather(const void *src, void *dst, int64_t block_bytes, const TapMap &tm, hipStream_t st) {
  const bool v16 = block_bytes % 16 == 0 && (((uintptr_t)src | (uintptr_t)dst) & 15) == 0;
  const int ub = v16 ? 16 : 2;
  const int upb = (int)(block_bytes / ub);
  const int wpb = (upb + 255) / 256;
  if (tm.n <= 0 || wpb <= 0) return;
  if (v16)
    launch_k((tap_gather_kernel<uint4>), dim3((unsigned)(tm.n * wpb)), dim3(256), st, (const uint4 *)src, (uint4 *)dst,
             upb, wpb, tm);
  else
    launch_k((tap_gather_kernel<uint16_t>), dim3((unsigned)(tm.n * wpb)), dim3(256), st, (const uint16_t *)src,
             (uint16_t *)dst, upb, wpb, tm);
}

void run_stride_scatter(const void *src, void *dst, int64_t B, int Hs, int Ws, int Ds, int H, int W, int D, int sh,
                        int sw, int sd, int64_t row_bytes, hipStream_t st) {
  const uintptr_t al = (uintptr_t)src | (uintptr_t)dst;
  // the widest unit that divides a row and both addresses
  for (int ub : {16, 8, 4}) {
    if (row_bytes % ub == 0 && al % ub == 0) {
      const int upr = (int)(row_bytes / ub);
      if (ub == 16) launch_scatter<uint4>(src, dst, B, Hs, Ws, Ds, H, W, D, sh, sw, sd, upr, st);
      else if (ub == 8) launch_scatter<uint2>(src, dst, B, Hs, Ws, Ds, H, W, D, sh, sw, sd, upr, st);
      else launch_scatter<uint32_t>(src, dst, B, Hs, Ws, Ds, H, W, D, sh, sw, sd, upr, st);
      return;
    }
  }
  launch_scatter<uint16_t>(src, dst, B, Hs, Ws, Ds, H, W, D, sh, sw, sd, (int)(row_bytes / 2), st);
}

template void run_upsample<double>(bool, const double *, double *, int64_t, int, int, int, int, int, int, int,
                                   hipStream_t);
template void run_upsample<bf16_t>(bool, const bf16_t *, bf16_t *, int64_t, int, int, int, int, int, int, int,
                                   hipStream_t);
template void run_upsample<f16_t>(bool, const f16_t *, f16_t *, int64_t, int, int, int, int, int, int, int,
                                  hipStream_t);

}  // namespace owl_amd

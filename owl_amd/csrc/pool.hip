// Max / average pooling for gfx950, bit-identical to Owl's C
// (owl_ndarray_pool_impl.h with the macros of owl_ndarray_pool_stub.c:16-51).
//
// These are bandwidth kernels: one thread per output cell (forward) or per
// input cell (backward), each covering VEC consecutive channels so a wave
// reads and writes contiguous 16-byte chunks of the NHWC / NDHWC arrays.
// Within a cell the arithmetic is Owl's, in Owl's order:
//   * the window is scanned H ("cols") outer, W ("rows"), D ("dpts") inner;
//   * max: acc = -inf; if (acc < t) acc = t   -> first strict max, NaN never wins;
//   * avg: acc = 0; acc += t; acc / count_in_bounds  (FP32 stays FP32);
//   * BF16 / FP16 (device entry points only; Owl's pooling is FP32 / FP64)
//     compare and add in FP32 and round once;
//   * backward is a gather: an input cell adds the contributions of the
//     windows that cover it in ascending (i, j, l) order, the order in which
//     Owl's scatter loop (:139-198, :377-440) adds them, starting from 0.
// Owl's max backward sends the gradient of a window with no candidate (all
// NaN / -inf) to flat index 0 of the whole input (max_idx starts at 0,
// :167); such windows raise a flag and a one-workgroup fix-up kernel rebuilds
// element 0 in Owl's order.
#include "owl_amd_common.h"
#include "pool_kernels.h"

namespace owl_amd {

// Non-temporal store of VEC values of T (16 bytes when vectorised, else one
// value), for the backward passes' dx and the argmax indices: written once,
// they stream past L2 (max-pool backward 6310 -> 6368 GB/s, argmax 4348 ->
// 4467, average-pool backward 2250 -> 2308 on a 64x56x56x128 layer).
typedef unsigned int pool_u32x4 __attribute__((ext_vector_type(4)));
template <typename T, int VEC>
__device__ __forceinline__ void store_nt(T *dst, const T (&v)[VEC]) {
  if constexpr (sizeof(T) * VEC == 16) {
    pool_u32x4 t;
    __builtin_memcpy(&t, v, 16);
    __builtin_nontemporal_store(t, (pool_u32x4 *)dst);
  } else if constexpr (VEC == 1 && sizeof(T) == 2) {
    __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, v[0]), (unsigned short *)dst);
  } else if constexpr (VEC == 1) {
    __builtin_nontemporal_store(v[0], dst);
  } else {
    __builtin_memcpy(dst, v, sizeof(T) * VEC);
  }
}

template <typename T> struct PoolAcc { typedef float type; };
template <> struct PoolAcc<double> { typedef double type; };

template <typename T> __device__ __forceinline__ T neg_inf();
template <> __device__ __forceinline__ float neg_inf<float>() { return -__builtin_inff(); }
template <> __device__ __forceinline__ double neg_inf<double>() { return -__builtin_inf(); }

struct Win {
  int lo, hi;  // in-bounds index range [lo, hi) of the window along one axis
};

__device__ __forceinline__ Win window(int o, int s, int p, int k, int n) {
  int st = o * s - p;
  int lo = st < 0 ? 0 : st;
  int hi = st + k > n ? n : st + k;
  return {lo, hi};
}

// ---------------------------------------------------------------- forward
template <typename T, int VEC, int MODE>  // MODE 0 avg, 1 max, 2 max + argmax
__global__ __launch_bounds__(256) void pool_fwd_kernel(PoolParams p, const T *__restrict__ x,
                                                       T *__restrict__ y,
                                                       int64_t *__restrict__ arg) {
  typedef typename PoolAcc<T>::type A;
  const int CV = p.C / VEC;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)p.B * p.Ho * p.Wo * p.Do * CV;
  if (gid >= total) return;
  int cv = (int)(gid % CV);
  int64_t t = gid / CV;
  int l = (int)(t % p.Do); t /= p.Do;
  int j = (int)(t % p.Wo); t /= p.Wo;
  int i = (int)(t % p.Ho);
  int64_t b = t / p.Ho;
  Win wh = window(i, p.sh, p.ph, p.kh, p.H);
  Win ww = window(j, p.sw, p.pw, p.kw, p.W);
  Win wd = window(l, p.sd, p.pd, p.kd, p.D);
  const int c0 = cv * VEC;
  A acc[VEC];
  int64_t mi[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    acc[v] = MODE == 0 ? (A)0 : (A)neg_inf<A>();
    mi[v] = -1;
  }
  for (int h = wh.lo; h < wh.hi; ++h)
    for (int w = ww.lo; w < ww.hi; ++w)
      for (int d = wd.lo; d < wd.hi; ++d) {
        int64_t id = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.C + c0;
        T tv[VEC];
        __builtin_memcpy(tv, x + id, sizeof(tv));
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          A tval = (A)tv[v];
          if (MODE == 0) {
            acc[v] += tval;
          } else if (acc[v] < tval) {
            acc[v] = tval;
            mi[v] = id + v;
          }
        }
      }
  int cnt = (wh.hi - wh.lo) * (ww.hi - ww.lo) * (wd.hi - wd.lo);
  if (wh.hi < wh.lo || ww.hi < ww.lo || wd.hi < wd.lo) cnt = 0;
  int64_t od = gid * VEC;
  T out[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) out[v] = (T)(MODE == 0 ? acc[v] / (A)cnt : acc[v]);
  // plain stores: non-temporal ones measured 3% slower here (6336 against
  // 6546 GB/s on a 64x56x56x128 max-pool)
  __builtin_memcpy(y + od, out, sizeof(out));
  if constexpr (MODE == 2) {
    if constexpr (VEC % 2 == 0) {
      // VEC int64 indices as 16-byte non-temporal stores
#pragma unroll
      for (int v = 0; v < VEC; v += 2) {
        pool_u32x4 t;
        __builtin_memcpy(&t, &mi[v], 16);
        __builtin_nontemporal_store(t, (pool_u32x4 *)(arg + od + v));
      }
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) __builtin_nontemporal_store(mi[v], arg + od + v);
    }
  }
}

// ---------------------------------------------------------------- backward
// Range of output indices o whose window covers input index h along one axis:
// o*s - p <= h < o*s - p + k  ->  ceil((h + p - k + 1) / s) <= o <= floor((h + p) / s).
__device__ __forceinline__ void covering(int h, int s, int p, int k, int O, int &lo, int &hi) {
  int num = h + p - k + 1;
  lo = num <= 0 ? 0 : (num + s - 1) / s;
  int top = h + p;
  hi = top < 0 ? -1 : top / s;
  if (hi > O - 1) hi = O - 1;
}

template <typename T, int VEC, bool IS_MAX>
__global__ __launch_bounds__(256) void pool_bwd_kernel(PoolParams p, const T *__restrict__ x,
                                                       const T *__restrict__ dy,
                                                       T *__restrict__ dx, int *orphan_flag) {
  typedef typename PoolAcc<T>::type A;
  const int CV = p.C / VEC;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)p.B * p.H * p.W * p.D * CV;
  if (gid >= total) return;
  int cv = (int)(gid % CV);
  int64_t t = gid / CV;
  int d = (int)(t % p.D); t /= p.D;
  int w = (int)(t % p.W); t /= p.W;
  int h = (int)(t % p.H);
  int64_t b = t / p.H;
  const int c0 = cv * VEC;
  const int64_t my_id = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.C + c0;

  int ilo, ihi, jlo, jhi, llo, lhi;
  covering(h, p.sh, p.ph, p.kh, p.Ho, ilo, ihi);
  covering(w, p.sw, p.pw, p.kw, p.Wo, jlo, jhi);
  covering(d, p.sd, p.pd, p.kd, p.Do, llo, lhi);

  A acc[VEC];  // the element type for FP32 / FP64 (Owl's arithmetic); FP32 for 16-bit
#pragma unroll
  for (int v = 0; v < VEC; ++v) acc[v] = (A)0;

  for (int i = ilo; i <= ihi; ++i) {
    Win wh = window(i, p.sh, p.ph, p.kh, p.H);
    for (int j = jlo; j <= jhi; ++j) {
      Win ww = window(j, p.sw, p.pw, p.kw, p.W);
      for (int l = llo; l <= lhi; ++l) {
        Win wd = window(l, p.sd, p.pd, p.kd, p.D);
        int64_t od = (((b * p.Ho + i) * (int64_t)p.Wo + j) * p.Do + l) * (int64_t)p.C + c0;
        T g[VEC];
        __builtin_memcpy(g, dy + od, sizeof(g));
        if (IS_MAX) {
          A best[VEC];
          int64_t bi[VEC];
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            best[v] = neg_inf<A>();
            bi[v] = -1;
          }
          for (int hh = wh.lo; hh < wh.hi; ++hh)
            for (int w2 = ww.lo; w2 < ww.hi; ++w2)
              for (int dd = wd.lo; dd < wd.hi; ++dd) {
                int64_t id = (((b * p.H + hh) * (int64_t)p.W + w2) * p.D + dd) * (int64_t)p.C + c0;
                T tv[VEC];
                __builtin_memcpy(tv, x + id, sizeof(tv));
#pragma unroll
                for (int v = 0; v < VEC; ++v)
                  if (best[v] < (A)tv[v]) {
                    best[v] = (A)tv[v];
                    bi[v] = id + v;
                  }
              }
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            if (bi[v] == my_id + v) acc[v] += (A)g[v];
            if (bi[v] < 0) *orphan_flag = 1;  // Owl sends this gradient to index 0
          }
        } else {
          int cnt = (wh.hi - wh.lo) * (ww.hi - ww.lo) * (wd.hi - wd.lo);
#pragma unroll
          for (int v = 0; v < VEC; ++v) acc[v] += (A)g[v] / (A)cnt;
        }
      }
    }
  }
  T o[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) o[v] = (T)acc[v];
  store_nt<T, VEC>(dx + my_id, o);
}

// Max backward when the windows tile the input exactly (window == stride,
// no padding, every cell in exactly one window): one thread per window and
// VEC channels reads the window once and writes all of its cells (dy at the
// first strict max, 0 elsewhere).  Each cell then receives Owl's single
// addition 0 + dy, so the result is the same bits as the gather kernel.
template <typename T, int VEC>
__global__ __launch_bounds__(256) void maxpool_bwd_tiled_kernel(PoolParams p, const T *__restrict__ x,
                                                                const T *__restrict__ dy,
                                                                T *__restrict__ dx, int *orphan_flag) {
  typedef typename PoolAcc<T>::type A;
  const int CV = p.C / VEC;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)p.B * p.Ho * p.Wo * p.Do * CV;
  if (gid >= total) return;
  int cv = (int)(gid % CV);
  int64_t t = gid / CV;
  int l = (int)(t % p.Do); t /= p.Do;
  int j = (int)(t % p.Wo); t /= p.Wo;
  int i = (int)(t % p.Ho);
  int64_t b = t / p.Ho;
  const int c0 = cv * VEC;
  A best[VEC];
  int widx[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) {
    best[v] = neg_inf<A>();
    widx[v] = -1;
  }
  int n = 0;
  for (int h = i * p.sh; h < i * p.sh + p.kh; ++h)
    for (int w = j * p.sw; w < j * p.sw + p.kw; ++w)
      for (int d = l * p.sd; d < l * p.sd + p.kd; ++d, ++n) {
        int64_t id = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.C + c0;
        T tv[VEC];
        __builtin_memcpy(tv, x + id, sizeof(tv));
#pragma unroll
        for (int v = 0; v < VEC; ++v)
          if (best[v] < (A)tv[v]) {
            best[v] = (A)tv[v];
            widx[v] = n;
          }
      }
  T g[VEC];
  __builtin_memcpy(g, dy + gid * VEC, sizeof(g));
#pragma unroll
  for (int v = 0; v < VEC; ++v)
    if (widx[v] < 0) *orphan_flag = 1;  // Owl sends this gradient to index 0
  n = 0;
  for (int h = i * p.sh; h < i * p.sh + p.kh; ++h)
    for (int w = j * p.sw; w < j * p.sw + p.kw; ++w)
      for (int d = l * p.sd; d < l * p.sd + p.kd; ++d, ++n) {
        int64_t id = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.C + c0;
        T o[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) o[v] = widx[v] == n ? (T)((T)0 + g[v]) : (T)0;
        store_nt<T, VEC>(dx + id, o);
      }
}

// Average backward when the windows tile the input exactly: one thread per
// window and VEC channels reads dy once and writes every cell of the window
// with Owl's single addition 0 + dy / count (count = the whole window, all
// cells in bounds), the same bits as the gather kernel (average-pool 2x2x2
// backward on an 8x32x32x32x64 FP32 layer: the gather ran at 1.9 TB/s).
template <typename T, int VEC>
__global__ __launch_bounds__(256) void avgpool_bwd_tiled_kernel(PoolParams p, const T *__restrict__ dy,
                                                                T *__restrict__ dx) {
  typedef typename PoolAcc<T>::type A;
  const int CV = p.C / VEC;
  int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = (int64_t)p.B * p.Ho * p.Wo * p.Do * CV;
  if (gid >= total) return;
  int cv = (int)(gid % CV);
  int64_t t = gid / CV;
  int l = (int)(t % p.Do); t /= p.Do;
  int j = (int)(t % p.Wo); t /= p.Wo;
  int i = (int)(t % p.Ho);
  int64_t b = t / p.Ho;
  const int c0 = cv * VEC;
  const A cnt = (A)(p.kh * p.kw * p.kd);
  T g[VEC], o[VEC];
  __builtin_memcpy(g, dy + gid * VEC, sizeof(g));
#pragma unroll
  for (int v = 0; v < VEC; ++v) o[v] = (T)((A)0 + (A)g[v] / cnt);
  for (int h = i * p.sh; h < i * p.sh + p.kh; ++h)
    for (int w = j * p.sw; w < j * p.sw + p.kw; ++w)
      for (int d = l * p.sd; d < l * p.sd + p.kd; ++d) {
        int64_t id = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.C + c0;
        store_nt<T, VEC>(dx + id, o);
      }
}

// Rebuild dx[0] in Owl's order when some window had no max candidate.  One
// workgroup walks the windows in Owl's loop order (b, i, j, l, c) 256 at a
// time; thread 0 adds the flagged gradients in that order.
template <typename T>
__global__ __launch_bounds__(256) void pool_bwd_fixup_kernel(PoolParams p, const T *__restrict__ x,
                                                             const T *__restrict__ dy,
                                                             T *__restrict__ dx, int *orphan_flag) {
  typedef typename PoolAcc<T>::type A;
  if (*orphan_flag == 0) return;
  __shared__ int hit[256];
  __shared__ A val[256];
  int64_t total = (int64_t)p.B * p.Ho * p.Wo * p.Do * p.C;
  A acc = (A)0;
  for (int64_t base = 0; base < total; base += 256) {
    int64_t gid = base + threadIdx.x;
    int flag = 0;
    A g = (A)0;
    if (gid < total) {
      int c = (int)(gid % p.C);
      int64_t t = gid / p.C;
      int l = (int)(t % p.Do); t /= p.Do;
      int j = (int)(t % p.Wo); t /= p.Wo;
      int i = (int)(t % p.Ho);
      int64_t b = t / p.Ho;
      Win wh = window(i, p.sh, p.ph, p.kh, p.H);
      Win ww = window(j, p.sw, p.pw, p.kw, p.W);
      Win wd = window(l, p.sd, p.pd, p.kd, p.D);
      A best = neg_inf<A>();
      int64_t bi = 0;
      for (int hh = wh.lo; hh < wh.hi; ++hh)
        for (int w2 = ww.lo; w2 < ww.hi; ++w2)
          for (int dd = wd.lo; dd < wd.hi; ++dd) {
            int64_t id = (((b * p.H + hh) * (int64_t)p.W + w2) * p.D + dd) * (int64_t)p.C + c;
            A tv = (A)x[id];
            if (best < tv) {
              best = tv;
              bi = id;
            }
          }
      flag = bi == 0;
      g = (A)dy[gid];
    }
    hit[threadIdx.x] = flag;
    val[threadIdx.x] = g;
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < 256; ++k)
        if (hit[k]) acc += val[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    dx[0] = (T)acc;
    *orphan_flag = 0;
  }
}

__global__ void set_flag_kernel(int *p) { *p = 1; }

// ---------------------------------------------------------------- launchers
template <typename T>
static int pick_vec(const PoolParams &p, const void *a, const void *b, const void *c) {
  const int v = 16 / (int)sizeof(T);
  uintptr_t al = (uintptr_t)a | (uintptr_t)b | (uintptr_t)c;
  if (p.C % v == 0 && (al % 16) == 0) return v;
  return 1;
}

template <typename T>
void run_pool_fwd(const PoolParams &p, int mode, const T *x, T *y, int64_t *arg, hipStream_t st) {
  // argmax: VEC channels per thread as well, its int64 indices written as
  // 16-byte pairs (argmax 2x2 on a 64x56x56x128 FP32 layer: one channel per
  // thread ran at 4.3 TB/s)
  const int vec = pick_vec<T>(p, x, y, mode == 2 ? (const void *)arg : nullptr);
  int64_t total = (int64_t)p.B * p.Ho * p.Wo * p.Do * (p.C / vec);
  dim3 grid((unsigned)((total + 255) / 256)), block(256);
  constexpr int V = 16 / (int)sizeof(T);
  if (mode == 0) {
    if (vec == V) hipLaunchKernelGGL((pool_fwd_kernel<T, V, 0>), grid, block, 0, st, p, x, y, arg);
    else hipLaunchKernelGGL((pool_fwd_kernel<T, 1, 0>), grid, block, 0, st, p, x, y, arg);
  } else if (mode == 1) {
    if (vec == V) hipLaunchKernelGGL((pool_fwd_kernel<T, V, 1>), grid, block, 0, st, p, x, y, arg);
    else hipLaunchKernelGGL((pool_fwd_kernel<T, 1, 1>), grid, block, 0, st, p, x, y, arg);
  } else {
    if (vec == V) hipLaunchKernelGGL((pool_fwd_kernel<T, V, 2>), grid, block, 0, st, p, x, y, arg);
    else hipLaunchKernelGGL((pool_fwd_kernel<T, 1, 2>), grid, block, 0, st, p, x, y, arg);
  }
}

template <typename T>
void run_pool_bwd(const PoolParams &p, bool is_max, const T *x, const T *dy, T *dx, int *flag,
                  hipStream_t st) {
  int vec = pick_vec<T>(p, x, dy, dx);
  int64_t total = (int64_t)p.B * p.H * p.W * p.D * (p.C / vec);
  dim3 grid((unsigned)((total + 255) / 256)), block(256);
  constexpr int V = 16 / (int)sizeof(T);
  const bool tiled = p.kh == p.sh && p.kw == p.sw && p.kd == p.sd && p.ph == 0 && p.pw == 0 &&
                     p.pd == 0 && p.H == p.Ho * p.sh && p.W == p.Wo * p.sw && p.D == p.Do * p.sd;
  if (is_max && tiled) {
    int64_t wt = (int64_t)p.B * p.Ho * p.Wo * p.Do * (p.C / vec);
    dim3 g2((unsigned)((wt + 255) / 256));
    if (vec == V) hipLaunchKernelGGL((maxpool_bwd_tiled_kernel<T, V>), g2, block, 0, st, p, x, dy, dx, flag);
    else hipLaunchKernelGGL((maxpool_bwd_tiled_kernel<T, 1>), g2, block, 0, st, p, x, dy, dx, flag);
    hipLaunchKernelGGL((pool_bwd_fixup_kernel<T>), dim3(1), dim3(256), 0, st, p, x, dy, dx, flag);
  } else if (is_max) {
    if (vec == V) hipLaunchKernelGGL((pool_bwd_kernel<T, V, true>), grid, block, 0, st, p, x, dy, dx, flag);
    else hipLaunchKernelGGL((pool_bwd_kernel<T, 1, true>), grid, block, 0, st, p, x, dy, dx, flag);
    hipLaunchKernelGGL((pool_bwd_fixup_kernel<T>), dim3(1), dim3(256), 0, st, p, x, dy, dx, flag);
  } else if (tiled) {
    int64_t wt = (int64_t)p.B * p.Ho * p.Wo * p.Do * (p.C / vec);
    dim3 g2((unsigned)((wt + 255) / 256));
    if (vec == V) hipLaunchKernelGGL((avgpool_bwd_tiled_kernel<T, V>), g2, block, 0, st, p, dy, dx);
    else hipLaunchKernelGGL((avgpool_bwd_tiled_kernel<T, 1>), g2, block, 0, st, p, dy, dx);
  } else {
    if (vec == V) hipLaunchKernelGGL((pool_bwd_kernel<T, V, false>), grid, block, 0, st, p, x, dy, dx, flag);
    else hipLaunchKernelGGL((pool_bwd_kernel<T, 1, false>), grid, block, 0, st, p, x, dy, dx, flag);
  }
}

// A window with no in-bounds cell (only possible with caller-chosen pads)
// is an orphan too; the host detects it from the sizes and raises the flag.
void run_set_flag(int *flag, hipStream_t st) {
  hipLaunchKernelGGL(set_flag_kernel, dim3(1), dim3(1), 0, st, flag);
}

template void run_pool_fwd<float>(const PoolParams &, int, const float *, float *, int64_t *, hipStream_t);
template void run_pool_fwd<double>(const PoolParams &, int, const double *, double *, int64_t *, hipStream_t);
template void run_pool_bwd<float>(const PoolParams &, bool, const float *, const float *, float *, int *, hipStream_t);
template void run_pool_bwd<double>(const PoolParams &, bool, const double *, const double *, double *, int *, hipStream_t);
template void run_pool_fwd<bf16_t>(const PoolParams &, int, const bf16_t *, bf16_t *, int64_t *, hipStream_t);
template void run_pool_fwd<f16_t>(const PoolParams &, int, const f16_t *, f16_t *, int64_t *, hipStream_t);
template void run_pool_bwd<bf16_t>(const PoolParams &, bool, const bf16_t *, const bf16_t *, bf16_t *, int *, hipStream_t);
template void run_pool_bwd<f16_t>(const PoolParams &, bool, const f16_t *, const f16_t *, f16_t *, int *, hipStream_t);

}  // namespace owl_amd

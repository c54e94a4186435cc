// Kernel-side parameter block and launcher declarations for conv.hip.
#pragma once
#include <hip/hip_ext.h>

#include "owl_amd_common.h"

namespace owl_amd {

// One implicit GEMM (see conv.hip for the M / N / K mapping of each pass).
struct ConvParams {
  int B, H, W, D, Ci;
  int kh, kw, kd, Co;
  int Ho, Wo, Do;
  int sh, sw, sd;
  int rh, rw, rd;
  int ph, pw, pd;
  int64_t M;
  int N;
  int64_t K;
  const void *a;  // forward: x, backward input: dy, backward kernel: x
  const void *b;  // forward: w, backward input: w,  backward kernel: dy
  void *c;        // y, dx, dw (or split-K workspace)
  int64_t ldc;
  const float *scale_a;  // FP8 forward: per-tensor scales of x and w (else null)
  const float *scale_b;
  int split;            // backward kernel: number of K splits
  int64_t k_per_split;  // multiple of the K-step
  // forward epilogue (owl_amd_conv_fwd_bias_act): y = act(conv + bias[n]);
  // bias in the accumulator type (float, or double for FP64), pre receives
  // conv + bias before the activation (same layout and type as y)
  const void *bias;
  void *pre;
  int act;  // OWL_AMD_ACT_*
  // forward, 16-bit halo kernel only: 2x2 stride-2 max-pool of y written
  // from the epilogue (owl_amd_conv_fwd_bias_act_pool), else null
  void *pool;
};

template <int PASS, typename TI, typename TO>
void run_igemm(const ConvParams &p, bool aligned, hipStream_t st);

// One problem of a grouped (ragged) forward launch.
struct GroupEntry {
  ConvParams p;        // p.split > 1: p.c is the problem's FP32 partials
  int64_t tile_begin;  // first tile of this problem in the concatenated list
  int64_t tiles_mn;    // output tiles of one split
  int cfg;             // 0: full tile edge, 1: half (for small problems)
  int aligned;
};
template <typename TI, typename TO>
void run_group_fwd(const GroupEntry *dev_tab, int nprob, int64_t total, int grid, hipStream_t st);
template <typename TI> int group_tile_edge(int cfg);
template <typename TI> int group_k_step();  // K per step of the grouped tile body
template <typename TA, typename TO>
void run_splitk_reduce(const TA *ws, TO *out, int64_t n, int splits, hipStream_t st);
template <typename T>
void run_fp8_quantize(const T *src, int64_t n, unsigned char *dst, float *scale, hipStream_t st);
void run_e4m3_encode(const float *src, int64_t n, unsigned char *dst, hipStream_t st);
// conv_fast.hip: forward / unit-stride backward input with A fragments read
// straight from global memory.  shape: 0 -> 256x64, 1 -> 256x128,
// 2 -> 128x256, 3 -> 128x128 (BM x BN); pd = K-steps of A in flight (1-3).
template <int PASS, typename TI, typename TO>
void run_fast(const ConvParams &p, int shape, int pd, hipStream_t st);
int fast_bk(int elem_bytes);
// conv_fast.hip: stride-1 2-D forward / backward input with an LDS-staged
// input halo (16-bit storage); returns false when the problem does not fit.
template <int PASS, typename TI, typename TO>
bool run_halo(const ConvParams &p, hipStream_t st);
bool halo_fits(const ConvParams &p, int pass, int in_bytes, int out_bytes);
// Share of a halo launch's MFMA work on real outputs (0 if it does not fit).
double halo_fill(const ConvParams &p, int pass, int in_bytes, int out_bytes);
// FP8 e4m3 forward on the halo kernel (conv_fast.hip): x as [..][Ci] bytes,
// wp the kernel repacked to [tap][co][ci] (conv_fp8.hip), per-tensor scales
// on the device; false when the layer does not fit (2-D, stride 1,
// Ci % 64 == 0, Co % 8 == 0).
bool halo_fits_fp8(const ConvParams &p, int out_bytes);
template <typename TO>
bool run_halo_fp8(const ConvParams &p, const void *wp, hipStream_t st);
void run_fp8_pack_krows(const unsigned char *w, unsigned char *wp, int64_t T, int Ci, int Co, hipStream_t st);
// conv_fast.hip: backward kernel for 16-bit and FP32 storage (es = bytes per
// value) via LDS-staged [pixel][channel] rows (tiles of 128 or 192 rows
// (tap, ci) x 128 or 64 output channels x one K-step of pixels; p.split /
// p.k_per_split as for run_igemm<2>).
int wgrad_tile_m(int64_t M, int es);
// Groups of four waves per weight-gradient workgroup (2: each group sums
// half of a split's pixels; 1 for light 16-bit tiles; OWL_AMD_WGRAD_KG = 1 or
// 2 forces one).
int wgrad_kgroups(const ConvParams &p, int es);
template <typename TI, typename TO>
void run_wgrad(const ConvParams &p, hipStream_t st);
// True when run_wgrad takes the halo form (conv_wgrad_halo_kernel).
bool wgrad_halo_fits(const ConvParams &p);
template <typename TI> int tile_bm();
template <typename TI> int tile_bk();

// Zero fill of bytes at p on st (epilogue.hip): a kernel, not a memset node
// when captured.
void run_zero_fill(void *p, size_t bytes, hipStream_t st);

// Adjoint of the fused bias / activation epilogue (epilogue.hip); partial
// holds bias_act_chunks(npix) x C values of A when db is wanted.
int bias_act_chunks(int64_t npix);
template <typename T, typename A>
void run_bias_act_bwd(int act, int64_t npix, int C, const T *pre, const T *dy, T *dz, A *partial, A *db,
                      hipStream_t st);
int maxpool_relu_bwd_chunks(int64_t nwin);
template <typename T, typename A>
void run_maxpool_relu_bwd(int64_t B, int H, int W, int C, const T *y, const T *pre, const T *dp, T *dz,
                          A *partial, A *db, hipStream_t st);

// 2-D nearest-neighbour upsampling (resample.hip): backward = false maps
// src = x [B,H,W,C] -> dst = y [B,Ho,Wo,C]; true maps dy -> dx.
template <typename T>
void run_upsample(bool backward, const T *src, T *dst, int64_t B, int H, int W, int C, int Ho, int Wo, int sh,
                  int sw, hipStream_t st);
// resample.hip: dst[b, h, w, d] = src[b, h / sh, w / sw, d / sd] on the
// strided grid, zero elsewhere (rows of row_bytes bytes, a multiple of 2).
void run_stride_scatter(const void *src, void *dst, int64_t B, int Hs, int Ws, int Ds, int H, int W, int D, int sh,
                        int sw, int sd, int64_t row_bytes, hipStream_t st);
// resample.hip: the phase grids of a strided backward input (phase q =
// (ch, cw, cd) holds the dx pixels (sh m + ch, sw n + cw, sd l + cd) as a
// dense [B][mh][mw][md] grid at byte offset off[q] of src, or zeros when
// off[q] < 0) interleaved into dx; first[q] = the workgroups of the phases
// before q (B mh rows each), first[n] their total.
constexpr int MAX_PHASES = 64;
struct PhaseMap {
  int n;
  int ch[MAX_PHASES], cw[MAX_PHASES], cd[MAX_PHASES];
  int mh[MAX_PHASES], mw[MAX_PHASES], md[MAX_PHASES];
  int64_t off[MAX_PHASES];
  int64_t first[MAX_PHASES + 1];
};
void run_phase_interleave(const void *src, void *dst, int H, int W, int D, int sh, int sw, int sd, int64_t row_bytes,
                          const PhaseMap &pm, hipStream_t st);
// resample.hip: dst block j (block_bytes each) = src block src[j], j < n.
constexpr int MAX_TAPS = 64;
struct TapMap {
  int n;
  int src[MAX_TAPS];
};
void run_tap_gather(const void *src, void *dst, int64_t block_bytes, const TapMap &tm, hipStream_t st);
// resample.hip: dst[b, m, n, l] = src[b, sh m + ch, sw n + cw, sd l + cd], dst
// the dense [B][Mh][Mw][Md] grid of one stride phase (rows of row_bytes bytes).
void run_stride_gather(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                       int sw, int sd, int ch, int cw, int cd, int64_t row_bytes, hipStream_t st);
// resample.hip: space-to-depth of x ([B][Mh][Mw][Md][sh sw sd C], phase-major
// channel blocks of row_bytes, zeros past the image), and the matching kernel
// relayout: block j (block_bytes) of dst = tap tap[j] of src, zeros for -1.
void run_space_to_depth(const void *src, void *dst, int64_t B, int H, int W, int D, int Mh, int Mw, int Md, int sh,
                        int sw, int sd, int64_t row_bytes, hipStream_t st);
constexpr int MAX_S2D_BLOCKS = 256;
struct S2dTaps {
  int n;
  int tap[MAX_S2D_BLOCKS];
};
void run_s2d_kernel(const void *src, void *dst, int64_t block_bytes, const S2dTaps &tm, hipStream_t st);

// complex.hip: complex kernels as real ones (see there).  pattern 0 expands
// the forward's kernel, 1 the backward input's; fold turns the real weight
// gradient [T][2Ci][2Co] back into complex [T][Ci][Co].
template <typename R>
void run_cplx_expand(const R *w, R *wr, int64_t T, int Ci, int Co, int pattern, hipStream_t st);
template <typename R>
void run_cplx_fold(const R *dwr, R *dw, int64_t T, int Ci, int Co, hipStream_t st);

// Kernel timing (owl_amd_kernel_timing).  While timing is enabled, a KTimer
// scope holds a start / stop event pair, and the first launch_k() inside the
// scope (on this thread) hands them to hipExtLaunchKernel, which stamps them
// with that kernel's own start and end -- the interval a kernel trace
// reports, without the dispatch latency an event recorded around the launch
// would add.  Scopes without a launch_k() record nothing.
void ktime_begin(int kclass, const char *name, hipStream_t st, void **token);
void ktime_end(void *token, hipStream_t st);
bool ktime_take(hipEvent_t *start, hipEvent_t *stop);
struct KTimer {
  void *tok = nullptr;
  hipStream_t st;
  KTimer(int kclass, const char *name, hipStream_t s) : st(s) { ktime_begin(kclass, name, s, &tok); }
  ~KTimer() {
    if (tok) ktime_end(tok, st);
  }
  KTimer(const KTimer &) = delete;
  KTimer &operator=(const KTimer &) = delete;
};

// Launch with no dynamic LDS on stream st; the arguments are converted to
// the kernel's parameter types.
template <typename F, typename... Args>
inline void launch_k(F kernel, dim3 grid, dim3 block, hipStream_t st, Args... args) {
  hipEvent_t a = nullptr, b = nullptr;
  ktime_take(&a, &b);
  hipExtLaunchKernelGGL(kernel, grid, block, 0, st, a, b, 0, args...);
}

}  // namespace owl_amd

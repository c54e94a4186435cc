/*
 * owl_amd.h -- C ABI of the MI355X (gfx950) implementation of Owl's
 * convolution and pooling hot path.
 *
 * Two layers, both plain C (pointers, 64-bit sizes, status codes):
 *
 * 1. Owl drop-in entry points, one per C stub that Owl's OCaml externals bind
 *    (src/owl/core/owl_ndarray_conv.ml, owl_ndarray_pool.ml).  Each takes the
 *    reference stub's arguments in the reference's order, with host (Bigarray)
 *    data pointers, and is synchronous: on return the output host buffer is
 *    complete, as Owl's CPU code guarantees.  The only difference from the
 *    reference is the error path: a status code instead of exit(1)
 *    (owl_ndarray_conv_impl.h:299-300); the OCaml shim turns it into
 *    caml_failwith.  Naming: stub_<dtype>_ndarray_<op>_native  ->
 *    owl_amd_<dtype>_ndarray_<op>.
 *
 * 2. Device entry points (owl_amd_conv_*, owl_amd_pool_*) on device pointers
 *    and a caller stream, used by the drop-ins and by callers that keep
 *    tensors resident in HBM.  They enqueue work and return without waiting.
 *
 * Layouts are Owl's: input [B][H][W][D][Ci] (Owl calls H "cols", W "rows",
 * D "dpts"), kernel [kh][kw][kd][Ci][Co], output [B][Ho][Wo][Do][Co]; a 2-D
 * call is the 3-D one with D = kd = Do = 1.
 */
#ifndef OWL_AMD_H
#define OWL_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  OWL_AMD_OK = 0,
  OWL_AMD_ERR_ARG = 1,         /* invalid sizes / pointers */
  OWL_AMD_ERR_HIP = 2,         /* HIP runtime error */
  OWL_AMD_ERR_UNSUPPORTED = 3, /* dtype / variant not provided */
  OWL_AMD_ERR_NOMEM = 4        /* device allocation failed */
};

typedef enum {
  OWL_AMD_F32 = 0,
  OWL_AMD_F64 = 1,
  OWL_AMD_BF16 = 2,
  OWL_AMD_F16 = 3,
  OWL_AMD_FP8E4M3 = 4,
  OWL_AMD_C32 = 5, /* complex, FP32 parts (Owl's complex32) */
  OWL_AMD_C64 = 6  /* complex, FP64 parts (Owl's complex64) */
} owl_amd_dtype;

/* Complex elements: real part then imaginary part, the layout of C99
 * `_Complex float` / `_Complex double` that Owl's complex stubs use. */
typedef struct {
  float re, im;
} owl_amd_complex32;
typedef struct {
  double re, im;
} owl_amd_complex64;

typedef struct {
  int64_t B, H, W, D, Ci; /* input  [B][H][W][D][Ci] */
  int64_t kh, kw, kd, Co; /* kernel [kh][kw][kd][Ci][Co] */
  int64_t Ho, Wo, Do;     /* output [B][Ho][Wo][Do][Co] */
  int64_t sh, sw, sd;     /* strides */
  int64_t rh, rw, rd;     /* dilation rates (Owl: *_in_stride) */
  int64_t ph, pw, pd;     /* leading pads */
} owl_amd_conv_desc;

typedef struct {
  int64_t B, H, W, D, C;
  int64_t kh, kw, kd; /* window */
  int64_t Ho, Wo, Do;
  int64_t sh, sw, sd;
  int64_t ph, pw, pd;
} owl_amd_pool_desc;

/* ---------------------------------------------------------------- library */
int owl_amd_init(int device);
const char *owl_amd_last_error(void);
/* Number of device buffers and bytes cached by the drop-in layer. */
int owl_amd_cache_stats(int64_t *n_buffers, int64_t *bytes);
/* The drop-in layer's hipGraph cache (no counterpart in Owl): while on, a
 * staging context runs a call kind / shape / dtype directly the first time
 * it sees it, captures the device work of the second such call (for calls
 * whose arrays fit the pinned chunks: the input DMAs, kernels and output
 * DMAs; otherwise the kernels) and replays the graph for every later call
 * on the same device buffers; host-side copies stay outside the graph.
 * Off by default (replays measured no faster than enqueueing on MI355X);
 * OWL_AMD_DROPIN_GRAPHS=1 or owl_amd_dropin_graphs(1) turns it on, (0) off;
 * any other mode only queries.  Returns the current setting.  Bypassed
 * while kernel timing is on.  graph_stats: graphs held now; captures /
 * replays: totals since load. */
int owl_amd_dropin_graphs(int mode);
int owl_amd_dropin_graph_stats(int64_t *graphs, int64_t *captures, int64_t *replays);
void owl_amd_cache_release(void);

/* Kernel timing (diagnostics; no counterpart in the reference).  While
 * enabled, each convolution call brackets its main compute kernel with a
 * pair of HIP events on the stream it launches on; the auxiliary kernels
 * (split-K reduction) are timed as their own class.
 * owl_amd_kernel_timing(1) clears the totals and starts recording, (0)
 * stops.  owl_amd_kernel_time waits for the recorded events and returns the
 * total milliseconds and launch count of one class, and the name of the
 * kernel that class last launched. */
enum {
  OWL_AMD_KT_FWD = 0,        /* main kernel of owl_amd_conv_fwd */
  OWL_AMD_KT_BWD_INPUT = 1,  /* main kernel of owl_amd_conv_bwd_input */
  OWL_AMD_KT_BWD_KERNEL = 2, /* main kernel of owl_amd_conv_bwd_kernel */
  OWL_AMD_KT_AUX = 3,        /* split-K reduction and other auxiliary kernels */
  OWL_AMD_KT_COUNT = 4
};
int owl_amd_kernel_timing(int enable);
int owl_amd_kernel_time(int kclass, double *total_ms, int64_t *launches, const char **name);

/* ---------------------------------------------------------------- device API */
/* Forward convolution y = x (*) w.  Replaces the arithmetic of
 * owl_ndarray_conv_impl.h:246-453 (spatial), :1211-1319 (cuboid_im2col),
 * :3097-3190 / :3417-3533 (dilated).  in_t: F32, F64, BF16, F16 (inputs and
 * kernel); out_t: F32 or in_t. */
int owl_amd_conv_fwd(const owl_amd_conv_desc *d, int in_t, int out_t,
                     const void *x, const void *w, void *y, void *stream);

/* Forward convolution with the conv layer's bias and activation fused into
 * the epilogue: y = act(conv(x, w) + bias[co]).  Owl builds this layer as
 * conv2d + b followed by an activation (owl_neural_neuron.ml:1253 Conv2D.run,
 * owl_algodiff_ops.ml:762-778 relu); relu is fmax(v, 0) as
 * owl_ndarray_maths_stub.c:3559-3569 (NaN -> 0).  bias: Co values in the
 * accumulator type (float; double when out_t is F64), or null.  pre: null,
 * or a second output receiving conv + bias before the activation (the
 * backward pass needs it: Owl's relu adjoint looks at its input's sign). */
enum { OWL_AMD_ACT_NONE = 0, OWL_AMD_ACT_RELU = 1 };
int owl_amd_conv_fwd_bias_act(const owl_amd_conv_desc *d, int in_t, int out_t,
                              const void *x, const void *w, const void *bias, int act,
                              void *y, void *pre, void *stream);

/* The conv layer followed by a 2x2 stride-2 VALID max-pool layer, the pair
 * of a VGG block's end: Conv2D.run then the activation, then MaxPool2D.run =
 * max_pool2d VALID [2;2] [2;2] (owl_neural_neuron.ml:1253, :2024;
 * owl_dense_ndarray_generic.ml:6905; owl_ndarray_pool_impl.h:9-92).  Writes
 * what owl_amd_conv_fwd_bias_act writes, and pooled[B][Ho/2][Wo/2][Co] =
 * the max of each window of y, bit-identical to owl_amd_pool_fwd over y.
 * 2-D only (D = Do = kd = 1).  y may be null when only the pooled layer is
 * wanted (inference).  For BF16 / F16 stride-1 layers the halo kernel pools
 * its output tile before it leaves the chip (y is then not read back);
 * otherwise the convolution runs, then the pooling kernel (into a library
 * buffer when y is null; calls on one device that overlap then share it). */
int owl_amd_conv_fwd_bias_act_pool(const owl_amd_conv_desc *d, int in_t, int out_t,
                                   const void *x, const void *w, const void *bias, int act,
                                   void *y, void *pre, void *pooled, void *stream);

/* 2-D nearest-neighbour upsampling of NHWC tensors by integer scales
 * (sh along H = Owl's cols, sw along W = rows).  Forward:
 * y[b,i,j,c] = x[b, min(i/sh, H-1), min(j/sw, W-1), c] -- Owl's upsampling2d is
 * repeat x [|1;sh;sw;1|] (owl_dense_ndarray_generic.ml:8223-8234), i.e.
 * Ho = H*sh, Wo = W*sw.  Backward: replaces owl_ndarray_upsampling_impl.h:9-60
 * (spatial_backward): dx[b,hi,wi,c] = sum of dy over the output cells mapping
 * to (hi,wi), in Owl's order (bit-identical for F32 / F64).
 * dt: F32, F64, BF16, F16. */
int owl_amd_upsampling2d_fwd(int dt, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho,
                             int64_t Wo, int64_t sh, int64_t sw, const void *x, void *y, void *stream);
int owl_amd_upsampling2d_bwd(int dt, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho,
                             int64_t Wo, int64_t sh, int64_t sw, const void *dy, void *dx, void *stream);

/* Adjoint of y = act(z + bias) for z laid out [npix][C] (C fastest):
 * dz = dy * (signum(pre) + 1) / 2 for relu (owl_algodiff_ops.ml:774: 1 above
 * zero, 1/2 at zero, 0 below, NaN propagates), dz = dy for no activation;
 * db[c] = sum over pixels of dz[., c] when db is non-null (float; double for
 * F64), summed in a fixed order.  dt: F32, F64, BF16, F16 for pre, dy, dz.
 * dz may alias dy. */
int owl_amd_bias_act_bwd(int dt, int act, int64_t npix, int64_t C, const void *pre,
                         const void *dy, void *dz, void *db, void *stream);

/* Adjoint of pooled = max_pool2d VALID [2;2] [2;2] (relu(z + bias)), the
 * backward of owl_amd_conv_fwd_bias_act_pool: Owl's max-pool backward
 * (owl_ndarray_pool_impl.h:102-202: each window's gradient to its first
 * strict maximum in y) followed by the relu adjoint and bias gradient of
 * owl_amd_bias_act_bwd, in one pass without the intermediate dy.  y =
 * relu(pre) and pre are [B][H][W][C] with H and W even, dp is
 * [B][H/2][W/2][C].  dz is bit-identical to owl_amd_pool_bwd then
 * owl_amd_bias_act_bwd (F16: up to the sign of zeros, IEEE's -0 for a
 * negative gradient times a zero factor being kept here); db is summed in
 * another fixed order (each thread takes a window's four pixels), so it
 * agrees to rounding and is the same on every call.  act: RELU only.
 * dt: F32, F64, BF16, F16. */
int owl_amd_maxpool_bias_act_bwd(int dt, int act, int64_t B, int64_t H, int64_t W, int64_t C,
                                 const void *y, const void *pre, const void *dp, void *dz,
                                 void *db, void *stream);

/* FP8 e4m3 forward with per-tensor scales: x = x_q * x_scale, w = w_q * w_scale
 * (the scales are what an amax pass produced; owl_amd_fp8_quantize below).
 * out_t: F32 or BF16. */
int owl_amd_conv_fwd_fp8(const owl_amd_conv_desc *d, int out_t, const void *xq,
                         const void *wq, const float *x_scale, const float *w_scale,
                         void *y, void *stream);

/* Per-tensor amax -> scale = amax / 448 and e4m3 quantisation (saturating,
 * round to nearest even).  src_t: F32 or BF16.  scale is a device float. */
int owl_amd_fp8_quantize(int src_t, const void *src, int64_t n, void *dst_q,
                         float *scale, void *stream);
/* Plain float -> e4m3 encoding (no scaling), for conversion checks. */
int owl_amd_fp8_encode(const float *src, int64_t n, void *dst, void *stream);

/* dx = backward_input(w, dy): owl_ndarray_conv_impl.h:465-664, :1456-1554,
 * :3317-3406, :3678-3784.  Gather formulation, no atomics. */
int owl_amd_conv_bwd_input(const owl_amd_conv_desc *d, int in_t, int out_t,
                           const void *w, const void *dy, void *dx, void *stream);

/* dw = backward_kernel(x, dy): owl_ndarray_conv_impl.h:675-888, :1331-1444,
 * :3203-3306, :3545-3666.  Split-K over B*Ho*Wo*Do with FP32 partials and a
 * deterministic compensated (Neumaier) reduction.  workspace may be NULL, in
 * which case the library uses its own cached device buffer. */
size_t owl_amd_conv_bwd_kernel_workspace(const owl_amd_conv_desc *d, int in_t);
int owl_amd_conv_bwd_kernel(const owl_amd_conv_desc *d, int in_t, int out_t,
                            const void *x, const void *dy, void *dw,
                            void *workspace, size_t ws_bytes, void *stream);

/* Both backward passes of a layer, dx and dw from x, w and dy: the pair
 * Owl's reverse rule for a convolution computes (owl_algodiff_ops.ml:2364,
 * conv2d_backward_input a b s ca, conv2d_backward_kernel a b s ca): the
 * weight gradient, then the backward input, on the caller's stream, with
 * the results of owl_amd_conv_bwd_kernel and owl_amd_conv_bwd_input (dx_t,
 * dw_t: their out_t).  workspace as for owl_amd_conv_bwd_kernel. */
int owl_amd_conv_bwd(const owl_amd_conv_desc *d, int in_t, int dx_t, int dw_t, const void *x,
                     const void *w, const void *dy, void *dx, void *dw, void *workspace,
                     size_t ws_bytes, void *stream);

/* The backward-kernel pass for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same arguments, kernels and
 * results as owl_amd_conv_bwd_kernel without a workspace, but its split-K
 * partials go to a file-scope device array of
 * OWL_AMD_NOALLOC_SCRATCH_BYTES (32 MiB) that every GPU the library runs on
 * holds until the process ends; the call itself allocates nothing (no
 * hipMalloc / hipFree, no library buffer).  A problem whose partials
 * (owl_amd_conv_bwd_kernel_workspace) exceed the scratch is rejected with
 * OWL_AMD_ERR_ARG (the split plans keep every problem's partials at or
 * below 25 MiB, so this is a guard).  Real kinds only (F32, F64, BF16, F16).  Every call on one
 * GPU shares that GPU's scratch: calls that overlap on one GPU (different
 * streams or host threads) race. */
#define OWL_AMD_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
int owl_amd_conv_bwd_kernel_noalloc(const owl_amd_conv_desc *d, int in_t, int out_t,
                                    const void *x, const void *dy, void *dw, void *stream);

/* Ragged batches of forward convolutions in one persistent launch (no
 * counterpart in Owl, which loops over calls; this is the batched form of
 * owl_amd_conv_fwd).  Each problem has a plan: kind 0 / 1 = a full / half
 * tile edge in the persistent launch, with 1..16 K splits (FP32 partials,
 * summed in split order by the deterministic reduction after the launch),
 * or kind 2 = a launch of the single-problem kernels of its own after it.
 * create: validates the n descriptors and takes each plan from the
 * autotuner's cache for its shape class, else a size heuristic (no splits);
 * autotune: times every candidate plan per shape class and caches the
 * fastest; plan / set_plan: read or impose the per-problem kinds and split
 * counts (plan reports the effective split count, which can come out below
 * the requested one); run: the launches over all problems, then waits;
 * capture / replay: record a run in a hipGraph on a non-default stream and
 * replay it (autotune and set_plan drop the graph: capture again).  Unsplit
 * problems in the persistent launch equal the single-problem implicit-GEMM
 * kernel bit for bit; own launches equal owl_amd_conv_fwd.  in_t / out_t:
 * BF16 -> BF16 or F32, F16 -> F16, F32 -> F32. */
typedef struct owl_amd_conv_group owl_amd_conv_group;
int owl_amd_conv_group_create(owl_amd_conv_group **g, int n,
                              const owl_amd_conv_desc *descs, int in_t, int out_t);
int owl_amd_conv_group_autotune(owl_amd_conv_group *g, const void *const *xs,
                                const void *const *ws, void *const *ys, void *stream);
int owl_amd_conv_group_run(owl_amd_conv_group *g, const void *const *xs,
                           const void *const *ws, void *const *ys, void *stream);
int owl_amd_conv_group_capture(owl_amd_conv_group *g, const void *const *xs,
                               const void *const *ws, void *const *ys, void *stream);
int owl_amd_conv_group_replay(owl_amd_conv_group *g, void *stream);
int owl_amd_conv_group_info(const owl_amd_conv_group *g, int *tile_edges,
                            int64_t *total_tiles);
int owl_amd_conv_group_plan(const owl_amd_conv_group *g, int *kinds, int *splits);
int owl_amd_conv_group_set_plan(owl_amd_conv_group *g, const int *kinds, const int *splits);
void owl_amd_conv_group_destroy(owl_amd_conv_group *g);

/* Pooling (owl_ndarray_pool_impl.h): F32 / F64 results bit-identical to
 * Owl's C.  BF16 / F16 (no counterpart in Owl, whose pooling is F32 / F64)
 * compare and add in FP32 and round once. */
int owl_amd_pool_fwd(const owl_amd_pool_desc *d, int t, int is_max,
                     const void *x, void *y, void *stream);
int owl_amd_pool_argmax(const owl_amd_pool_desc *d, int t, const void *x,
                        void *y, int64_t *argmax, void *stream);
int owl_amd_pool_bwd(const owl_amd_pool_desc *d, int t, int is_max,
                     const void *x, const void *dy, void *dx, void *stream);

/* ---------------------------------------------------------------- Owl drop-ins */
/* Conv 2-D: replaces stub_<T>_ndarray_conv_spatial_native
 * (owl_ndarray_conv_impl.h:246-253; external owl_ndarray_conv.ml:13-32). */
#define OWL_AMD_CONV2D_FWD_ARGS(T)                                                  \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t in_channel, int64_t kernel_cols,                  \
      int64_t kernel_rows, int64_t output_cols, int64_t output_rows,                \
      int64_t out_channel, int64_t row_stride, int64_t col_stride, int64_t padding, \
      int64_t row_in_stride, int64_t col_in_stride
/* Conv 2-D backward: owl_ndarray_conv_impl.h:465-471 / :675-681 (no padding
 * argument; input is written by backward_input, kernel by backward_kernel). */
#define OWL_AMD_CONV2D_BWD_ARGS(T)                                                  \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t in_channel, int64_t kernel_cols,                  \
      int64_t kernel_rows, int64_t output_cols, int64_t output_rows,                \
      int64_t out_channel, int64_t row_stride, int64_t col_stride,                  \
      int64_t row_in_stride, int64_t col_in_stride
/* Conv 3-D: owl_ndarray_conv_impl.h:1211-1220 (forward), :1331-1338 (bwd). */
#define OWL_AMD_CONV3D_FWD_ARGS(T)                                                  \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t input_dpts, int64_t in_channel,                   \
      int64_t kernel_cols, int64_t kernel_rows, int64_t kernel_dpts,                \
      int64_t output_cols, int64_t output_rows, int64_t output_dpts,                \
      int64_t out_channel, int64_t dpt_stride, int64_t row_stride,                  \
      int64_t col_stride, int64_t padding
#define OWL_AMD_CONV3D_BWD_ARGS(T)                                                  \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t input_dpts, int64_t in_channel,                   \
      int64_t kernel_cols, int64_t kernel_rows, int64_t kernel_dpts,                \
      int64_t output_cols, int64_t output_rows, int64_t output_dpts,                \
      int64_t out_channel, int64_t dpt_stride, int64_t row_stride,                  \
      int64_t col_stride
/* Dilated 3-D: owl_ndarray_conv_impl.h:3417-3427 (forward), :3545-3553 (bwd). */
#define OWL_AMD_DCONV3D_FWD_ARGS(T)                                                 \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t input_dpts, int64_t in_channel,                   \
      int64_t kernel_cols, int64_t kernel_rows, int64_t kernel_dpts,                \
      int64_t output_cols, int64_t output_rows, int64_t output_dpts,                \
      int64_t out_channel, int64_t dpt_stride, int64_t row_stride,                  \
      int64_t col_stride, int64_t dpt_in_stride, int64_t row_in_stride,             \
      int64_t col_in_stride, int64_t padding
#define OWL_AMD_DCONV3D_BWD_ARGS(T)                                                 \
  T *input, T *kernel, T *output, int64_t batches, int64_t input_cols,              \
      int64_t input_rows, int64_t input_dpts, int64_t in_channel,                   \
      int64_t kernel_cols, int64_t kernel_rows, int64_t kernel_dpts,                \
      int64_t output_cols, int64_t output_rows, int64_t output_dpts,                \
      int64_t out_channel, int64_t dpt_stride, int64_t row_stride,                  \
      int64_t col_stride, int64_t dpt_in_stride, int64_t row_in_stride,             \
      int64_t col_in_stride
/* Pool 2-D forward: owl_ndarray_pool_impl.h:9-15. */
#define OWL_AMD_POOL2D_FWD_ARGS(T)                                                  \
  T *input, T *output, int64_t batches, int64_t input_cols, int64_t input_rows,     \
      int64_t in_channel, int64_t kernel_cols, int64_t kernel_rows,                 \
      int64_t output_cols, int64_t output_rows, int64_t row_stride,                 \
      int64_t col_stride, int64_t padding, int64_t row_in_stride,                   \
      int64_t col_in_stride
/* Pool 2-D backward: owl_ndarray_pool_impl.h:102-108 (pad_rows = W pad,
 * pad_cols = H pad, from calc_conv2d_padding). */
#define OWL_AMD_POOL2D_BWD_ARGS(T)                                                  \
  T *input, T *output_back, T *input_back, int64_t batches, int64_t input_cols,     \
      int64_t input_rows, int64_t in_channel, int64_t kernel_cols,                  \
      int64_t kernel_rows, int64_t output_cols, int64_t output_rows,                \
      int64_t row_stride, int64_t col_stride, int64_t pad_rows, int64_t pad_cols
/* Max pool 2-D with argmax: owl_ndarray_pool_impl.h:459-465. */
#define OWL_AMD_POOL2D_ARG_ARGS(T)                                                  \
  T *input, T *output, int64_t *argmax, int64_t batches, int64_t input_cols,        \
      int64_t input_rows, int64_t in_channel, int64_t kernel_cols,                  \
      int64_t kernel_rows, int64_t output_cols, int64_t output_rows,                \
      int64_t row_stride, int64_t col_stride, int64_t pad_rows, int64_t pad_cols
/* Pool 3-D forward: owl_ndarray_pool_impl.h:212-219 (strides dpt,row,col). */
#define OWL_AMD_POOL3D_FWD_ARGS(T)                                                  \
  T *input, T *output, int64_t batches, int64_t input_cols, int64_t input_rows,     \
      int64_t input_dpts, int64_t in_channel, int64_t kernel_cols,                  \
      int64_t kernel_rows, int64_t kernel_dpts, int64_t output_cols,                \
      int64_t output_rows, int64_t output_dpts, int64_t dpt_stride,                 \
      int64_t row_stride, int64_t col_stride, int64_t padding
/* Pool 3-D backward: owl_ndarray_pool_impl.h:324-331 (strides col,row,dpt). */
#define OWL_AMD_POOL3D_BWD_ARGS(T)                                                  \
  T *input, T *output_back, T *input_back, int64_t batches, int64_t input_cols,     \
      int64_t input_rows, int64_t input_dpts, int64_t in_channel,                   \
      int64_t kernel_cols, int64_t kernel_rows, int64_t kernel_dpts,                \
      int64_t output_cols, int64_t output_rows, int64_t output_dpts,                \
      int64_t col_stride, int64_t row_stride, int64_t dpt_stride, int64_t padding

/* Upsampling backward: owl_ndarray_upsampling_impl.h:9-14 (input = dx, written;
 * output = dy, read). */
#define OWL_AMD_UPSAMPLE_ARGS(T)                                                    \
  T *input, T *output, int64_t batches, int64_t input_cols, int64_t input_rows,     \
      int64_t in_channel, int64_t output_cols, int64_t output_rows,                 \
      int64_t col_scale, int64_t row_scale

#define OWL_AMD_DECLARE_DROPINS(N, T)                                                   \
  /* owl_ndarray_conv_impl.h:246 spatial, :903 spatial_im2col, :1570 mec, :2440 naive */ \
  int owl_amd_##N##_ndarray_conv_spatial(OWL_AMD_CONV2D_FWD_ARGS(T));                    \
  int owl_amd_##N##_ndarray_conv_spatial_im2col(OWL_AMD_CONV2D_FWD_ARGS(T));             \
  int owl_amd_##N##_ndarray_conv_spatial_mec(OWL_AMD_CONV2D_FWD_ARGS(T));                \
  int owl_amd_##N##_ndarray_conv_spatial_naive(OWL_AMD_CONV2D_FWD_ARGS(T));              \
  /* :465 / :1116 / :1828 / :2642 */                                                    \
  int owl_amd_##N##_ndarray_conv_spatial_backward_input(OWL_AMD_CONV2D_BWD_ARGS(T));     \
  int owl_amd_##N##_ndarray_conv_spatial_backward_input_im2col(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_spatial_backward_input_mec(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_spatial_backward_input_naive(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  /* :675 / :1006 / :1699 / :2546 */                                                    \
  int owl_amd_##N##_ndarray_conv_spatial_backward_kernel(OWL_AMD_CONV2D_BWD_ARGS(T));    \
  int owl_amd_##N##_ndarray_conv_spatial_backward_kernel_im2col(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_spatial_backward_kernel_mec(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_spatial_backward_kernel_naive(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  /* :1211 cuboid_im2col, :1957 mec, :2735 naive */                                     \
  int owl_amd_##N##_ndarray_conv_cuboid_im2col(OWL_AMD_CONV3D_FWD_ARGS(T));              \
  int owl_amd_##N##_ndarray_conv_cuboid_mec(OWL_AMD_CONV3D_FWD_ARGS(T));                 \
  int owl_amd_##N##_ndarray_conv_cuboid_naive(OWL_AMD_CONV3D_FWD_ARGS(T));               \
  /* :1456 / :2278 / :2977 */                                                           \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_input_im2col(OWL_AMD_CONV3D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_input_mec(OWL_AMD_CONV3D_BWD_ARGS(T));  \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_input_naive(OWL_AMD_CONV3D_BWD_ARGS(T)); \
  /* :1331 / :2118 / :2860 */                                                           \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_kernel_im2col(OWL_AMD_CONV3D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_kernel_mec(OWL_AMD_CONV3D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_cuboid_backward_kernel_naive(OWL_AMD_CONV3D_BWD_ARGS(T)); \
  /* :3097 / :3317 / :3203 dilated spatial */                                           \
  int owl_amd_##N##_ndarray_conv_dilated_spatial_im2col(OWL_AMD_CONV2D_FWD_ARGS(T));     \
  int owl_amd_##N##_ndarray_conv_dilated_spatial_backward_input_im2col(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_dilated_spatial_backward_kernel_im2col(OWL_AMD_CONV2D_BWD_ARGS(T)); \
  /* :3417 / :3678 / :3545 dilated cuboid */                                            \
  int owl_amd_##N##_ndarray_conv_dilated_cuboid_im2col(OWL_AMD_DCONV3D_FWD_ARGS(T));     \
  int owl_amd_##N##_ndarray_conv_dilated_cuboid_backward_input_im2col(OWL_AMD_DCONV3D_BWD_ARGS(T)); \
  int owl_amd_##N##_ndarray_conv_dilated_cuboid_backward_kernel_im2col(OWL_AMD_DCONV3D_BWD_ARGS(T)); \
  /* owl_ndarray_pool_impl.h:9 spatial, :102 spatial_backward, :212 cuboid,          */ \
  /* :324 cuboid_backward, :459 spatial_arg; macro sets owl_ndarray_pool_stub.c:16-51 */ \
  int owl_amd_##N##_ndarray_maxpool_spatial(OWL_AMD_POOL2D_FWD_ARGS(T));                 \
  int owl_amd_##N##_ndarray_avgpool_spatial(OWL_AMD_POOL2D_FWD_ARGS(T));                 \
  int owl_amd_##N##_ndarray_maxpool_spatial_backward(OWL_AMD_POOL2D_BWD_ARGS(T));        \
  int owl_amd_##N##_ndarray_avgpool_spatial_backward(OWL_AMD_POOL2D_BWD_ARGS(T));        \
  int owl_amd_##N##_ndarray_maxpool_spatial_arg(OWL_AMD_POOL2D_ARG_ARGS(T));             \
  int owl_amd_##N##_ndarray_maxpool_cuboid(OWL_AMD_POOL3D_FWD_ARGS(T));                  \
  int owl_amd_##N##_ndarray_avgpool_cuboid(OWL_AMD_POOL3D_FWD_ARGS(T));                  \
  int owl_amd_##N##_ndarray_maxpool_cuboid_backward(OWL_AMD_POOL3D_BWD_ARGS(T));         \
  int owl_amd_##N##_ndarray_avgpool_cuboid_backward(OWL_AMD_POOL3D_BWD_ARGS(T));        \
  /* owl_ndarray_upsampling_impl.h:9 spatial_backward (dx from dy); spatial is   */    \
  /* the forward with the same arguments (output written), which Owl computes    */    \
  /* with repeat (owl_dense_ndarray_generic.ml:8223) and has no stub for         */    \
  int owl_amd_##N##_ndarray_upsampling_spatial_backward(OWL_AMD_UPSAMPLE_ARGS(T));       \
  int owl_amd_##N##_ndarray_upsampling_spatial(OWL_AMD_UPSAMPLE_ARGS(T));

OWL_AMD_DECLARE_DROPINS(float32, float)
OWL_AMD_DECLARE_DROPINS(float64, double)

/* Complex dtypes (owl_ndarray_conv_stub.c:91-122, owl_ndarray_pool_stub.c:92-165,
 * owl_ndarray_upsampling_stub.c:31-45): the same drop-ins for complex32 and
 * complex64, and OWL_AMD_C32 / OWL_AMD_C64 on the device entry points.
 *  - Convolutions run as the real convolution of the interleaved channels
 *    (2 Ci -> 2 Co) with the kernel expanded on the device into the real
 *    matrix of the complex-linear map; Owl's complex GEMMs use plain (not
 *    conjugate) transposes, so the backward passes are the plain transposes
 *    too.  Results agree with Owl's complex arithmetic to FP rounding.
 *  - Average pooling and upsampling are the real kernels on 2 C channels
 *    (complex sums and division by the count are componentwise): bit-exact.
 *  - Max pooling (and its backward and argmax) returns OWL_AMD_ERR_UNSUPPORTED:
 *    Owl's complex max starts from -INFINITY, i.e. -inf + 0i whose modulus is
 *    +inf, so its modulus-then-argument comparison (owl_core.h:125,133) never
 *    replaces it; the shim keeps Owl's CPU stub, whose result callers get today.
 * owl_amd_complex_supported() returns 1. */
OWL_AMD_DECLARE_DROPINS(complex32, owl_amd_complex32)
OWL_AMD_DECLARE_DROPINS(complex64, owl_amd_complex64)
int owl_amd_complex_supported(void);

#ifdef __cplusplus
}
#endif
#endif /* OWL_AMD_H */

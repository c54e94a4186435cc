// C ABI of libowl_amd: argument checking, GEMM-view set-up, dtype dispatch,
// split-K planning, and the Owl drop-in entry points that own device buffers.
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "conv_kernels.h"
#include "owl_amd_common.h"
#include "pool_kernels.h"

using namespace owl_amd;

// ------------------------------------------------------------------ errors
static thread_local char g_err[512] = "";

extern "C" void owl_amd_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

extern "C" const char *owl_amd_last_error(void) { return g_err; }

#define ARG_FAIL(...)                \
  do {                               \
    owl_amd_set_error(__VA_ARGS__);  \
    return OWL_AMD_ERR_ARG;          \
  } while (0)

// ------------------------------------------------------------------ kernel timing
namespace owl_amd {
namespace {
struct KRec {
  int kclass;
  const char *name;
  hipEvent_t a, b;
  bool taken;  // handed to a launch
  KRec *prev;  // enclosing open scope on this thread
};
thread_local KRec *t_kt_open = nullptr;
std::mutex g_kt_mu;
std::atomic<bool> g_kt_on{false};
std::vector<KRec *> g_kt_pending;
double g_kt_ms[OWL_AMD_KT_COUNT];
int64_t g_kt_n[OWL_AMD_KT_COUNT];
const char *g_kt_name[OWL_AMD_KT_COUNT];

// Events are pooled (created when timing is enabled, reused after each
// read) so that recording adds no event creation to the launch path.
std::vector<hipEvent_t> g_kt_pool;

hipEvent_t kt_event() {
  hipEvent_t e = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_kt_mu);
    if (!g_kt_pool.empty()) {
      e = g_kt_pool.back();
      g_kt_pool.pop_back();
      return e;
    }
  }
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

// caller holds g_kt_mu
void kt_free(KRec *r) {
  if (r->a) g_kt_pool.push_back(r->a);
  if (r->b) g_kt_pool.push_back(r->b);
  delete r;
}
}  // namespace

void ktime_begin(int kclass, const char *name, hipStream_t st, void **token) {
  (void)st;
  *token = nullptr;
  if (!g_kt_on.load(std::memory_order_relaxed)) return;
  KRec *r = new KRec{kclass, name, kt_event(), kt_event(), false, t_kt_open};
  if (!r->a || !r->b) {
    std::lock_guard<std::mutex> lk(g_kt_mu);
    kt_free(r);
    return;
  }
  t_kt_open = r;
  *token = r;
}

bool ktime_take(hipEvent_t *start, hipEvent_t *stop) {
  KRec *r = t_kt_open;
  if (!r || r->taken) return false;
  r->taken = true;
  *start = r->a;
  *stop = r->b;
  return true;
}

void ktime_end(void *token, hipStream_t st) {
  (void)st;
  KRec *r = (KRec *)token;
  t_kt_open = r->prev;
  std::lock_guard<std::mutex> lk(g_kt_mu);
  if (!r->taken) {
    kt_free(r);
    return;
  }
  g_kt_pending.push_back(r);
}
}  // namespace owl_amd

extern "C" int owl_amd_kernel_timing(int enable) {
  std::lock_guard<std::mutex> lk(g_kt_mu);
  g_kt_on.store(enable != 0);
  if (enable) {
    for (KRec *r : g_kt_pending) kt_free(r);
    g_kt_pending.clear();
    while (g_kt_pool.size() < 512) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        break;
      }
      g_kt_pool.push_back(e);
    }
    for (int i = 0; i < OWL_AMD_KT_COUNT; ++i) {
      g_kt_ms[i] = 0.0;
      g_kt_n[i] = 0;
      g_kt_name[i] = nullptr;
    }
  }
  return OWL_AMD_OK;
}

extern "C" int owl_amd_kernel_time(int kclass, double *total_ms, int64_t *launches, const char **name) {
  if (kclass < 0 || kclass >= OWL_AMD_KT_COUNT) ARG_FAIL("kernel class %d out of range", kclass);
  std::lock_guard<std::mutex> lk(g_kt_mu);
  int rc = OWL_AMD_OK;
  for (KRec *r : g_kt_pending) {
    float ms = 0.0f;
    if (rc == OWL_AMD_OK) {
      hipError_t e = hipEventSynchronize(r->b);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, r->a, r->b);
      if (e != hipSuccess) {
        owl_amd_set_error("kernel timing: %s", hipGetErrorString(e));
        rc = OWL_AMD_ERR_HIP;
      } else {
        g_kt_ms[r->kclass] += ms;
        g_kt_n[r->kclass] += 1;
        g_kt_name[r->kclass] = r->name;
      }
    }
    kt_free(r);
  }
  g_kt_pending.clear();
  if (total_ms) *total_ms = g_kt_ms[kclass];
  if (launches) *launches = g_kt_n[kclass];
  if (name) *name = g_kt_name[kclass] ? g_kt_name[kclass] : "";
  return rc;
}

// ------------------------------------------------------------------ per-device state
namespace {

struct DevBuf {
  void *ptr = nullptr;
  size_t bytes = 0;
};

struct DeviceState {
  bool ready = false;
  hipStream_t stream = nullptr;  // default priority, used by the drop-ins
  int *flag = nullptr;           // pooling orphan flag (one int)
  DevBuf ws;                     // split-K workspace
  DevBuf red;                    // bias-gradient partial sums
  DevBuf cplx;                   // complex convolutions: expanded kernel / real weight gradient
  DevBuf f8w;                    // FP8 halo forward: kernel repacked to [tap][co][ci]
  DevBuf ytmp;                   // fused conv + pool without y, when the pooling runs apart
  DevBuf dsub;                   // strided backward input: dx's phase grids
  DevBuf wsub;                   // strided backward input: the kernel's taps grouped by phase
  DevBuf xsub;                   // strided forward with dilation a multiple of the stride: x's one phase
};

// Recursive: the drop-ins hold it while calling the device API, which takes it
// again for the split-K workspace and the pooling flag.
std::recursive_mutex g_mu;
DeviceState g_dev[64];
// Library scratch buffers replaced by larger ones.  A HIP graph captured
// earlier may still hold their addresses, so they are only freed by
// owl_amd_cache_release (which invalidates such graphs); guarded by g_mu.
std::vector<void *> g_retired;

int current_device(int *dev) {
  hipError_t e = hipGetDevice(dev);
  if (e != hipSuccess) {
    owl_amd_set_error("hipGetDevice: %s", hipGetErrorString(e));
    return OWL_AMD_ERR_HIP;
  }
  if (*dev < 0 || *dev >= 64) ARG_FAIL("device index %d out of range", *dev);
  return OWL_AMD_OK;
}

int ensure_state(DeviceState **out) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  DeviceState &s = g_dev[dev];
  if (!s.ready) {
    OWL_CHECK_HIP(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking));
    OWL_CHECK_HIP(hipMalloc(&s.flag, 256));
    OWL_CHECK_HIP(hipMemset(s.flag, 0, 256));
    OWL_CHECK_HIP(hipDeviceSynchronize());
    s.ready = true;
  }
  *out = &s;
  return OWL_AMD_OK;
}

// Grow-only device buffer; the caller holds g_mu.  A scratch buffer that
// device-API calls hand to kernels (retire = true) is kept alive when it is
// replaced -- a captured graph may still use it -- and grows at least 2x so
// that a run of slowly growing calls retires few buffers.  The drop-ins'
// staging slots (used only inside one synchronous call) are freed at once,
// after the stream that used them has finished.
int ensure_buf(DevBuf &b, size_t bytes, hipStream_t st, bool retire = true) {
  if (b.bytes >= bytes) return OWL_AMD_OK;
  size_t want = bytes < 256 ? 256 : bytes;
  if (b.ptr) {
    if (retire) {
      g_retired.push_back(b.ptr);
      if (want < 2 * b.bytes) want = 2 * b.bytes;
    } else {
      OWL_CHECK_HIP(hipStreamSynchronize(st));
      OWL_CHECK_HIP(hipFree(b.ptr));
    }
    b.ptr = nullptr;
    b.bytes = 0;
  }
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    owl_amd_set_error("hipMalloc(%zu): %s", want, hipGetErrorString(e));
    return OWL_AMD_ERR_NOMEM;
  }
  b.bytes = want;
  return OWL_AMD_OK;
}

size_t dtype_size(int t) {
  switch (t) {
    case OWL_AMD_F32: return 4;
    case OWL_AMD_F64: return 8;
    case OWL_AMD_BF16:
    case OWL_AMD_F16: return 2;
    case OWL_AMD_FP8E4M3: return 1;
    case OWL_AMD_C32: return 8;
    case OWL_AMD_C64: return 16;
  }
  return 0;
}

bool fits_int(int64_t v) { return v >= 0 && v <= 0x7fffffffLL; }

int check_conv(const owl_amd_conv_desc *d) {
  if (!d) ARG_FAIL("null descriptor");
  const int64_t v[] = {d->B, d->H, d->W, d->D, d->Ci, d->kh, d->kw, d->kd, d->Co,
                       d->Ho, d->Wo, d->Do, d->ph, d->pw, d->pd};
  for (int64_t x : v)
    if (!fits_int(x)) ARG_FAIL("conv size out of range (negative or > INT32_MAX)");
  if (d->sh < 1 || d->sw < 1 || d->sd < 1) ARG_FAIL("conv stride must be >= 1");
  if (d->rh < 1 || d->rw < 1 || d->rd < 1) ARG_FAIL("conv dilation must be >= 1");
  if (d->kh < 1 || d->kw < 1 || d->kd < 1) ARG_FAIL("conv kernel extent must be >= 1");
  if (!fits_int(d->B * d->H * d->W * d->D) || !fits_int(d->B * d->Ho * d->Wo * d->Do) ||
      !fits_int(d->kh * d->kw * d->kd * d->Ci * d->Co))
    ARG_FAIL("conv tensor too large for 32-bit pixel indexing");
  return OWL_AMD_OK;
}

ConvParams make_params(const owl_amd_conv_desc *d) {
  ConvParams p;
  memset(&p, 0, sizeof(p));  // scale_a / scale_b null: not FP8
  p.B = (int)d->B; p.H = (int)d->H; p.W = (int)d->W; p.D = (int)d->D; p.Ci = (int)d->Ci;
  p.kh = (int)d->kh; p.kw = (int)d->kw; p.kd = (int)d->kd; p.Co = (int)d->Co;
  p.Ho = (int)d->Ho; p.Wo = (int)d->Wo; p.Do = (int)d->Do;
  p.sh = (int)d->sh; p.sw = (int)d->sw; p.sd = (int)d->sd;
  p.rh = (int)d->rh; p.rw = (int)d->rw; p.rd = (int)d->rd;
  p.ph = (int)d->ph; p.pw = (int)d->pw; p.pd = (int)d->pd;
  p.split = 1;
  return p;
}

bool aligned16(const void *a, const void *b, const void *c) {
  return (((uintptr_t)a | (uintptr_t)b | (uintptr_t)c) & 15) == 0;
}

template <int PASS>
int dispatch_igemm(const ConvParams &p, int in_t, int out_t, bool aligned, hipStream_t st) {
  KTimer kt(PASS, "conv_igemm_kernel", st);
  if (in_t == OWL_AMD_F32 && out_t == OWL_AMD_F32) run_igemm<PASS, float, float>(p, aligned, st);
  else if (in_t == OWL_AMD_F64 && out_t == OWL_AMD_F64) run_igemm<PASS, double, double>(p, aligned, st);
  else if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_BF16) run_igemm<PASS, bf16_t, bf16_t>(p, aligned, st);
  else if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_F32) run_igemm<PASS, bf16_t, float>(p, aligned, st);
  else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F16) run_igemm<PASS, f16_t, f16_t>(p, aligned, st);
  else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F32) run_igemm<PASS, f16_t, float>(p, aligned, st);
  else {
    owl_amd_set_error("unsupported dtype pair in=%d out=%d", in_t, out_t);
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    owl_amd_set_error("kernel launch: %s", hipGetErrorString(e));
    return OWL_AMD_ERR_HIP;
  }
  return OWL_AMD_OK;
}

int env_int(const char *name, int dflt) {
  const char *e = getenv(name);
  return e ? atoi(e) : dflt;
}

int vec_of(int t) { return 16 / (int)dtype_size(t); }

// Kernel family selection: OWL_AMD_CONV_PATH=generic forces the LDS-staged
// kernel of conv.hip everywhere (used to cross-check the two families).
bool fast_path_enabled() {
  static int v = -1;
  if (v < 0) {
    const char *e = getenv("OWL_AMD_CONV_PATH");
    v = (e && strcmp(e, "generic") == 0) ? 0 : 1;
  }
  return v == 1;
}

// OWL_AMD_HALO=0 disables the halo kernel (cross-checks, tuning).
bool halo_enabled() {
  static int v = -1;
  if (v < 0) v = env_int("OWL_AMD_HALO", 1);
  return v != 0;
}

int fast_shape_for(int N) { return N <= 64 ? 0 : (N <= 128 ? 1 : 2); }



// A-prefetch depth of the fast kernels (OWL_AMD_FAST_PD overrides).
int fast_pd() {
  static int v = -1;
  if (v < 0) v = env_int("OWL_AMD_FAST_PD", 2);
  return v;
}

// Forward / backward-input on the conv_fast.hip kernels when the problem
// fits them; returns false (nothing launched) otherwise.
template <int PASS>
bool try_fast(const ConvParams &p, int in_t, int out_t, bool aligned, hipStream_t st) {
  if (!fast_path_enabled() || !aligned) return false;
  if (in_t != OWL_AMD_F32 && in_t != OWL_AMD_BF16 && in_t != OWL_AMD_F16) return false;
  const int C = PASS == 0 ? p.Ci : p.Co;
  if (PASS == 1 && (p.sh != 1 || p.sw != 1 || p.sd != 1)) return false;
  // 32-bit byte offsets: operands and the GEMM rows must stay below 2 GiB
  const int64_t es = (int64_t)dtype_size(in_t);
  const int64_t a_elems = PASS == 0 ? (int64_t)p.B * p.H * p.W * p.D * p.Ci
                                    : (int64_t)p.B * p.Ho * p.Wo * p.Do * p.Co;
  const int64_t c_elems = p.M * (int64_t)p.ldc;
  if (a_elems * es >= (1LL << 31) || (int64_t)p.kh * p.kw * p.kd * p.Ci * p.Co * es >= (1LL << 31) ||
      c_elems * (int64_t)dtype_size(out_t) >= (1LL << 31))
    return false;
  // stride-1 problems with 16-bit storage: the halo kernel (2-D, or 3-D as
  // planes).  A 2-D backward input whose halo tiles would be less than 40%
  // full (small images: a tile holds whole rows of one image) runs on the
  // direct-A kernel instead where that applies, which tiles pixels across
  // images: 256x7x7x512 -> 512 3x3 BF16 254 -> 100 us, 128x7x7x512 118 -> 94 us
  // (10x10 and 14x14 images, and the forward, stay faster on the halo kernel).
  const bool d2 = p.D == 1 && p.Do == 1 && p.kd == 1;
  const bool sparse_halo = PASS == 1 && d2 && C % fast_bk((int)dtype_size(in_t)) == 0 &&
                           halo_fill(p, PASS, (int)dtype_size(in_t), (int)dtype_size(out_t)) < 0.4;
  if (halo_enabled() && !sparse_halo && halo_fits(p, PASS, (int)dtype_size(in_t), (int)dtype_size(out_t))) {
    bool done = false;
    if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_BF16) done = run_halo<PASS, bf16_t, bf16_t>(p, st);
    else if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_F32) done = run_halo<PASS, bf16_t, float>(p, st);
    else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F16) done = run_halo<PASS, f16_t, f16_t>(p, st);
    else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F32) done = run_halo<PASS, f16_t, float>(p, st);
    if (done) return true;
  }
  // the direct-A kernel takes whole K-steps of channels
  if (C % fast_bk((int)dtype_size(in_t)) != 0) return false;
  const int shape = fast_shape_for(p.N);
  KTimer kt(PASS, "conv_fast_kernel", st);
  if (in_t == OWL_AMD_F32 && out_t == OWL_AMD_F32) run_fast<PASS, float, float>(p, shape, fast_pd(), st);
  else if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_BF16) run_fast<PASS, bf16_t, bf16_t>(p, shape, fast_pd(), st);
  else if (in_t == OWL_AMD_BF16 && out_t == OWL_AMD_F32) run_fast<PASS, bf16_t, float>(p, shape, fast_pd(), st);
  else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F16) run_fast<PASS, f16_t, f16_t>(p, shape, fast_pd(), st);
  else if (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F32) run_fast<PASS, f16_t, float>(p, shape, fast_pd(), st);
  else return false;
  return true;
}

int tile_k(int t) {
  switch (t) {
    case OWL_AMD_F32: return tile_bk<float>();
    case OWL_AMD_F64: return tile_bk<double>();
    case OWL_AMD_BF16: return tile_bk<bf16_t>();
    default: return tile_bk<f16_t>();
  }
}
int tile_m(int t) {
  switch (t) {
    case OWL_AMD_F32: return tile_bm<float>();
    case OWL_AMD_F64: return tile_bm<double>();
    case OWL_AMD_BF16: return tile_bm<bf16_t>();
    default: return tile_bm<f16_t>();
  }
}

// Split-K plan for the backward-kernel GEMM (M = T*Ci, N = Co, K = pixels):
// one round of workgroups over the 256 CUs -- three per CU for the 16-bit
// kernel's 128-row tile, two for its 192-row tile (208-214 VGPRs) and for
// the implicit-GEMM kernels -- each split at least 16 K-steps long, so the
// FP32 partials stay a small fraction of the inputs' bytes.  The plan
// depends only on the shape and the input type, so the workspace query
// agrees with the call.
void plan_splitk(const ConvParams &p, int in_t, int *split, int64_t *kps) {
  // the weight-gradient kernel's tiles (16-bit and FP32 storage), else the
  // implicit GEMM's
  const bool w16 = in_t == OWL_AMD_BF16 || in_t == OWL_AMD_F16 || in_t == OWL_AMD_F32;
  const int es = in_t == OWL_AMD_F32 ? 4 : 2;
  const int bm = w16 ? wgrad_tile_m(p.M, es) : tile_m(in_t), bn = w16 ? 128 : tile_m(in_t), bk = tile_k(in_t);
  int64_t tiles = ((p.M + bm - 1) / bm) * (int64_t)((p.N + bn - 1) / bn);
  static const int env_target = env_int("OWL_AMD_SPLIT_TARGET", 0);  // workgroups to aim for (tuning)
  // three 128-row wgrad workgroups fit a CU, two of the 192-row ones; with
  // two wave groups per workgroup (each summing half a split) one 512-thread
  // workgroup per CU, 256 of them, whole ones only
  const int kg = w16 ? wgrad_kgroups(p, es) : 1;
  const int target = env_target > 0 ? env_target : (kg == 2 ? 256 : (w16 && bm == 128 ? 768 : 512));
  int64_t want = kg == 2 ? target / tiles : (target + tiles - 1) / tiles;
  int64_t max_split = (p.K + 16 * kg * bk - 1) / (16 * kg * bk);
  static const int cap = env_int("OWL_AMD_MAX_SPLIT", 1 << 30);  // diagnostics / tuning
  if (max_split > cap) max_split = cap;
  if (want > max_split) want = max_split;
  if (want < 1) want = 1;
  int64_t k = (p.K + want - 1) / want;
  k = (k + kg * bk - 1) / (kg * bk) * (kg * bk);  // halves of whole K-steps
  *kps = k;
  *split = (int)((p.K + k - 1) / k);
}

}  // namespace

// ------------------------------------------------------------------ init
extern "C" int owl_amd_init(int device) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  OWL_CHECK_HIP(hipSetDevice(device));
  DeviceState *s;
  return ensure_state(&s);
}

extern "C" int owl_amd_complex_supported(void) { return 1; }

// ------------------------------------------------------------------ complex kinds
// A complex convolution is the real one over the interleaved channels
// (2 Ci -> 2 Co) with the kernel expanded into the real matrix of the
// complex-linear map (complex.hip).  The expanded kernel, or the real weight
// gradient before it is folded, lives in a per-device grow-only buffer, so
// complex calls on one device that overlap on different streams race.
namespace {
bool is_complex(int t) { return t == OWL_AMD_C32 || t == OWL_AMD_C64; }
int real_of(int t) { return t == OWL_AMD_C32 ? OWL_AMD_F32 : OWL_AMD_F64; }
owl_amd_conv_desc widen(const owl_amd_conv_desc *d) {
  owl_amd_conv_desc r = *d;
  r.Ci *= 2;
  r.Co *= 2;
  return r;
}
int f8w_scratch(size_t bytes, void **out, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  DeviceState *s;
  int rc = ensure_state(&s);
  if (rc) return rc;
  rc = ensure_buf(s->f8w, bytes, st);
  if (rc) return rc;
  *out = s->f8w.ptr;
  return OWL_AMD_OK;
}
int cplx_scratch(size_t bytes, void **out, hipStream_t st) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  DeviceState *s;
  int rc = ensure_state(&s);
  if (rc) return rc;
  rc = ensure_buf(s->cplx, bytes, st);
  if (rc) return rc;
  *out = s->cplx.ptr;
  return OWL_AMD_OK;
}

// pass 0: a = x, b = w, c = y; pass 1: a = w, b = dy, c = dx; pass 2: a = x, b = dy, c = dw
int complex_conv(int pass, const owl_amd_conv_desc *d, int in_t, int out_t, const void *a, const void *b, void *c,
                 void *ws, size_t ws_bytes, void *stream) {
  int rc = check_conv(d);
  if (rc) return rc;
  if (in_t != out_t) ARG_FAIL("complex convolutions read and write the same kind (%d, %d)", in_t, out_t);
  const int rt = real_of(in_t);
  const owl_amd_conv_desc r = widen(d);
  const int64_t T = d->kh * d->kw * d->kd;
  const int64_t nk = T * d->Ci * d->Co;  // complex kernel elements
  hipStream_t st = (hipStream_t)stream;
  if (nk == 0) {  // nothing to expand: the real call handles the empty cases
    if (pass == 0) return owl_amd_conv_fwd(&r, rt, rt, a, b, c, stream);
    if (pass == 1) return owl_amd_conv_bwd_input(&r, rt, rt, a, b, c, stream);
    return owl_amd_conv_bwd_kernel(&r, rt, rt, a, b, c, ws, ws_bytes, stream);
  }
  if (!a || !b || !c) ARG_FAIL("null pointer");
  if (d->Ci > INT32_MAX / 2 || d->Co > INT32_MAX / 2) ARG_FAIL("channel count above 2^30");
  void *scr;
  rc = cplx_scratch((size_t)nk * 4 * dtype_size(rt), &scr, st);
  if (rc) return rc;
  const int ci = (int)d->Ci, co = (int)d->Co;
  if (pass == 0 || pass == 1) {
    const void *w = pass == 0 ? b : a;
    if (rt == OWL_AMD_F32) run_cplx_expand<float>((const float *)w, (float *)scr, T, ci, co, pass, st);
    else run_cplx_expand<double>((const double *)w, (double *)scr, T, ci, co, pass, st);
    OWL_CHECK_HIP(hipGetLastError());
    return pass == 0 ? owl_amd_conv_fwd(&r, rt, rt, a, scr, c, stream)
                     : owl_amd_conv_bwd_input(&r, rt, rt, scr, b, c, stream);
  }
  rc = owl_amd_conv_bwd_kernel(&r, rt, rt, a, b, scr, ws, ws_bytes, stream);
  if (rc) return rc;
  if (rt == OWL_AMD_F32) run_cplx_fold<float>((const float *)scr, (float *)c, T, ci, co, st);
  else run_cplx_fold<double>((const double *)scr, (double *)c, T, ci, co, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}
}  // namespace

// ------------------------------------------------------------------ device API: conv
namespace {
// A forward with few output tiles over a long K (small late-stage layers,
// single images) runs as split-K on the implicit GEMM: each split sums one
// range of whole K-steps of (tap, channel) into FP32 (FP64) partials in the
// library's split-K buffer, and the deterministic split-K reduction sums them
// in split order into y.  Plain forwards only (no fused epilogue);
// OWL_AMD_FWD_SPLITK=0 turns it off (read on every call).  Returns -1 when the
// problem does not take it.  5x7x7x512 -> 256 3x3 stride 2 rate 2 took 124 us
// and 16x14x14x512 -> 32 5x5 rate 2 150 us as one launch of 1-10 workgroups.
// The split count for a forward of M x N outputs over K, or 0 when it stays
// unsplit (more than 8 output tiles of the implicit GEMM -- larger problems
// fill the GPU on the halo and direct-A kernels --, fewer than 64 K-steps, or
// off).
int64_t fwd_split_count(int64_t M, int64_t N, int64_t K, int in_t) {
  if (env_int("OWL_AMD_FWD_SPLITK", 1) == 0) return 0;
  const int bm = tile_m(in_t), bk = tile_k(in_t);
  const int64_t tiles = ((M + bm - 1) / bm) * ((N + bm - 1) / bm);
  const int64_t ksteps = (K + bk - 1) / bk;
  if (tiles > 8 || ksteps < 64) return 0;
  int64_t split = 256 / tiles;
  if (split > ksteps / 8) split = ksteps / 8;
  return split < 2 ? 0 : split;
}

// The plain forward of a descriptor takes the split-K path (its partials in
// the per-device split-K buffer).
bool fwd_splitk_applies(const owl_amd_conv_desc *d, int t) {
  return dtype_size(t) && fwd_split_count(d->B * d->Ho * d->Wo * d->Do, d->Co, d->kh * d->kw * d->kd * d->Ci, t) > 0;
}

int fwd_splitk(ConvParams p, int in_t, int out_t, bool aligned, hipStream_t st) {
  if (!dtype_size(in_t) || !dtype_size(out_t)) return -1;
  const int bk = tile_k(in_t);
  const int64_t ksteps = (p.K + bk - 1) / bk;
  int64_t split = fwd_split_count(p.M, p.N, p.K, in_t);
  if (split == 0) return -1;
  const int64_t kps = (ksteps + split - 1) / split * bk;
  split = (p.K + kps - 1) / kps;
  const int64_t es = in_t == OWL_AMD_F64 ? 8 : 4;
  const size_t bytes = (size_t)(split * p.M * p.N * es);
  void *ws = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->ws, bytes, st);
    if (rc) return rc;
    ws = s->ws.ptr;
  }
  void *y = p.c;
  p.split = (int)split;
  p.k_per_split = kps;
  p.c = ws;
  int rc = dispatch_igemm<0>(p, in_t, out_t, aligned, st);
  if (rc) return rc;
  const int64_t n = p.M * (int64_t)p.N;
  KTimer kt(OWL_AMD_KT_AUX, "splitk_reduce_kernel", st);
  if (in_t == OWL_AMD_F64) run_splitk_reduce<double, double>((const double *)ws, (double *)y, n, p.split, st);
  else if (out_t == OWL_AMD_F32) run_splitk_reduce<float, float>((const float *)ws, (float *)y, n, p.split, st);
  else if (out_t == OWL_AMD_BF16) run_splitk_reduce<float, bf16_t>((const float *)ws, (bf16_t *)y, n, p.split, st);
  else run_splitk_reduce<float, f16_t>((const float *)ws, (f16_t *)y, n, p.split, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// A strided forward whose dilation is a multiple of the stride along every
// strided dimension (OWL_AMD_STRIDED_FWD_PHASE=0 turns the path off; read on
// every call): x[s i + r a - p] always has residue c = (-p) mod s, so the
// convolution reads only x's phase c, x_c[m] = x[s m + c], and equals the
// stride-1 convolution of x_c (M = ceil((H - c) / s) pixels) with rate r / s
// and pad (p + c) / s.  That phase is gathered into a per-device library
// buffer and the stride-1 forward (the halo kernel where it fits) runs on it;
// every y value is the same sum over (tap, ci), in that kernel's order.  A
// 24x112x112x512 -> 128 5x5 layer with stride 2 and rate 2 (BF16): 510.7 us on
// the direct-A kernel.
bool strided_fwd_one_phase(const owl_amd_conv_desc *d) {
  const int64_t s3[3] = {d->sh, d->sw, d->sd}, r3[3] = {d->rh, d->rw, d->rd}, p3[3] = {d->ph, d->pw, d->pd},
                n3[3] = {d->H, d->W, d->D};
  if (d->sh == 1 && d->sw == 1 && d->sd == 1) return false;
  for (int a = 0; a < 3; ++a) {
    if (r3[a] % s3[a] != 0) return false;
    if ((s3[a] - p3[a] % s3[a]) % s3[a] >= n3[a]) return false;  // the phase has no pixels
  }
  return env_int("OWL_AMD_STRIDED_FWD_PHASE", 1) != 0;
}

int conv_fwd_one_phase(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x, const void *w,
                       const void *bias, int act, void *y, void *pre, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  const int64_t s3[3] = {d->sh, d->sw, d->sd}, p3[3] = {d->ph, d->pw, d->pd}, n3[3] = {d->H, d->W, d->D};
  int64_t c[3], M[3];
  for (int a = 0; a < 3; ++a) {
    c[a] = (s3[a] - p3[a] % s3[a]) % s3[a];
    M[a] = (n3[a] - c[a] + s3[a] - 1) / s3[a];
  }
  const int64_t row_bytes = d->Ci * (int64_t)dtype_size(in_t);
  void *xs = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->xsub, (size_t)(d->B * M[0] * M[1] * M[2] * row_bytes), st);
    if (rc) return rc;
    xs = s->xsub.ptr;
  }
  {
    KTimer kt(OWL_AMD_KT_AUX, "stride_gather_kernel", st);
    run_stride_gather(x, xs, d->B, (int)d->H, (int)d->W, (int)d->D, (int)M[0], (int)M[1], (int)M[2], (int)d->sh,
                      (int)d->sw, (int)d->sd, (int)c[0], (int)c[1], (int)c[2], row_bytes, st);
  }
  OWL_CHECK_HIP(hipGetLastError());
  owl_amd_conv_desc q = *d;
  q.H = M[0];
  q.W = M[1];
  q.D = M[2];
  q.rh = d->rh / d->sh;
  q.rw = d->rw / d->sw;
  q.rd = d->rd / d->sd;
  q.ph = (d->ph + c[0]) / d->sh;
  q.pw = (d->pw + c[1]) / d->sw;
  q.pd = (d->pd + c[2]) / d->sd;
  q.sh = q.sw = q.sd = 1;
  return owl_amd_conv_fwd_bias_act(&q, in_t, out_t, xs, w, bias, act, y, pre, stream);
}

// Space-to-depth forward of a strided 2-D convolution with unit dilation
// (16-bit storage; OWL_AMD_S2D=0 turns it off, 2 takes it wherever it applies
// regardless of size, for tests; read on every call).  With
// a - p = s u + c, y[i] = sum over phases c and u of x_c[i + u] w[s u + c + p],
// x_c[m] = x[s m + c]: a stride-1 convolution of X = x's phases stacked as
// channels ([B][ceil(H / sh)][ceil(W / sw)][sh sw Ci]) with U = u_max - u_min
// + 1 taps per dimension (u_min = -ceil(p / s), u_max = floor((k - 1 - p) /
// s)), pad -u_min, and the kernel's taps moved to (u - u_min, phase) blocks,
// zero where s u + c + p falls outside the kernel.  The zero taps cost
// U_h U_w sh sw / (kh kw) times the MACs, so the path is taken when that is at
// most 1.5 (5x5 and larger kernels at stride 2) and the stride-1 problem fits
// the halo kernel: the ragged batch's 11x112x112x256 -> 256 5x5 stride-2 layer
// ran 392 us on the direct-A kernel.
struct S2dPlan {
  int64_t umin[2], U[2], M[2];
};
bool s2d_plan(const owl_amd_conv_desc *d, int in_t, int out_t, S2dPlan *pl, owl_amd_conv_desc *q) {
  if (in_t != OWL_AMD_BF16 && in_t != OWL_AMD_F16) return false;
  if (d->sh == 1 && d->sw == 1) return false;
  if (d->D != 1 || d->kd != 1 || d->sd != 1 || d->Do != 1 || d->pd != 0) return false;
  if (d->rh != 1 || d->rw != 1) return false;
  const int mode = env_int("OWL_AMD_S2D", 1);  // 0 off, 1 where it pays, 2 wherever it applies (tests)
  if (mode == 0) return false;
  const int64_t s2[2] = {d->sh, d->sw}, k2[2] = {d->kh, d->kw}, p2[2] = {d->ph, d->pw}, n2[2] = {d->H, d->W};
  for (int a = 0; a < 2; ++a) {
    if (p2[a] > k2[a] - 1) return false;
    pl->umin[a] = -((p2[a] + s2[a] - 1) / s2[a]);
    pl->U[a] = (k2[a] - 1 - p2[a]) / s2[a] - pl->umin[a] + 1;
    pl->M[a] = (n2[a] + s2[a] - 1) / s2[a];
  }
  const int64_t nph = d->sh * d->sw;
  if (pl->U[0] * pl->U[1] * nph * 2 > 3 * d->kh * d->kw) return false;  // more than 1.5x the MACs
  // enough output tiles for the stride-1 kernel (a 24x7x7x256 5x5 layer ran
  // 122 against 67 us), and enough work per byte of x that rewriting x costs
  // little beside it (12x112x112x256 -> 128 5x5: 144 us plus ~30 us of
  // rewriting against 157 us; 11x112x112x256 -> 256: 220 + ~30 against 389)
  const double out_px = (double)d->B * d->Ho * d->Wo;
  const double flop = 2.0 * out_px * d->Co * d->kh * d->kw * d->Ci, xbytes = (double)d->B * d->H * d->W * d->Ci * 2.0;
  if (mode == 1 && (out_px < 16384.0 || flop < 1024.0 * xbytes)) return false;
  if (pl->U[0] * pl->U[1] * nph > MAX_S2D_BLOCKS) return false;
  *q = *d;
  q->H = pl->M[0];
  q->W = pl->M[1];
  q->Ci = nph * d->Ci;
  q->kh = pl->U[0];
  q->kw = pl->U[1];
  q->sh = q->sw = 1;
  q->ph = -pl->umin[0];
  q->pw = -pl->umin[1];
  if (check_conv(q)) return false;
  ConvParams cp = make_params(q);
  return halo_fits(cp, 0, 2, (int)dtype_size(out_t));
}

// The forward takes the space-to-depth path (its phases and kernel in
// per-device library buffers).
bool s2d_uses_buffers(const owl_amd_conv_desc *d, int t) {
  S2dPlan pl;
  owl_amd_conv_desc q;
  return s2d_plan(d, t, t, &pl, &q);
}

int conv_fwd_s2d(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x, const void *w, const void *bias,
                 int act, void *y, void *pre, void *stream, const S2dPlan &pl, const owl_amd_conv_desc &q) {
  hipStream_t st = (hipStream_t)stream;
  const int64_t es = 2, nph = d->sh * d->sw;
  const int64_t xbytes = d->B * pl.M[0] * pl.M[1] * nph * d->Ci * es;
  const int64_t block = d->Ci * d->Co * es;
  S2dTaps tm;
  memset(&tm, 0, sizeof(tm));
  for (int64_t th = 0; th < pl.U[0]; ++th)
    for (int64_t tw = 0; tw < pl.U[1]; ++tw)
      for (int64_t ch = 0; ch < d->sh; ++ch)
        for (int64_t cw = 0; cw < d->sw; ++cw) {
          const int64_t ah = d->sh * (th + pl.umin[0]) + ch + d->ph, aw = d->sw * (tw + pl.umin[1]) + cw + d->pw;
          tm.tap[tm.n++] = (ah >= 0 && ah < d->kh && aw >= 0 && aw < d->kw) ? (int)(ah * d->kw + aw) : -1;
        }
  void *xs = nullptr, *ws = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->xsub, (size_t)xbytes, st);
    if (rc) return rc;
    rc = ensure_buf(s->wsub, (size_t)(tm.n * block), st);
    if (rc) return rc;
    xs = s->xsub.ptr;
    ws = s->wsub.ptr;
  }
  {
    KTimer kt(OWL_AMD_KT_AUX, "space_to_depth_kernel", st);
    run_space_to_depth(x, xs, d->B, (int)d->H, (int)d->W, 1, (int)pl.M[0], (int)pl.M[1], 1, (int)d->sh, (int)d->sw,
                       1, d->Ci * es, st);
    run_s2d_kernel(w, ws, block, tm, st);
  }
  OWL_CHECK_HIP(hipGetLastError());
  return owl_amd_conv_fwd_bias_act(&q, in_t, out_t, xs, ws, bias, act, y, pre, stream);
}
}  // namespace

extern "C" int owl_amd_conv_fwd(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x,
                                const void *w, void *y, void *stream) {
  if (is_complex(in_t) || is_complex(out_t)) return complex_conv(0, d, in_t, out_t, x, w, y, nullptr, 0, stream);
  return owl_amd_conv_fwd_bias_act(d, in_t, out_t, x, w, nullptr, OWL_AMD_ACT_NONE, y, nullptr, stream);
}

extern "C" int owl_amd_conv_fwd_bias_act(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x,
                                         const void *w, const void *bias, int act, void *y, void *pre,
                                         void *stream) {
  int rc = check_conv(d);
  if (rc) return rc;
  if (is_complex(in_t) || is_complex(out_t)) {
    owl_amd_set_error("the fused bias / activation epilogue takes real kinds");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  if (act != OWL_AMD_ACT_NONE && act != OWL_AMD_ACT_RELU) ARG_FAIL("unknown activation %d", act);
  if (pre && pre == y) ARG_FAIL("pre must not alias y");
  ConvParams p = make_params(d);
  p.M = d->B * d->Ho * d->Wo * d->Do;
  p.N = p.Co;
  p.K = d->kh * d->kw * d->kd * d->Ci;
  if (p.M == 0 || p.N == 0) return OWL_AMD_OK;
  if (!x || !w || !y) ARG_FAIL("null pointer");
  if (dtype_size(in_t) && dtype_size(out_t) && strided_fwd_one_phase(d))
    return conv_fwd_one_phase(d, in_t, out_t, x, w, bias, act, y, pre, stream);
  {
    S2dPlan pl;
    owl_amd_conv_desc q;
    if (dtype_size(out_t) && s2d_plan(d, in_t, out_t, &pl, &q))
      return conv_fwd_s2d(d, in_t, out_t, x, w, bias, act, y, pre, stream, pl, q);
  }
  p.a = x; p.b = w; p.c = y; p.ldc = p.Co;
  p.bias = bias; p.pre = pre; p.act = act;
  const int v = vec_of(in_t);
  bool aligned = dtype_size(in_t) && p.Ci % v == 0 && p.Co % v == 0 && aligned16(x, w, y) &&
                 aligned16(pre, nullptr, nullptr);
  if (!bias && act == OWL_AMD_ACT_NONE && !pre) {
    const int rc2 = fwd_splitk(p, in_t, out_t, aligned, (hipStream_t)stream);
    if (rc2 != -1) return rc2;
  }
  if (try_fast<0>(p, in_t, out_t, aligned, (hipStream_t)stream)) {
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  return dispatch_igemm<0>(p, in_t, out_t, aligned, (hipStream_t)stream);
}

extern "C" int owl_amd_conv_fwd_bias_act_pool(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x,
                                              const void *w, const void *bias, int act, void *y, void *pre,
                                              void *pooled, void *stream) {
  int rc = check_conv(d);
  if (rc) return rc;
  if (is_complex(in_t) || is_complex(out_t)) {
    owl_amd_set_error("the fused bias / activation / pooling epilogue takes real kinds");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  if (d->D != 1 || d->Do != 1 || d->kd != 1) {
    owl_amd_set_error("the fused pooling takes 2-D convolutions");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  if (act != OWL_AMD_ACT_NONE && act != OWL_AMD_ACT_RELU) ARG_FAIL("unknown activation %d", act);
  if (pre && pre == y) ARG_FAIL("pre must not alias y");
  const int64_t Hp = d->Ho / 2, Wp = d->Wo / 2;
  if (d->B * d->Ho * d->Wo * d->Co == 0) return OWL_AMD_OK;
  if (!x || !w || (!pooled && Hp * Wp > 0)) ARG_FAIL("null pointer");
  if (pooled && (pooled == y || pooled == pre)) ARG_FAIL("pooled must not alias y or pre");
  hipStream_t st = (hipStream_t)stream;
  // 16-bit stride-1 layers: the halo kernel pools its staged output tile
  if (in_t == out_t && (in_t == OWL_AMD_BF16 || in_t == OWL_AMD_F16) && Hp * Wp > 0 && fast_path_enabled() &&
      halo_enabled() && d->Ci % 8 == 0 && d->Co % 8 == 0 && aligned16(x, w, y) && aligned16(pre, pooled, nullptr)) {
    ConvParams p = make_params(d);
    p.M = d->B * d->Ho * d->Wo;
    p.N = p.Co;
    p.K = d->kh * d->kw * d->Ci;
    p.a = x; p.b = w; p.c = y; p.ldc = p.Co;
    p.bias = bias; p.pre = pre; p.act = act; p.pool = pooled;
    const bool done = in_t == OWL_AMD_BF16 ? run_halo<0, bf16_t, bf16_t>(p, st) : run_halo<0, f16_t, f16_t>(p, st);
    if (done) {
      OWL_CHECK_HIP(hipGetLastError());
      return OWL_AMD_OK;
    }
  }
  // otherwise the convolution, then the pooling kernel over y (a library
  // buffer when the caller does not want y)
  void *yy = y;
  if (!yy) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->ytmp, (size_t)(d->B * d->Ho * d->Wo * d->Co) * dtype_size(out_t), st);
    if (rc) return rc;
    yy = s->ytmp.ptr;
  }
  rc = owl_amd_conv_fwd_bias_act(d, in_t, out_t, x, w, bias, act, yy, pre, stream);
  if (rc || Hp * Wp == 0) return rc;
  const owl_amd_pool_desc q = {d->B, d->Ho, d->Wo, 1, d->Co, 2, 2, 1, Hp, Wp, 1, 2, 2, 1, 0, 0, 0};
  return owl_amd_pool_fwd(&q, out_t, 1, yy, pooled, stream);
}

namespace {
// OWL_AMD_STRIDED_1X1=0 sends strided 1x1 backward inputs to the implicit
// GEMM over every dx pixel (cross-checks, tuning); read on every call.
bool strided_1x1_enabled() { return env_int("OWL_AMD_STRIDED_1X1", 1) != 0; }

// The backward input takes the strided 1x1 path below, whose subsampled
// grid lives in a per-device library buffer.
bool strided_1x1_bwd_input(const owl_amd_conv_desc *d) {
  return d->kh == 1 && d->kw == 1 && d->kd == 1 && (d->sh > 1 || d->sw > 1 || d->sd > 1) && d->ph == 0 &&
         d->pw == 0 && d->pd == 0 && strided_1x1_enabled();
}

// OWL_AMD_STRIDED_PHASES=0 sends the other strided backward inputs to the
// implicit GEMM over every dx pixel (cross-checks, tuning); read on every call.
// The phase path takes unit dilation, at most MAX_PHASES stride phases and
// MAX_TAPS kernel taps; its phase grids and grouped kernel live in
// per-device library buffers.
bool strided_phase_bwd_input(const owl_amd_conv_desc *d) {
  return (d->sh > 1 || d->sw > 1 || d->sd > 1) && d->rh == 1 && d->rw == 1 && d->rd == 1 &&
         d->sh * d->sw * d->sd <= MAX_PHASES && d->kh * d->kw * d->kd <= MAX_TAPS && !strided_1x1_bwd_input(d) &&
         env_int("OWL_AMD_STRIDED_PHASES", 1) != 0;
}

// Backward input of a 1x1 convolution with strides and no padding: dx is
// dy . w^T on the strided grid (h = i sh, w = j sw, d = l sd) and zero off
// it.  The stride-1 backward input of the subsampled problem writes the
// grid's values into a library buffer (the halo or direct-A kernels, not
// the implicit GEMM over all of dx, which also computed every zero), then
// one pass writes dx: the buffer on the grid, zeros elsewhere.  Each grid
// value is the dot product over Co that the generic path forms (summed in
// the stride-1 kernel's order).  ResNet's 128x28x28x256 -> 512 stride-2
// downsampling layer (BF16): 80 us on the implicit GEMM.
int bwd_input_strided_1x1(const owl_amd_conv_desc *d, int in_t, int out_t, const void *w, const void *dy, void *dx,
                          void *stream) {
  hipStream_t st = (hipStream_t)stream;
  owl_amd_conv_desc s1 = *d;
  s1.H = d->Ho;
  s1.W = d->Wo;
  s1.D = d->Do;
  s1.sh = s1.sw = s1.sd = 1;
  const int64_t row_bytes = d->Ci * (int64_t)dtype_size(out_t);
  const size_t bytes = (size_t)(d->B * d->Ho * d->Wo * d->Do) * (size_t)row_bytes;
  void *sub = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->dsub, bytes, st);
    if (rc) return rc;
    sub = s->dsub.ptr;
  }
  int rc = owl_amd_conv_bwd_input(&s1, in_t, out_t, w, dy, sub, stream);
  if (rc) return rc;
  {
    KTimer kt(OWL_AMD_KT_AUX, "stride_scatter_kernel", st);
    run_stride_scatter(sub, dx, d->B, (int)d->Ho, (int)d->Wo, (int)d->Do, (int)d->H, (int)d->W, (int)d->D,
                       (int)d->sh, (int)d->sw, (int)d->sd, row_bytes, st);
  }
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// Backward input of a strided convolution (unit dilation) by output
// phases.  Along one dimension with stride s, kernel k and leading pad p, dx
// pixel h = s m + c (phase c) receives exactly the taps a = phi + s t with
// phi = (c + p) mod s, from dy pixel m - t + delta, delta = (c + p - phi) / s:
// a stride-1 backward input of the grid m < M = ceil((H - c) / s) with the
// T = ceil((k - phi) / s) taps of that phase and pad delta.  Each phase that
// has taps runs as such a problem (on the halo or direct-A kernels, instead
// of the implicit GEMM over every dx pixel, a 1 / (sh sw sd) share of whose
// work has taps) with the kernel's taps of the phase gathered into a library
// buffer, its grid into another, and one pass interleaves the grids into dx
// (phases without taps: zeros).  Every dx value is the same sum over (tap,
// co) that the generic path forms, in the stride-1 kernel's order.  ResNet's
// 64x56x56x128 -> 128 3x3 stride-2 layer (BF16): 154.5 us on the implicit GEMM.
int bwd_input_strided_phases(const owl_amd_conv_desc *d, int in_t, int out_t, const void *w, const void *dy,
                             void *dx, void *stream) {
  hipStream_t st = (hipStream_t)stream;
  const int64_t sd3[3] = {d->sh, d->sw, d->sd}, k3[3] = {d->kh, d->kw, d->kd}, p3[3] = {d->ph, d->pw, d->pd},
                n3[3] = {d->H, d->W, d->D};
  int64_t phi[3][MAX_PHASES], T[3][MAX_PHASES], del[3][MAX_PHASES], M[3][MAX_PHASES];
  for (int a = 0; a < 3; ++a)
    for (int64_t c = 0; c < sd3[a]; ++c) {
      phi[a][c] = (c + p3[a]) % sd3[a];
      T[a][c] = phi[a][c] < k3[a] ? (k3[a] - phi[a][c] + sd3[a] - 1) / sd3[a] : 0;
      del[a][c] = (c + p3[a] - phi[a][c]) / sd3[a];
      M[a][c] = c < n3[a] ? (n3[a] - c + sd3[a] - 1) / sd3[a] : 0;
    }
  const int64_t es_out = (int64_t)dtype_size(out_t);
  const int64_t row_bytes = d->Ci * es_out;
  const int64_t wblock = d->Ci * d->Co * (int64_t)dtype_size(in_t);
  PhaseMap pm;
  memset(&pm, 0, sizeof(pm));
  TapMap tm;
  memset(&tm, 0, sizeof(tm));
  struct Sub {
    owl_amd_conv_desc desc;
    int64_t woff, ooff;
  } subs[MAX_PHASES];
  int nsub = 0;
  int64_t obytes = 0;
  for (int64_t ch = 0; ch < d->sh; ++ch)
    for (int64_t cw = 0; cw < d->sw; ++cw)
      for (int64_t cd = 0; cd < d->sd; ++cd) {
        const int q = pm.n++;
        pm.ch[q] = (int)ch;
        pm.cw[q] = (int)cw;
        pm.cd[q] = (int)cd;
        pm.mh[q] = (int)M[0][ch];
        pm.mw[q] = (int)M[1][cw];
        pm.md[q] = (int)M[2][cd];
        pm.first[q + 1] = pm.first[q] + d->B * M[0][ch];
        const int64_t cells = d->B * M[0][ch] * M[1][cw] * M[2][cd];
        if (cells == 0 || T[0][ch] == 0 || T[1][cw] == 0 || T[2][cd] == 0) {
          pm.off[q] = -1;  // no taps (or no pixels): zeros
          continue;
        }
        pm.off[q] = obytes;
        Sub &u = subs[nsub++];
        u.ooff = obytes;
        u.woff = tm.n * wblock;
        obytes += cells * row_bytes;
        for (int64_t th = 0; th < T[0][ch]; ++th)
          for (int64_t tw = 0; tw < T[1][cw]; ++tw)
            for (int64_t td = 0; td < T[2][cd]; ++td)
              tm.src[tm.n++] = (int)(((phi[0][ch] + d->sh * th) * d->kw + phi[1][cw] + d->sw * tw) * d->kd +
                                     phi[2][cd] + d->sd * td);
        u.desc = *d;
        u.desc.H = M[0][ch];
        u.desc.W = M[1][cw];
        u.desc.D = M[2][cd];
        u.desc.kh = T[0][ch];
        u.desc.kw = T[1][cw];
        u.desc.kd = T[2][cd];
        u.desc.sh = u.desc.sw = u.desc.sd = 1;
        u.desc.ph = del[0][ch];
        u.desc.pw = del[1][cw];
        u.desc.pd = del[2][cd];
      }
  char *grids = nullptr, *wg = nullptr;
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    if (obytes) {
      rc = ensure_buf(s->dsub, (size_t)obytes, st);
      if (rc) return rc;
      rc = ensure_buf(s->wsub, (size_t)(tm.n * wblock), st);
      if (rc) return rc;
      grids = (char *)s->dsub.ptr;
      wg = (char *)s->wsub.ptr;
    }
  }
  if (tm.n) {
    KTimer kt(OWL_AMD_KT_AUX, "tap_gather_kernel", st);
    run_tap_gather(w, wg, wblock, tm, st);
  }
  for (int i = 0; i < nsub; ++i) {
    const int rc = owl_amd_conv_bwd_input(&subs[i].desc, in_t, out_t, wg + subs[i].woff, dy, grids + subs[i].ooff,
                                          stream);
    if (rc) return rc;
  }
  {
    KTimer kt(OWL_AMD_KT_AUX, "phase_interleave_kernel", st);
    run_phase_interleave(grids, dx, (int)d->H, (int)d->W, (int)d->D, (int)d->sh, (int)d->sw, (int)d->sd, row_bytes,
                         pm, st);
  }
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}
}  // namespace

extern "C" int owl_amd_conv_bwd_input(const owl_amd_conv_desc *d, int in_t, int out_t,
                                      const void *w, const void *dy, void *dx, void *stream) {
  if (is_complex(in_t) || is_complex(out_t)) return complex_conv(1, d, in_t, out_t, w, dy, dx, nullptr, 0, stream);
  int rc = check_conv(d);
  if (rc) return rc;
  ConvParams p = make_params(d);
  p.M = d->B * d->H * d->W * d->D;
  p.N = p.Ci;
  p.K = d->kh * d->kw * d->kd * d->Co;
  if (p.M == 0 || p.N == 0) return OWL_AMD_OK;
  if (!w || !dy || !dx) ARG_FAIL("null pointer");
  if (strided_1x1_bwd_input(d) && d->Co > 0 && dtype_size(in_t) && dtype_size(out_t))
    return bwd_input_strided_1x1(d, in_t, out_t, w, dy, dx, stream);
  if (strided_phase_bwd_input(d) && d->Co > 0 && dtype_size(in_t) && dtype_size(out_t))
    return bwd_input_strided_phases(d, in_t, out_t, w, dy, dx, stream);
  p.a = dy; p.b = w; p.c = dx; p.ldc = p.Ci;
  const int v = vec_of(in_t);
  bool aligned = dtype_size(in_t) && p.Co % v == 0 && p.Ci % 4 == 0 && aligned16(dy, w, dx);
  if (try_fast<1>(p, in_t, out_t, aligned, (hipStream_t)stream)) {
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  aligned = dtype_size(in_t) && p.Co % v == 0 && aligned16(dy, w, nullptr);
  return dispatch_igemm<1>(p, in_t, out_t, aligned, (hipStream_t)stream);
}

static void bwd_kernel_params(const owl_amd_conv_desc *d, ConvParams &p) {
  p = make_params(d);
  p.M = d->kh * d->kw * d->kd * d->Ci;
  p.N = p.Co;
  p.K = d->B * d->Ho * d->Wo * d->Do;
}

extern "C" size_t owl_amd_conv_bwd_kernel_workspace(const owl_amd_conv_desc *d, int in_t) {
  if (check_conv(d)) return 0;
  if (is_complex(in_t)) {
    const owl_amd_conv_desc r = widen(d);
    return owl_amd_conv_bwd_kernel_workspace(&r, real_of(in_t));
  }
  ConvParams p;
  bwd_kernel_params(d, p);
  if (p.M == 0 || p.N == 0 || p.K == 0) return 0;
  int split;
  int64_t kps;
  plan_splitk(p, in_t, &split, &kps);
  if (split <= 1) return 0;
  size_t acc = in_t == OWL_AMD_F64 ? 8 : 4;
  return (size_t)split * (size_t)p.M * (size_t)p.N * acc;
}

static int bwd_kernel_impl(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x, const void *dy, void *dw,
                           void *workspace, size_t ws_bytes, hipStream_t st) {
  int rc = check_conv(d);
  if (rc) return rc;
  ConvParams p;
  bwd_kernel_params(d, p);
  if (p.M == 0 || p.N == 0) return OWL_AMD_OK;
  if (!x || !dy || !dw) ARG_FAIL("null pointer");
  if (p.K == 0) {
    run_zero_fill(dw, (size_t)p.M * p.N * dtype_size(out_t), st);
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  plan_splitk(p, in_t, &p.split, &p.k_per_split);
  p.a = x; p.b = dy; p.ldc = p.Co;
  const int v = vec_of(in_t);
  bool aligned = dtype_size(in_t) && p.Ci % v == 0 && p.Co % v == 0 && aligned16(x, dy, nullptr);
  // 16-bit and FP32 storage: the weight-gradient kernel of conv_fast.hip
  // (32-bit byte offsets, so operands below 2 GiB; FP32 in, FP32 out)
  const int64_t es = (int64_t)dtype_size(in_t);
  const bool tr = fast_path_enabled() && aligned &&
                  (in_t == OWL_AMD_BF16 || in_t == OWL_AMD_F16 || (in_t == OWL_AMD_F32 && out_t == OWL_AMD_F32)) &&
                  (int64_t)p.B * p.H * p.W * p.D * p.Ci * es < (1LL << 31) && p.K * (int64_t)p.Co * es < (1LL << 31) &&
                  p.M * (int64_t)p.N * 4 < (1LL << 31) && p.K < (1LL << 31) &&
                  (out_t == in_t || out_t == OWL_AMD_F32);
  const char *wg_name = in_t != OWL_AMD_F32 && wgrad_halo_fits(p) ? "conv_wgrad_halo_kernel" : "conv_wgrad_kernel";
  if (p.split <= 1) {
    p.c = dw;
    if (tr) {
      KTimer kt(OWL_AMD_KT_BWD_KERNEL, wg_name, st);
      if (in_t == OWL_AMD_F32) {
        run_wgrad<float, float>(p, st);
      } else if (in_t == OWL_AMD_BF16) {
        if (out_t == OWL_AMD_BF16) run_wgrad<bf16_t, bf16_t>(p, st); else run_wgrad<bf16_t, float>(p, st);
      } else {
        if (out_t == OWL_AMD_F16) run_wgrad<f16_t, f16_t>(p, st); else run_wgrad<f16_t, float>(p, st);
      }
      OWL_CHECK_HIP(hipGetLastError());
      return OWL_AMD_OK;
    }
    return dispatch_igemm<2>(p, in_t, out_t, aligned, st);
  }
  size_t need = owl_amd_conv_bwd_kernel_workspace(d, in_t);
  void *ws = workspace;
  if (!ws) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->ws, need, st);
    if (rc) return rc;
    ws = s->ws.ptr;
  } else if (ws_bytes < need) {
    ARG_FAIL("workspace too small: %zu < %zu bytes", ws_bytes, need);
  }
  p.c = ws;
  if (tr) {
    KTimer kt(OWL_AMD_KT_BWD_KERNEL, wg_name, st);
    if (in_t == OWL_AMD_F32) run_wgrad<float, float>(p, st);
    else if (in_t == OWL_AMD_BF16) run_wgrad<bf16_t, float>(p, st);
    else run_wgrad<f16_t, float>(p, st);
    OWL_CHECK_HIP(hipGetLastError());
  } else {
    rc = dispatch_igemm<2>(p, in_t, out_t, aligned, st);
    if (rc) return rc;
  }
  const int64_t n = p.M * p.N;
  KTimer kt(OWL_AMD_KT_AUX, "splitk_reduce_kernel", st);
  if (in_t == OWL_AMD_F64) run_splitk_reduce<double, double>((const double *)ws, (double *)dw, n, p.split, st);
  else if (out_t == OWL_AMD_F32) run_splitk_reduce<float, float>((const float *)ws, (float *)dw, n, p.split, st);
  else if (out_t == OWL_AMD_BF16) run_splitk_reduce<float, bf16_t>((const float *)ws, (bf16_t *)dw, n, p.split, st);
  else run_splitk_reduce<float, f16_t>((const float *)ws, (f16_t *)dw, n, p.split, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_bwd_kernel(const owl_amd_conv_desc *d, int in_t, int out_t,
                                       const void *x, const void *dy, void *dw, void *workspace,
                                       size_t ws_bytes, void *stream) {
  if (is_complex(in_t) || is_complex(out_t))
    return complex_conv(2, d, in_t, out_t, x, dy, dw, workspace, ws_bytes, stream);
  return bwd_kernel_impl(d, in_t, out_t, x, dy, dw, workspace, ws_bytes, (hipStream_t)stream);
}

// Both backward passes of a layer (the pair Owl's reverse rule computes,
// owl_algodiff_ops.ml:2364), on the caller's stream: the weight gradient
// (main kernel and split-K reduction), then the backward input.  Running
// the reduction on a side stream beside the backward-input kernel measured
// slower (64x56x56x64 -> 128 BF16 step 123.6 against 121.7 us): the two
// compete for the CUs' slots instead of filling the other's gaps.
extern "C" int owl_amd_conv_bwd(const owl_amd_conv_desc *d, int in_t, int dx_t, int dw_t, const void *x,
                                const void *w, const void *dy, void *dx, void *dw, void *workspace, size_t ws_bytes,
                                void *stream) {
  int rc = owl_amd_conv_bwd_kernel(d, in_t, dw_t, x, dy, dw, workspace, ws_bytes, stream);
  return rc ? rc : owl_amd_conv_bwd_input(d, in_t, dx_t, w, dy, dx, stream);
}

// ------------------------------------------------------------------ backward kernel without allocation
// The split-K partials of owl_amd_conv_bwd_kernel_noalloc live in this
// file-scope device array: OWL_AMD_NOALLOC_SCRATCH_BYTES per GPU, held by
// every GPU the library runs on until the process ends.  In host code the
// array's name is only a stand-in for it: its device address, one per GPU,
// comes from hipGetSymbolAddress and is what the kernels receive.
static __device__ __align__(256) unsigned char scratch[OWL_AMD_NOALLOC_SCRATCH_BYTES];

namespace {
void *g_scratch_ptr[64];  // device address of scratch on each device (guarded by g_mu)

// Device address of this device's scratch array, looked up once per device
// with that device current.
int scratch_ptr(void **out) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  if (!g_scratch_ptr[dev]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      (void)hipGetLastError();
      owl_amd_set_error("hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
      return OWL_AMD_ERR_HIP;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return OWL_AMD_OK;
}
}  // namespace

extern "C" int owl_amd_conv_bwd_kernel_noalloc(const owl_amd_conv_desc *d, int in_t, int out_t, const void *x,
                                               const void *dy, void *dw, void *stream) {
  if (is_complex(in_t) || is_complex(out_t)) {
    owl_amd_set_error("owl_amd_conv_bwd_kernel_noalloc takes real kinds (complex kinds expand the kernel "
                      "into a library buffer)");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  int rc = check_conv(d);
  if (rc) return rc;
  const size_t need = owl_amd_conv_bwd_kernel_workspace(d, in_t);
  if (need > (size_t)OWL_AMD_NOALLOC_SCRATCH_BYTES)
    ARG_FAIL("split-K partials need %zu bytes, above the %d-byte scratch of owl_amd_conv_bwd_kernel_noalloc; "
             "call owl_amd_conv_bwd_kernel with a workspace", need, (int)OWL_AMD_NOALLOC_SCRATCH_BYTES);
  // every partial the kernels read is written first by the same call, so
  // nothing in the scratch needs clearing between calls
  void *ws = nullptr;
  if (need > 0) {
    rc = scratch_ptr(&ws);
    if (rc) return rc;
  }
  return owl_amd_conv_bwd_kernel(d, in_t, out_t, x, dy, dw, ws, need > 0 ? (size_t)OWL_AMD_NOALLOC_SCRATCH_BYTES : 0,
                                 stream);
}

// ------------------------------------------------------------------ fused epilogue adjoint
extern "C" int owl_amd_maxpool_bias_act_bwd(int dt, int act, int64_t B, int64_t H, int64_t W, int64_t C,
                                           const void *y, const void *pre, const void *dp, void *dz, void *db,
                                           void *stream) {
  if (act != OWL_AMD_ACT_RELU) {
    owl_amd_set_error("the max-pool / relu adjoint takes relu layers (y >= 0: every window has a maximum)");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  if (B < 0 || H < 0 || W < 0 || C <= 0 || C > INT32_MAX || H > INT32_MAX || W > INT32_MAX)
    ARG_FAIL("bad sizes B=%lld H=%lld W=%lld C=%lld", (long long)B, (long long)H, (long long)W, (long long)C);
  if (H % 2 || W % 2) {
    owl_amd_set_error("the max-pool / relu adjoint takes even H and W (every pixel in one window)");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  if (dt != OWL_AMD_F32 && dt != OWL_AMD_F64 && dt != OWL_AMD_BF16 && dt != OWL_AMD_F16) {
    owl_amd_set_error("unsupported dtype %d", dt);
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t nwin = B * (H / 2) * (W / 2);
  if (nwin == 0) {
    if (db) run_zero_fill(db, (size_t)C * (dt == OWL_AMD_F64 ? 8 : 4), st);
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  if (!y || !pre || !dp || !dz) ARG_FAIL("null pointer");
  if (dz == y || dz == pre || dz == dp) ARG_FAIL("dz must not alias y, pre or dp");
  void *partial = nullptr;
  if (db) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->red, (size_t)maxpool_relu_bwd_chunks(nwin) * (size_t)C * (dt == OWL_AMD_F64 ? 8 : 4), st);
    if (rc) return rc;
    partial = s->red.ptr;
  }
  KTimer kt(OWL_AMD_KT_AUX, "maxpool_relu_bwd_kernel", st);
  const int h = (int)H, w = (int)W, c = (int)C;
  if (dt == OWL_AMD_F32)
    run_maxpool_relu_bwd<float, float>(B, h, w, c, (const float *)y, (const float *)pre, (const float *)dp,
                                       (float *)dz, (float *)partial, (float *)db, st);
  else if (dt == OWL_AMD_F64)
    run_maxpool_relu_bwd<double, double>(B, h, w, c, (const double *)y, (const double *)pre, (const double *)dp,
                                         (double *)dz, (double *)partial, (double *)db, st);
  else if (dt == OWL_AMD_BF16)
    run_maxpool_relu_bwd<bf16_t, float>(B, h, w, c, (const bf16_t *)y, (const bf16_t *)pre, (const bf16_t *)dp,
                                        (bf16_t *)dz, (float *)partial, (float *)db, st);
  else
    run_maxpool_relu_bwd<f16_t, float>(B, h, w, c, (const f16_t *)y, (const f16_t *)pre, (const f16_t *)dp,
                                       (f16_t *)dz, (float *)partial, (float *)db, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_bias_act_bwd(int dt, int act, int64_t npix, int64_t C, const void *pre, const void *dy,
                                    void *dz, void *db, void *stream) {
  if (act != OWL_AMD_ACT_NONE && act != OWL_AMD_ACT_RELU) ARG_FAIL("unknown activation %d", act);
  if (npix < 0 || C <= 0 || C > INT32_MAX) ARG_FAIL("bad sizes npix=%lld C=%lld", (long long)npix, (long long)C);
  if (dt != OWL_AMD_F32 && dt != OWL_AMD_F64 && dt != OWL_AMD_BF16 && dt != OWL_AMD_F16) {
    owl_amd_set_error("unsupported dtype %d", dt);
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  if (npix == 0) {  // an empty layer: pre, dy and dz may be null
    if (db) run_zero_fill(db, (size_t)C * (dt == OWL_AMD_F64 ? 8 : 4), st);
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  if (!dy || !dz || (act == OWL_AMD_ACT_RELU && !pre)) ARG_FAIL("null pointer");
  void *partial = nullptr;
  if (db) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    int rc = ensure_state(&s);
    if (rc) return rc;
    rc = ensure_buf(s->red, (size_t)bias_act_chunks(npix) * (size_t)C * (dt == OWL_AMD_F64 ? 8 : 4), st);
    if (rc) return rc;
    partial = s->red.ptr;
  }
  KTimer kt(OWL_AMD_KT_AUX, "bias_act_bwd_kernel", st);
  const int c = (int)C;
  if (dt == OWL_AMD_F32)
    run_bias_act_bwd<float, float>(act, npix, c, (const float *)pre, (const float *)dy, (float *)dz,
                                   (float *)partial, (float *)db, st);
  else if (dt == OWL_AMD_F64)
    run_bias_act_bwd<double, double>(act, npix, c, (const double *)pre, (const double *)dy, (double *)dz,
                                     (double *)partial, (double *)db, st);
  else if (dt == OWL_AMD_BF16)
    run_bias_act_bwd<bf16_t, float>(act, npix, c, (const bf16_t *)pre, (const bf16_t *)dy, (bf16_t *)dz,
                                    (float *)partial, (float *)db, st);
  else
    run_bias_act_bwd<f16_t, float>(act, npix, c, (const f16_t *)pre, (const f16_t *)dy, (f16_t *)dz,
                                   (float *)partial, (float *)db, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// ------------------------------------------------------------------ upsampling
static int upsample(bool backward, int dt, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho, int64_t Wo,
                    int64_t sh, int64_t sw, const void *src, void *dst, void *stream) {
  if (B < 0 || H < 0 || W < 0 || C < 0 || Ho < 0 || Wo < 0 || sh < 1 || sw < 1)
    ARG_FAIL("bad upsampling sizes");
  // complex: the real kernel on the interleaved parts (sums are componentwise)
  if (is_complex(dt)) return upsample(backward, real_of(dt), B, H, W, 2 * C, Ho, Wo, sh, sw, src, dst, stream);
  if (H > INT32_MAX || W > INT32_MAX || C > INT32_MAX || Ho > INT32_MAX || Wo > INT32_MAX)
    ARG_FAIL("upsampling dimension above 2^31");
  if (dt != OWL_AMD_F32 && dt != OWL_AMD_F64 && dt != OWL_AMD_BF16 && dt != OWL_AMD_F16) {
    owl_amd_set_error("unsupported dtype %d", dt);
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  hipStream_t st = (hipStream_t)stream;
  const int64_t n_dst = backward ? B * H * W * C : B * Ho * Wo * C;
  if (n_dst == 0) return OWL_AMD_OK;
  if (!dst || (!src && (backward ? B * Ho * Wo * C : B * H * W * C) > 0)) ARG_FAIL("null pointer");
  if (H == 0 || W == 0) ARG_FAIL("empty input image with a non-empty output");
  if (backward && (Ho == 0 || Wo == 0)) {
    run_zero_fill(dst, (size_t)n_dst * dtype_size(dt), st);
    OWL_CHECK_HIP(hipGetLastError());
    return OWL_AMD_OK;
  }
  KTimer kt(OWL_AMD_KT_AUX, backward ? "upsample_bwd_kernel" : "upsample_fwd_kernel", st);
  const int h = (int)H, w = (int)W, c = (int)C, ho = (int)Ho, wo = (int)Wo, s0 = (int)sh, s1 = (int)sw;
  if (dt == OWL_AMD_F32) run_upsample<float>(backward, (const float *)src, (float *)dst, B, h, w, c, ho, wo, s0, s1, st);
  else if (dt == OWL_AMD_F64)
    run_upsample<double>(backward, (const double *)src, (double *)dst, B, h, w, c, ho, wo, s0, s1, st);
  else if (dt == OWL_AMD_BF16)
    run_upsample<bf16_t>(backward, (const bf16_t *)src, (bf16_t *)dst, B, h, w, c, ho, wo, s0, s1, st);
  else run_upsample<f16_t>(backward, (const f16_t *)src, (f16_t *)dst, B, h, w, c, ho, wo, s0, s1, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_upsampling2d_fwd(int dt, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho, int64_t Wo,
                                        int64_t sh, int64_t sw, const void *x, void *y, void *stream) {
  return upsample(false, dt, B, H, W, C, Ho, Wo, sh, sw, x, y, stream);
}

extern "C" int owl_amd_upsampling2d_bwd(int dt, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho, int64_t Wo,
                                        int64_t sh, int64_t sw, const void *dy, void *dx, void *stream) {
  return upsample(true, dt, B, H, W, C, Ho, Wo, sh, sw, dy, dx, stream);
}

// ------------------------------------------------------------------ device API: FP8
extern "C" int owl_amd_fp8_quantize(int src_t, const void *src, int64_t n, void *dst_q, float *scale,
                                    void *stream) {
  if (n < 0) ARG_FAIL("negative element count");
  if (n > 0 && (!src || !dst_q)) ARG_FAIL("null pointer");
  if (!scale) ARG_FAIL("null scale pointer");
  hipStream_t st = (hipStream_t)stream;
  if (src_t == OWL_AMD_F32) run_fp8_quantize<float>((const float *)src, n, (unsigned char *)dst_q, scale, st);
  else if (src_t == OWL_AMD_BF16) run_fp8_quantize<bf16_t>((const bf16_t *)src, n, (unsigned char *)dst_q, scale, st);
  else {
    owl_amd_set_error("fp8 quantize: source must be F32 or BF16");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_fp8_encode(const float *src, int64_t n, void *dst, void *stream) {
  if (n < 0 || (n > 0 && (!src || !dst))) ARG_FAIL("bad arguments");
  if (n) run_e4m3_encode(src, n, (unsigned char *)dst, (hipStream_t)stream);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_fwd_fp8(const owl_amd_conv_desc *d, int out_t, const void *xq, const void *wq,
                                    const float *x_scale, const float *w_scale, void *y, void *stream) {
  int rc = check_conv(d);
  if (rc) return rc;
  ConvParams p = make_params(d);
  p.M = d->B * d->Ho * d->Wo * d->Do;
  p.N = p.Co;
  p.K = d->kh * d->kw * d->kd * d->Ci;
  if (p.M == 0 || p.N == 0) return OWL_AMD_OK;
  if (!xq || !wq || !y || !x_scale || !w_scale) ARG_FAIL("null pointer");
  p.a = xq; p.b = wq; p.c = y; p.ldc = p.Co;
  p.scale_a = x_scale;
  p.scale_b = w_scale;
  bool aligned = p.Ci % 16 == 0 && p.Co % 16 == 0 && aligned16(xq, wq, nullptr);
  hipStream_t st = (hipStream_t)stream;
  // 2-D stride-1 layers with Ci % 64 == 0 and Co % 8 == 0:
  // the halo kernel with FP8 MFMAs, on the kernel repacked to [tap][co][ci]
  // in a per-device buffer (calls that overlap on one device race on it)
  if (fast_path_enabled() && halo_enabled() && (out_t == OWL_AMD_F32 || out_t == OWL_AMD_BF16) &&
      aligned16(xq, wq, y) && halo_fits_fp8(p, (int)dtype_size(out_t))) {
    const int64_t T = (int64_t)p.kh * p.kw * p.kd;
    void *wp = nullptr;
    rc = f8w_scratch((size_t)(T * p.Ci * p.Co), &wp, st);
    if (rc) return rc;
    run_fp8_pack_krows((const unsigned char *)wq, (unsigned char *)wp, T, p.Ci, p.Co, st);
    const bool ok = out_t == OWL_AMD_F32 ? run_halo_fp8<float>(p, wp, st) : run_halo_fp8<bf16_t>(p, wp, st);
    if (ok) {
      OWL_CHECK_HIP(hipGetLastError());
      return OWL_AMD_OK;
    }
  }
  KTimer kt(OWL_AMD_KT_FWD, "conv_igemm_kernel", st);
  if (out_t == OWL_AMD_F32) run_igemm<0, fp8e4m3_t, float>(p, aligned, st);
  else if (out_t == OWL_AMD_BF16) run_igemm<0, fp8e4m3_t, bf16_t>(p, aligned, st);
  else {
    owl_amd_set_error("fp8 forward: output must be F32 or BF16");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// ------------------------------------------------------------------ device API: pool
static int check_pool(const owl_amd_pool_desc *d, int t) {
  if (!d) ARG_FAIL("null descriptor");
  const int64_t v[] = {d->B, d->H, d->W, d->D, d->C, d->Ho, d->Wo, d->Do, d->ph, d->pw, d->pd};
  for (int64_t x : v)
    if (!fits_int(x)) ARG_FAIL("pool size out of range");
  if (d->kh < 1 || d->kw < 1 || d->kd < 1) ARG_FAIL("pool window must be >= 1");
  if (d->sh < 1 || d->sw < 1 || d->sd < 1) ARG_FAIL("pool stride must be >= 1");
  if (!fits_int(d->B * d->H * d->W * d->D * d->C) || !fits_int(d->B * d->Ho * d->Wo * d->Do * d->C))
    ARG_FAIL("pool tensor too large for 32-bit indexing");
  if (t != OWL_AMD_F32 && t != OWL_AMD_F64 && t != OWL_AMD_BF16 && t != OWL_AMD_F16 && !is_complex(t)) {
    owl_amd_set_error("pooling supports F32, F64, BF16, F16 and (average only) C32, C64");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  return OWL_AMD_OK;
}

static PoolParams make_pool(const owl_amd_pool_desc *d) {
  PoolParams p;
  p.B = (int)d->B; p.H = (int)d->H; p.W = (int)d->W; p.D = (int)d->D; p.C = (int)d->C;
  p.kh = (int)d->kh; p.kw = (int)d->kw; p.kd = (int)d->kd;
  p.Ho = (int)d->Ho; p.Wo = (int)d->Wo; p.Do = (int)d->Do;
  p.sh = (int)d->sh; p.sw = (int)d->sw; p.sd = (int)d->sd;
  p.ph = (int)d->ph; p.pw = (int)d->pw; p.pd = (int)d->pd;
  return p;
}

// Complex pooling: the average is the real one on the interleaved parts (the
// complex sum and the division by the count are componentwise); the maximum
// is not provided (Owl's starts from -inf + 0i, whose modulus is +inf, so
// its modulus-then-argument comparison never replaces it).
static int complex_pool(const owl_amd_pool_desc *d, int t, int is_max, owl_amd_pool_desc *r) {
  if (is_max) {
    owl_amd_set_error("complex max pooling is not provided: Owl's starts from -inf + 0i (modulus +inf), "
                      "which its modulus-then-argument comparison never replaces");
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  *r = *d;
  r->C *= 2;
  (void)t;
  return OWL_AMD_OK;
}

extern "C" int owl_amd_pool_fwd(const owl_amd_pool_desc *d, int t, int is_max, const void *x,
                                void *y, void *stream) {
  if (d && is_complex(t)) {
    owl_amd_pool_desc r;
    int rc = complex_pool(d, t, is_max, &r);
    return rc ? rc : owl_amd_pool_fwd(&r, real_of(t), 0, x, y, stream);
  }
  int rc = check_pool(d, t);
  if (rc) return rc;
  PoolParams p = make_pool(d);
  if ((int64_t)p.B * p.Ho * p.Wo * p.Do * p.C == 0) return OWL_AMD_OK;
  if (!x || !y) ARG_FAIL("null pointer");
  const int mode = is_max ? 1 : 0;
  hipStream_t st = (hipStream_t)stream;
  if (t == OWL_AMD_F32) run_pool_fwd<float>(p, mode, (const float *)x, (float *)y, nullptr, st);
  else if (t == OWL_AMD_F64) run_pool_fwd<double>(p, mode, (const double *)x, (double *)y, nullptr, st);
  else if (t == OWL_AMD_BF16) run_pool_fwd<bf16_t>(p, mode, (const bf16_t *)x, (bf16_t *)y, nullptr, st);
  else run_pool_fwd<f16_t>(p, mode, (const f16_t *)x, (f16_t *)y, nullptr, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_pool_argmax(const owl_amd_pool_desc *d, int t, const void *x, void *y,
                                   int64_t *argmax, void *stream) {
  if (d && is_complex(t)) {
    owl_amd_pool_desc r;
    return complex_pool(d, t, 1, &r);
  }
  int rc = check_pool(d, t);
  if (rc) return rc;
  PoolParams p = make_pool(d);
  if ((int64_t)p.B * p.Ho * p.Wo * p.Do * p.C == 0) return OWL_AMD_OK;
  if (!x || !y || !argmax) ARG_FAIL("null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (t == OWL_AMD_F32) run_pool_fwd<float>(p, 2, (const float *)x, (float *)y, argmax, st);
  else if (t == OWL_AMD_F64) run_pool_fwd<double>(p, 2, (const double *)x, (double *)y, argmax, st);
  else if (t == OWL_AMD_BF16) run_pool_fwd<bf16_t>(p, 2, (const bf16_t *)x, (bf16_t *)y, argmax, st);
  else run_pool_fwd<f16_t>(p, 2, (const f16_t *)x, (f16_t *)y, argmax, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// True if some window along one axis has no in-bounds cell.
static bool axis_has_empty(int64_t n, int64_t k, int64_t s, int64_t p, int64_t o) {
  if (o <= 0) return false;
  return (k - p <= 0) || ((o - 1) * s - p >= n);
}

// The pooling backward with its orphan flag (an int that is zero between
// calls) given, or the device's own when flag is null.
static int pool_bwd_flag(const owl_amd_pool_desc *d, int t, int is_max, const void *x, const void *dy, void *dx,
                         int *flag, void *stream) {
  if (d && is_complex(t)) {
    owl_amd_pool_desc r;
    int rc = complex_pool(d, t, is_max, &r);
    return rc ? rc : pool_bwd_flag(&r, real_of(t), 0, x, dy, dx, flag, stream);
  }
  int rc = check_pool(d, t);
  if (rc) return rc;
  PoolParams p = make_pool(d);
  hipStream_t st = (hipStream_t)stream;
  if ((int64_t)p.B * p.H * p.W * p.D * p.C == 0) return OWL_AMD_OK;
  if (!x || !dy || !dx) ARG_FAIL("null pointer");
  if (!flag) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    DeviceState *s;
    rc = ensure_state(&s);
    if (rc) return rc;
    flag = s->flag;
  }
  bool empty = axis_has_empty(d->H, d->kh, d->sh, d->ph, d->Ho) ||
               axis_has_empty(d->W, d->kw, d->sw, d->pw, d->Wo) ||
               axis_has_empty(d->D, d->kd, d->sd, d->pd, d->Do);
  if (is_max && empty && (int64_t)p.Ho * p.Wo * p.Do * p.B * p.C > 0) run_set_flag(flag, st);
  const bool mx = is_max != 0;
  if (t == OWL_AMD_F32) run_pool_bwd<float>(p, mx, (const float *)x, (const float *)dy, (float *)dx, flag, st);
  else if (t == OWL_AMD_F64) run_pool_bwd<double>(p, mx, (const double *)x, (const double *)dy, (double *)dx, flag, st);
  else if (t == OWL_AMD_BF16) run_pool_bwd<bf16_t>(p, mx, (const bf16_t *)x, (const bf16_t *)dy, (bf16_t *)dx, flag, st);
  else run_pool_bwd<f16_t>(p, mx, (const f16_t *)x, (const f16_t *)dy, (f16_t *)dx, flag, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

extern "C" int owl_amd_pool_bwd(const owl_amd_pool_desc *d, int t, int is_max, const void *x,
                                const void *dy, void *dx, void *stream) {
  return pool_bwd_flag(d, t, is_max, x, dy, dx, nullptr, stream);
}

// ------------------------------------------------------------------ Owl drop-ins
namespace {

// Leading pad exactly as Owl's C derives it:  p = (s*(o-1) + k_eff - n) / 2
// (C truncation), clamped at 0  (owl_ndarray_conv_impl.h:288-294, 502-505).
int64_t owl_pad(int64_t s, int64_t o, int64_t k_eff, int64_t n) {
  int64_t t = s * (o - 1) + k_eff - n;
  int64_t q = t / 2;  // C++ truncates toward zero like C
  return q < 0 ? 0 : q;
}

struct HostArg {
  const void *src;  // host data to upload (nullptr: output only)
  size_t bytes;
};

// ---- drop-in staging.  A drop-in call copies Owl's (pageable) Bigarray
// data through pinned host chunks: the host copies one chunk into pinned
// memory with several threads while the DMA engine moves the previous one,
// and the other way round for results.  Each call takes a staging context
// of its own (stream, device buffers, split-K workspace, pooling flag,
// pinned chunks, copy threads), so calls from different host threads on one
// device run concurrently; the library lock is held only to take and return
// a context.  Complex kinds keep the lock for the whole call (their
// expanded kernels live in one per-device buffer).
constexpr size_t kChunk = 8u << 20;  // pinned chunk (two per context)
constexpr int kMaxCtx = 8;           // staging contexts per device

// Parallel memcpy: the caller and up to nw worker threads each copy one
// stripe of at least 1 MiB.
class CopyPool {
 public:
  explicit CopyPool(int nw) {
    for (int i = 0; i < nw; ++i) th_.emplace_back([this, i] { work(i); });
  }
  ~CopyPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto &t : th_) t.join();
  }
  void copy(void *dst, const void *src, size_t n) {
    const size_t stripe_min = 1u << 20;
    int parts = (int)(n / stripe_min);
    if (parts > (int)th_.size() + 1) parts = (int)th_.size() + 1;
    if (parts <= 1) {
      memcpy(dst, src, n);
      return;
    }
    const size_t per = (n / parts + 63) & ~(size_t)63;
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = (char *)dst;
      src_ = (const char *)src;
      n_ = n;
      per_ = per;
      parts_ = parts;
      pending_ = parts - 1;
      ++gen_;
    }
    cv_.notify_all();
    memcpy(dst, src, per < n ? per : n);  // stripe 0
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

 private:
  void work(int i) {
    uint64_t seen = 0;
    for (;;) {
      char *d;
      const char *s;
      size_t n, per;
      int parts;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        d = dst_; s = src_; n = n_; per = per_; parts = parts_;
      }
      const int part = i + 1;
      if (part < parts) {
        const size_t off = (size_t)part * per;
        if (off < n) memcpy(d + off, s + off, (off + per < n && part + 1 < parts) ? per : n - off);
        std::lock_guard<std::mutex> lk(m_);
        if (--pending_ == 0) done_.notify_all();
      }
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  bool stop_ = false;
  uint64_t gen_ = 0;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t n_ = 0, per_ = 0;
  int parts_ = 0, pending_ = 0;
};

struct DropinCtx {
  int dev = -1;
  bool busy = false;
  hipStream_t stream = nullptr;
  int *flag = nullptr;
  DevBuf slot[4];
  DevBuf ws;
  void *pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  CopyPool *copier = nullptr;
  // device work of repeated calls, captured once per (call kind, shape,
  // dtype, device buffers) and replayed: see dropin_launch
  std::map<std::string, hipGraphExec_t> graphs;
  std::set<std::string> seen;  // call keys run once directly on this context
};
// Graphs and seen keys kept per context before the cache starts over.
constexpr size_t kMaxGraphs = 64, kMaxSeen = 4096;
std::atomic<int64_t> g_graph_captures{0}, g_graph_replays{0};

void drop_graphs(DropinCtx &c) {
  for (auto &kv : c.graphs) (void)hipGraphExecDestroy(kv.second);
  c.graphs.clear();
}

// Drop-in graph capture: off unless OWL_AMD_DROPIN_GRAPHS=1 or
// owl_amd_dropin_graphs(1).  Measured on MI355X (FP32 drop-ins on host
// arrays): an [8,32,32,3] (*) [3,3,3,16] forward 78.7 us with graphs against
// 76.8 without, a [8,28,28,64] (*) [3,3,64,64] fwd + bwd + pool step 843
// against 790 us: replaying the captured copies and kernels is no faster
// than enqueueing them on this stack.
std::atomic<int> g_dropin_graphs{-1};
bool dropin_graphs_enabled() {
  int v = g_dropin_graphs.load(std::memory_order_relaxed);
  if (v < 0) {
    v = env_int("OWL_AMD_DROPIN_GRAPHS", 0) != 0;
    g_dropin_graphs.store(v);
  }
  return v != 0;
}
std::vector<DropinCtx *> g_ctx;  // guarded by g_mu
std::condition_variable_any g_ctx_free;

int copy_threads() {
  static const int n = [] {
    const int env = env_int("OWL_AMD_COPY_THREADS", 0);
    if (env > 0) return env > 32 ? 32 : env;
    const unsigned hw = std::thread::hardware_concurrency();
    const int t = hw >= 16 ? 4 : (hw >= 8 ? 3 : 1);
    return t;
  }();
  return n;
}

int make_ctx(int dev, DropinCtx **out) {
  DropinCtx *c = new DropinCtx();
  c->dev = dev;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&c->flag, 256);
  if (e == hipSuccess) e = hipMemset(c->flag, 0, 256);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipHostMalloc(&c->pin[i], kChunk, hipHostMallocDefault);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipGetLastError();
    owl_amd_set_error("drop-in staging set-up: %s", hipGetErrorString(e));
    for (int i = 0; i < 2; ++i) {
      if (c->pin[i]) (void)hipHostFree(c->pin[i]);
      if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    }
    if (c->flag) (void)hipFree(c->flag);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return OWL_AMD_ERR_HIP;
  }
  const int nt = copy_threads();
  c->copier = new CopyPool(nt > 1 ? nt - 1 : 0);
  *out = c;
  return OWL_AMD_OK;
}

// Take a free staging context of the current device (creating one, up to
// kMaxCtx; otherwise wait for one to be returned).
int acquire_ctx(DropinCtx **out) {
  int dev;
  int rc = current_device(&dev);
  if (rc) return rc;
  std::unique_lock<std::recursive_mutex> lk(g_mu);
  for (;;) {
    int count = 0;
    for (DropinCtx *c : g_ctx) {
      if (c->dev != dev) continue;
      ++count;
      if (!c->busy) {
        c->busy = true;
        *out = c;
        return OWL_AMD_OK;
      }
    }
    if (count < kMaxCtx) {
      DropinCtx *c;
      rc = make_ctx(dev, &c);
      if (rc) return rc;
      c->busy = true;
      g_ctx.push_back(c);
      *out = c;
      return OWL_AMD_OK;
    }
    g_ctx_free.wait(lk);
  }
}

void release_ctx(DropinCtx *c) {
  {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    c->busy = false;
  }
  g_ctx_free.notify_one();
}

// host -> device through the pinned chunks (the host copy of chunk k + 1
// overlaps the DMA of chunk k)
int upload(DropinCtx &c, void *dst, const void *src, size_t n) {
  for (size_t off = 0, k = 0; off < n; off += kChunk, ++k) {
    const size_t len = n - off < kChunk ? n - off : kChunk;
    const int b = (int)(k & 1);
    OWL_CHECK_HIP(hipEventSynchronize(c.ev[b]));  // the DMA that last read this chunk is done
    c.copier->copy(c.pin[b], (const char *)src + off, len);
    OWL_CHECK_HIP(hipMemcpyAsync((char *)dst + off, c.pin[b], len, hipMemcpyHostToDevice, c.stream));
    OWL_CHECK_HIP(hipEventRecord(c.ev[b], c.stream));
  }
  return OWL_AMD_OK;
}

// device -> host, after the work already on the stream (the DMA of chunk
// k + 1 overlaps the host copy of chunk k)
int download(DropinCtx &c, void *dst, const void *src, size_t n) {
  size_t prev_off = 0, prev_len = 0;
  int prev_b = -1;
  for (size_t off = 0, k = 0; off < n; off += kChunk, ++k) {
    const size_t len = n - off < kChunk ? n - off : kChunk;
    const int b = (int)(k & 1);
    OWL_CHECK_HIP(hipMemcpyAsync(c.pin[b], (const char *)src + off, len, hipMemcpyDeviceToHost, c.stream));
    OWL_CHECK_HIP(hipEventRecord(c.ev[b], c.stream));
    if (prev_b >= 0) {
      OWL_CHECK_HIP(hipEventSynchronize(c.ev[prev_b]));
      c.copier->copy((char *)dst + prev_off, c.pin[prev_b], prev_len);
    }
    prev_off = off;
    prev_len = len;
    prev_b = b;
  }
  if (prev_b >= 0) {
    OWL_CHECK_HIP(hipEventSynchronize(c.ev[prev_b]));
    c.copier->copy((char *)dst + prev_off, c.pin[prev_b], prev_len);
  }
  return OWL_AMD_OK;
}

// The device work of a drop-in call.  Calls with a key (the call kind,
// shape and dtype: the bytes of the arguments the kernels depend on) are
// run directly the first time a context sees the key -- which sizes its
// workspace -- then captured into a hipGraph on the context's stream
// (thread-local capture: other host threads keep allocating freely) and
// replayed by every later call whose device buffers (slots, workspace,
// flag) are the same addresses.  The work reads its inputs from those
// buffers (or from the context's pinned chunks, for small calls), so a
// replay after fresh host copies computes the new call.  Not used while
// kernel timing is on (its events would be captured).
template <typename F>
int dropin_launch(DropinCtx &c, const std::string *gkey, void **dev, F &launch) {
  if (!gkey || !dropin_graphs_enabled() || g_kt_on.load(std::memory_order_relaxed)) return launch(dev, c.stream, c);
  std::string k = *gkey;
  for (int i = 0; i < 4; ++i) k.append((const char *)&dev[i], sizeof(void *));
  k.append((const char *)&c.ws.ptr, sizeof(void *));
  k.append((const char *)&c.flag, sizeof(void *));
  auto it = c.graphs.find(k);
  if (it != c.graphs.end()) {
    OWL_CHECK_HIP(hipGraphLaunch(it->second, c.stream));
    g_graph_replays.fetch_add(1, std::memory_order_relaxed);
    return OWL_AMD_OK;
  }
  if (c.seen.size() >= kMaxSeen) c.seen.clear();
  if (c.seen.insert(*gkey).second) return launch(dev, c.stream, c);
  if (c.graphs.size() >= kMaxGraphs) {
    OWL_CHECK_HIP(hipStreamSynchronize(c.stream));
    drop_graphs(c);
  }
  hipGraph_t graph = nullptr;
  OWL_CHECK_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
  int rc = launch(dev, c.stream, c);
  hipError_t e = hipStreamEndCapture(c.stream, &graph);
  hipGraphExec_t exec = nullptr;
  if (rc == OWL_AMD_OK && e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (graph) (void)hipGraphDestroy(graph);
  if (rc) return rc;
  if (e != hipSuccess || !exec) {
    // not capturable: run it directly (and try again on the next call)
    (void)hipGetLastError();
    return launch(dev, c.stream, c);
  }
  c.graphs[k] = exec;
  g_graph_captures.fetch_add(1, std::memory_order_relaxed);
  OWL_CHECK_HIP(hipGraphLaunch(exec, c.stream));
  g_graph_replays.fetch_add(1, std::memory_order_relaxed);
  return OWL_AMD_OK;
}

// Key bytes of a drop-in call: a call-kind tag and the values its kernels
// depend on.
template <typename... Ts> std::string dropin_key(char tag, const Ts &...v) {
  std::string k(1, tag);
  (k.append((const char *)&v, sizeof(v)), ...);
  return k;
}

// One synchronous drop-in call: upload the host inputs into the context's
// device slots, run launch(slots, stream, context), download the outputs.
template <typename F>
int run_dropin_ctx(DropinCtx &c, std::initializer_list<HostArg> args, int out_slot, void *out_host,
                   size_t out_bytes, F launch, int extra_out_slot, void *extra_host, size_t extra_bytes,
                   const std::string *gkey) {
  void *dev[4] = {nullptr, nullptr, nullptr, nullptr};
  int i = 0, rc;
  // Small calls (every input fits one pinned chunk, every output the other):
  // the host copies the inputs into pin[0] first, and the device work --
  // the input DMAs, the kernels, the output DMAs into pin[1] -- is one
  // sequence on the stream, captured and replayed as a whole with a key.
  // The previous call on this context ended with a stream synchronisation,
  // so both chunks are free.
  size_t in_off[4] = {0, 0, 0, 0}, in_tot = 0;
  for (const HostArg &a : args) {
    if (a.src && a.bytes) in_off[i] = in_tot, in_tot += (a.bytes + 255) & ~(size_t)255;
    ++i;
  }
  const size_t extra_off = (out_bytes + 255) & ~(size_t)255;
  const bool small = in_tot <= kChunk && extra_off + (extra_out_slot >= 0 ? extra_bytes : 0) <= kChunk;
  i = 0;
  for (const HostArg &a : args) {
    rc = ensure_buf(c.slot[i], a.bytes, c.stream, false);
    if (rc) return rc;
    dev[i] = c.slot[i].ptr;
    if (a.src && a.bytes) {
      if (small) c.copier->copy((char *)c.pin[0] + in_off[i], a.src, a.bytes);
      else if ((rc = upload(c, dev[i], a.src, a.bytes))) return rc;
    }
    ++i;
  }
  if (small) {
    std::vector<HostArg> in(args);
    auto whole = [&](void **dv, hipStream_t st, DropinCtx &cx) -> int {
      for (size_t j = 0; j < in.size(); ++j)
        if (in[j].src && in[j].bytes)
          OWL_CHECK_HIP(hipMemcpyAsync(dv[j], (char *)cx.pin[0] + in_off[j], in[j].bytes, hipMemcpyHostToDevice, st));
      int r = launch(dv, st, cx);
      if (r) return r;
      if (out_bytes) OWL_CHECK_HIP(hipMemcpyAsync(cx.pin[1], dv[out_slot], out_bytes, hipMemcpyDeviceToHost, st));
      if (extra_out_slot >= 0 && extra_bytes)
        OWL_CHECK_HIP(hipMemcpyAsync((char *)cx.pin[1] + extra_off, dv[extra_out_slot], extra_bytes,
                                     hipMemcpyDeviceToHost, st));
      return OWL_AMD_OK;
    };
    rc = dropin_launch(c, gkey, dev, whole);
    OWL_CHECK_HIP(hipStreamSynchronize(c.stream));
    if (rc) return rc;
    if (out_bytes) c.copier->copy(out_host, c.pin[1], out_bytes);
    if (extra_out_slot >= 0 && extra_bytes) c.copier->copy(extra_host, (char *)c.pin[1] + extra_off, extra_bytes);
    return OWL_AMD_OK;
  }
  rc = dropin_launch(c, gkey, dev, launch);
  if (rc) {
    (void)hipStreamSynchronize(c.stream);
    return rc;
  }
  if (out_bytes && (rc = download(c, out_host, dev[out_slot], out_bytes))) return rc;
  if (extra_out_slot >= 0 && extra_bytes && (rc = download(c, extra_host, dev[extra_out_slot], extra_bytes)))
    return rc;
  OWL_CHECK_HIP(hipStreamSynchronize(c.stream));
  return OWL_AMD_OK;
}

// gkey: the call's graph key (dropin_launch), or null to launch directly
// every time (calls whose kernels use library-wide buffers: those may move).
template <typename F>
int run_dropin(std::initializer_list<HostArg> args, int out_slot, void *out_host, size_t out_bytes,
               F launch, int extra_out_slot = -1, void *extra_host = nullptr, size_t extra_bytes = 0,
               bool whole_call_locked = false, const std::string *gkey = nullptr) {
  DropinCtx *c;
  int rc = acquire_ctx(&c);
  if (rc) return rc;
  if (whole_call_locked) {
    std::lock_guard<std::recursive_mutex> lk(g_mu);
    rc = run_dropin_ctx(*c, args, out_slot, out_host, out_bytes, launch, extra_out_slot, extra_host, extra_bytes,
                        gkey);
  } else {
    rc = run_dropin_ctx(*c, args, out_slot, out_host, out_bytes, launch, extra_out_slot, extra_host, extra_bytes,
                        gkey);
  }
  if (rc) (void)hipStreamSynchronize(c->stream);
  release_ctx(c);
  return rc;
}

}  // namespace

extern "C" int owl_amd_cache_stats(int64_t *n, int64_t *bytes) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int64_t cnt = 0, tot = 0;
  for (auto &s : g_dev) {
    if (!s.ready) continue;
    for (const DevBuf *b : {&s.ws, &s.red, &s.cplx, &s.f8w, &s.ytmp, &s.dsub, &s.wsub, &s.xsub})
      if (b->ptr) { cnt++; tot += (int64_t)b->bytes; }
  }
  for (DropinCtx *c : g_ctx)
    for (const DevBuf *b : {&c->slot[0], &c->slot[1], &c->slot[2], &c->slot[3], &c->ws})
      if (b->ptr) { cnt++; tot += (int64_t)b->bytes; }
  if (n) *n = cnt;
  if (bytes) *bytes = tot;
  return OWL_AMD_OK;
}

extern "C" int owl_amd_dropin_graphs(int mode) {
  if (mode == 0 || mode == 1) g_dropin_graphs.store(mode);
  return dropin_graphs_enabled() ? 1 : 0;
}

extern "C" int owl_amd_dropin_graph_stats(int64_t *graphs, int64_t *captures, int64_t *replays) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int64_t n = 0;
  for (DropinCtx *c : g_ctx) n += (int64_t)c->graphs.size();
  if (graphs) *graphs = n;
  if (captures) *captures = g_graph_captures.load();
  if (replays) *replays = g_graph_replays.load();
  return OWL_AMD_OK;
}

extern "C" void owl_amd_cache_release(void) {
  std::lock_guard<std::recursive_mutex> lk(g_mu);
  int cur = -1;
  (void)hipGetDevice(&cur);
  for (int dev = 0; dev < 64; ++dev) {
    DeviceState &s = g_dev[dev];
    if (!s.ready) continue;
    if (hipSetDevice(dev) != hipSuccess) continue;
    (void)hipStreamSynchronize(s.stream);
    (void)hipDeviceSynchronize();
    for (DevBuf *b : {&s.ws, &s.red, &s.cplx, &s.f8w, &s.ytmp, &s.dsub, &s.wsub, &s.xsub}) {
      if (b->ptr) (void)hipFree(b->ptr);
      *b = DevBuf();
    }
  }
  for (void *p : g_retired) (void)hipFree(p);
  g_retired.clear();
  // idle drop-in staging contexts
  std::vector<DropinCtx *> keep;
  for (DropinCtx *c : g_ctx) {
    if (c->busy) {
      keep.push_back(c);
      continue;
    }
    if (hipSetDevice(c->dev) == hipSuccess) (void)hipStreamSynchronize(c->stream);
    drop_graphs(*c);
    for (DevBuf *b : {&c->slot[0], &c->slot[1], &c->slot[2], &c->slot[3], &c->ws})
      if (b->ptr) (void)hipFree(b->ptr);
    for (int i = 0; i < 2; ++i) {
      (void)hipHostFree(c->pin[i]);
      (void)hipEventDestroy(c->ev[i]);
    }
    (void)hipFree(c->flag);
    (void)hipStreamDestroy(c->stream);
    delete c->copier;
    delete c;
  }
  g_ctx.swap(keep);
  if (cur >= 0) (void)hipSetDevice(cur);
}

namespace {

template <typename T> constexpr int dt();
template <> constexpr int dt<float>() { return OWL_AMD_F32; }
template <> constexpr int dt<double>() { return OWL_AMD_F64; }
template <> constexpr int dt<owl_amd_complex32>() { return OWL_AMD_C32; }
template <> constexpr int dt<owl_amd_complex64>() { return OWL_AMD_C64; }

enum ConvPass { FWD = 0, BWD_IN = 1, BWD_KER = 2 };

// Common body of every conv drop-in.  `padding` < 0 means "backward: infer
// the pad from the shapes"; 0 = SAME, 1 = VALID (forward).
template <typename T>
int conv_dropin(ConvPass pass, T *input, T *kernel, T *output, int64_t B, int64_t H, int64_t W,
                int64_t D, int64_t Ci, int64_t kh, int64_t kw, int64_t kd, int64_t Ho, int64_t Wo,
                int64_t Do, int64_t Co, int64_t sh, int64_t sw, int64_t sd, int64_t rh, int64_t rw,
                int64_t rd, int64_t padding) {
  owl_amd_conv_desc d;
  d.B = B; d.H = H; d.W = W; d.D = D; d.Ci = Ci;
  d.kh = kh; d.kw = kw; d.kd = kd; d.Co = Co;
  d.Ho = Ho; d.Wo = Wo; d.Do = Do;
  d.sh = sh; d.sw = sw; d.sd = sd;
  d.rh = rh; d.rw = rw; d.rd = rd;
  if (padding == 1) {
    d.ph = d.pw = d.pd = 0;
  } else {
    d.ph = owl_pad(sh, Ho, kh + (kh - 1) * (rh - 1), H);
    d.pw = owl_pad(sw, Wo, kw + (kw - 1) * (rw - 1), W);
    d.pd = owl_pad(sd, Do, kd + (kd - 1) * (rd - 1), D);
  }
  int rc = check_conv(&d);
  if (rc) return rc;
  if (!input || !kernel || !output) ARG_FAIL("null pointer");
  const size_t nx = (size_t)(B * H * W * D * Ci) * sizeof(T);
  const size_t nw = (size_t)(kh * kw * kd * Ci * Co) * sizeof(T);
  const size_t ny = (size_t)(B * Ho * Wo * Do * Co) * sizeof(T);
  const int t = dt<T>();
  // complex kernels (expanded into one per-device buffer), strided backward
  // inputs (their phase grids in per-device buffers), strided forwards that
  // read one phase of x or go through space-to-depth, and split-K forwards
  // (per-device buffers all): the whole call locked, not captured
  const bool whole = is_complex(t) || (pass == BWD_IN && (strided_1x1_bwd_input(&d) || strided_phase_bwd_input(&d))) ||
                     (pass == FWD && (strided_fwd_one_phase(&d) || s2d_uses_buffers(&d, t) || fwd_splitk_applies(&d, t)));
  const std::string key = dropin_key('c', (int)pass, t, d);
  const std::string *gk = whole ? nullptr : &key;
  if (pass == FWD) {
    // Owl memsets the output first (owl_ndarray_conv_impl.h:285): an empty
    // GEMM (K == 0 cannot occur, M == 0 writes nothing) leaves zeros.
    return run_dropin({{input, nx}, {kernel, nw}, {nullptr, ny}}, 2, output, ny,
                      [&](void **dv, hipStream_t st, DropinCtx &) {
                        return owl_amd_conv_fwd(&d, t, t, dv[0], dv[1], dv[2], st);
                      }, -1, nullptr, 0, whole, gk);
  } else if (pass == BWD_IN) {
    return run_dropin({{kernel, nw}, {output, ny}, {nullptr, nx}}, 2, input, nx,
                      [&](void **dv, hipStream_t st, DropinCtx &) {
                        return owl_amd_conv_bwd_input(&d, t, t, dv[0], dv[1], dv[2], st);
                      }, -1, nullptr, 0, whole, gk);
  } else {
    return run_dropin({{input, nx}, {output, ny}, {nullptr, nw}}, 2, kernel, nw,
                      [&](void **dv, hipStream_t st, DropinCtx &c) {
                        // the split-K partials in the context's own workspace
                        const size_t need = owl_amd_conv_bwd_kernel_workspace(&d, t);
                        int r = ensure_buf(c.ws, need, st, false);
                        if (r) return r;
                        return owl_amd_conv_bwd_kernel(&d, t, t, dv[0], dv[1], dv[2], need ? c.ws.ptr : nullptr,
                                                       need, st);
                      }, -1, nullptr, 0, whole, gk);
  }
}

template <typename T>
int pool_dropin_fwd(bool is_max, T *input, T *output, int64_t B, int64_t H, int64_t W, int64_t D,
                    int64_t C, int64_t kh, int64_t kw, int64_t kd, int64_t Ho, int64_t Wo,
                    int64_t Do, int64_t sh, int64_t sw, int64_t sd, int64_t padding) {
  owl_amd_pool_desc d;
  d.B = B; d.H = H; d.W = W; d.D = D; d.C = C;
  d.kh = kh; d.kw = kw; d.kd = kd;
  d.Ho = Ho; d.Wo = Wo; d.Do = Do;
  d.sh = sh; d.sw = sw; d.sd = sd;
  if (padding == 1) {
    d.ph = d.pw = d.pd = 0;
  } else {  // owl_ndarray_pool_impl.h:42-48, :263-271
    d.ph = owl_pad(sh, Ho, kh, H);
    d.pw = owl_pad(sw, Wo, kw, W);
    d.pd = owl_pad(sd, Do, kd, D);
  }
  int rc = check_pool(&d, dt<T>());
  if (rc) return rc;
  owl_amd_pool_desc r;
  if (is_complex(dt<T>()) && (rc = complex_pool(&d, dt<T>(), is_max, &r))) return rc;
  if (!input || !output) ARG_FAIL("null pointer");
  const size_t nx = (size_t)(B * H * W * D * C) * sizeof(T);
  const size_t ny = (size_t)(B * Ho * Wo * Do * C) * sizeof(T);
  const std::string key = dropin_key('p', dt<T>(), (int)is_max, d);
  return run_dropin({{input, nx}, {nullptr, ny}}, 1, output, ny, [&](void **dv, hipStream_t st, DropinCtx &c) {
    return owl_amd_pool_fwd(&d, dt<T>(), is_max, dv[0], dv[1], st);
  }, -1, nullptr, 0, false, &key);
}

template <typename T>
int pool_dropin_bwd(bool is_max, T *input, T *output_back, T *input_back, int64_t B, int64_t H,
                    int64_t W, int64_t D, int64_t C, int64_t kh, int64_t kw, int64_t kd,
                    int64_t Ho, int64_t Wo, int64_t Do, int64_t sh, int64_t sw, int64_t sd,
                    int64_t ph, int64_t pw, int64_t pd) {
  owl_amd_pool_desc d;
  d.B = B; d.H = H; d.W = W; d.D = D; d.C = C;
  d.kh = kh; d.kw = kw; d.kd = kd;
  d.Ho = Ho; d.Wo = Wo; d.Do = Do;
  d.sh = sh; d.sw = sw; d.sd = sd;
  d.ph = ph < 0 ? 0 : ph;  // owl_ndarray_pool_impl.h:135-136
  d.pw = pw < 0 ? 0 : pw;
  d.pd = pd < 0 ? 0 : pd;
  int rc = check_pool(&d, dt<T>());
  if (rc) return rc;
  owl_amd_pool_desc r;
  if (is_complex(dt<T>()) && (rc = complex_pool(&d, dt<T>(), is_max, &r))) return rc;
  if (!input || !output_back || !input_back) ARG_FAIL("null pointer");
  const size_t nx = (size_t)(B * H * W * D * C) * sizeof(T);
  const size_t ny = (size_t)(B * Ho * Wo * Do * C) * sizeof(T);
  // Owl's average backward is called with the output buffer as a dummy input
  // (owl_ndarray_pool.ml:85-119); it is never read for the average.
  const std::string key = dropin_key('q', dt<T>(), (int)is_max, d);
  return run_dropin({{is_max ? (const void *)input : nullptr, is_max ? nx : 16}, {output_back, ny}, {nullptr, nx}}, 2,
                    input_back, nx, [&](void **dv, hipStream_t st, DropinCtx &c) {
                      return pool_bwd_flag(&d, dt<T>(), is_max, dv[0], dv[1], dv[2], c.flag, st);
                    }, -1, nullptr, 0, false, &key);
}

template <typename T>
int pool_dropin_arg(T *input, T *output, int64_t *argmax, int64_t B, int64_t H, int64_t W,
                    int64_t C, int64_t kh, int64_t kw, int64_t Ho, int64_t Wo, int64_t sh,
                    int64_t sw, int64_t ph, int64_t pw) {
  owl_amd_pool_desc d;
  d.B = B; d.H = H; d.W = W; d.D = 1; d.C = C;
  d.kh = kh; d.kw = kw; d.kd = 1;
  d.Ho = Ho; d.Wo = Wo; d.Do = 1;
  d.sh = sh; d.sw = sw; d.sd = 1;
  d.ph = ph < 0 ? 0 : ph;  // owl_ndarray_pool_impl.h:491-492
  d.pw = pw < 0 ? 0 : pw;
  d.pd = 0;
  int rc = check_pool(&d, dt<T>());
  if (rc) return rc;
  owl_amd_pool_desc r;
  if (is_complex(dt<T>()) && (rc = complex_pool(&d, dt<T>(), 1, &r))) return rc;
  if (!input || !output || !argmax) ARG_FAIL("null pointer");
  const size_t nx = (size_t)(B * H * W * C) * sizeof(T);
  const size_t ny = (size_t)(B * Ho * Wo * C) * sizeof(T);
  const size_t na = (size_t)(B * Ho * Wo * C) * sizeof(int64_t);
  const std::string key = dropin_key('a', dt<T>(), d);
  return run_dropin({{input, nx}, {nullptr, ny}, {nullptr, na}}, 1, output, ny,
                    [&](void **dv, hipStream_t st, DropinCtx &c) {
                      return owl_amd_pool_argmax(&d, dt<T>(), dv[0], dv[1], (int64_t *)dv[2], st);
                    },
                    2, argmax, na, false, &key);
}

}  // namespace

// Owl passes (row = W, col = H) and strides as (row_stride, col_stride); the
// 2-D drop-ins map them onto the H / W axes of the descriptor.
#define DEFINE_DROPINS(N, T)                                                                      \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial(OWL_AMD_CONV2D_FWD_ARGS(T)) {                \
    return conv_dropin<T>(FWD, input, kernel, output, batches, input_cols, input_rows, 1,        \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, 1, 1, 1, padding);             \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial_backward_input(OWL_AMD_CONV2D_BWD_ARGS(T)) { \
    return conv_dropin<T>(BWD_IN, input, kernel, output, batches, input_cols, input_rows, 1,     \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, 1, 1, 1, -1);                  \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial_backward_kernel(OWL_AMD_CONV2D_BWD_ARGS(T)) { \
    return conv_dropin<T>(BWD_KER, input, kernel, output, batches, input_cols, input_rows, 1,    \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, 1, 1, 1, -1);                  \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_spatial_im2col(OWL_AMD_CONV2D_FWD_ARGS(T)) { \
    return conv_dropin<T>(FWD, input, kernel, output, batches, input_cols, input_rows, 1,        \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, col_in_stride, row_in_stride,  \
                          1, padding);                                                           \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_spatial_backward_input_im2col(              \
      OWL_AMD_CONV2D_BWD_ARGS(T)) {                                                              \
    return conv_dropin<T>(BWD_IN, input, kernel, output, batches, input_cols, input_rows, 1,     \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, col_in_stride, row_in_stride,  \
                          1, -1);                                                                \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_spatial_backward_kernel_im2col(             \
      OWL_AMD_CONV2D_BWD_ARGS(T)) {                                                              \
    return conv_dropin<T>(BWD_KER, input, kernel, output, batches, input_cols, input_rows, 1,    \
                          in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, 1,  \
                          out_channel, col_stride, row_stride, 1, col_in_stride, row_in_stride,  \
                          1, -1);                                                                \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_cuboid_im2col(OWL_AMD_CONV3D_FWD_ARGS(T)) {          \
    return conv_dropin<T>(FWD, input, kernel, output, batches, input_cols, input_rows,           \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, 1, 1, 1, padding);                             \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_cuboid_backward_input_im2col(                       \
      OWL_AMD_CONV3D_BWD_ARGS(T)) {                                                              \
    return conv_dropin<T>(BWD_IN, input, kernel, output, batches, input_cols, input_rows,        \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, 1, 1, 1, -1);                                  \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_cuboid_backward_kernel_im2col(                      \
      OWL_AMD_CONV3D_BWD_ARGS(T)) {                                                              \
    return conv_dropin<T>(BWD_KER, input, kernel, output, batches, input_cols, input_rows,       \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, 1, 1, 1, -1);                                  \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_cuboid_im2col(OWL_AMD_DCONV3D_FWD_ARGS(T)) { \
    return conv_dropin<T>(FWD, input, kernel, output, batches, input_cols, input_rows,           \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, col_in_stride, row_in_stride, dpt_in_stride,   \
                          padding);                                                              \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_cuboid_backward_input_im2col(               \
      OWL_AMD_DCONV3D_BWD_ARGS(T)) {                                                             \
    return conv_dropin<T>(BWD_IN, input, kernel, output, batches, input_cols, input_rows,        \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, col_in_stride, row_in_stride, dpt_in_stride,   \
                          -1);                                                                   \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_dilated_cuboid_backward_kernel_im2col(              \
      OWL_AMD_DCONV3D_BWD_ARGS(T)) {                                                             \
    return conv_dropin<T>(BWD_KER, input, kernel, output, batches, input_cols, input_rows,       \
                          input_dpts, in_channel, kernel_cols, kernel_rows, kernel_dpts,         \
                          output_cols, output_rows, output_dpts, out_channel, col_stride,        \
                          row_stride, dpt_stride, col_in_stride, row_in_stride, dpt_in_stride,   \
                          -1);                                                                   \
  }                                                                                               \
  /* The im2col / mec / naive spellings compute the same function as the            */         \
  /* GEBP entry points (test/unit_conv_mec_naive.ml reuses one set of vectors).        */         \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial_im2col(OWL_AMD_CONV2D_FWD_ARGS(T)) {         \
    return owl_amd_##N##_ndarray_conv_spatial(input, kernel, output, batches, input_cols,        \
        input_rows, in_channel, kernel_cols, kernel_rows, output_cols, output_rows, out_channel, \
        row_stride, col_stride, padding, row_in_stride, col_in_stride);                          \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial_mec(OWL_AMD_CONV2D_FWD_ARGS(T)) {            \
    return owl_amd_##N##_ndarray_conv_spatial(input, kernel, output, batches, input_cols,        \
        input_rows, in_channel, kernel_cols, kernel_rows, output_cols, output_rows, out_channel, \
        row_stride, col_stride, padding, row_in_stride, col_in_stride);                          \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_conv_spatial_naive(OWL_AMD_CONV2D_FWD_ARGS(T)) {          \
    return owl_amd_##N##_ndarray_conv_spatial(input, kernel, output, batches, input_cols,        \
        input_rows, in_channel, kernel_cols, kernel_rows, output_cols, output_rows, out_channel, \
        row_stride, col_stride, padding, row_in_stride, col_in_stride);                          \
  }                                                                                               \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_input, im2col)                                          \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_input, mec)                                             \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_input, naive)                                           \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_kernel, im2col)                                         \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_kernel, mec)                                            \
  DEFINE_BWD_ALIAS(N, T, spatial_backward_kernel, naive)                                          \
  DEFINE_C3_ALIAS(N, T, cuboid, im2col, mec, FWD)                                                 \
  DEFINE_C3_ALIAS(N, T, cuboid, im2col, naive, FWD)                                               \
  DEFINE_C3_ALIAS(N, T, cuboid_backward_input, im2col, mec, BWD)                                  \
  DEFINE_C3_ALIAS(N, T, cuboid_backward_input, im2col, naive, BWD)                                \
  DEFINE_C3_ALIAS(N, T, cuboid_backward_kernel, im2col, mec, BWD)                                 \
  DEFINE_C3_ALIAS(N, T, cuboid_backward_kernel, im2col, naive, BWD)                               \
  /* pooling */                                                                                   \
  extern "C" int owl_amd_##N##_ndarray_maxpool_spatial(OWL_AMD_POOL2D_FWD_ARGS(T)) {             \
    return pool_dropin_fwd<T>(true, input, output, batches, input_cols, input_rows, 1,           \
                              in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, \
                              1, col_stride, row_stride, 1, padding);                            \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_avgpool_spatial(OWL_AMD_POOL2D_FWD_ARGS(T)) {             \
    return pool_dropin_fwd<T>(false, input, output, batches, input_cols, input_rows, 1,          \
                              in_channel, kernel_cols, kernel_rows, 1, output_cols, output_rows, \
                              1, col_stride, row_stride, 1, padding);                            \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_maxpool_spatial_backward(OWL_AMD_POOL2D_BWD_ARGS(T)) {    \
    return pool_dropin_bwd<T>(true, input, output_back, input_back, batches, input_cols,         \
                              input_rows, 1, in_channel, kernel_cols, kernel_rows, 1,            \
                              output_cols, output_rows, 1, col_stride, row_stride, 1, pad_cols,  \
                              pad_rows, 0);                                                      \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_avgpool_spatial_backward(OWL_AMD_POOL2D_BWD_ARGS(T)) {    \
    return pool_dropin_bwd<T>(false, input, output_back, input_back, batches, input_cols,        \
                              input_rows, 1, in_channel, kernel_cols, kernel_rows, 1,            \
                              output_cols, output_rows, 1, col_stride, row_stride, 1, pad_cols,  \
                              pad_rows, 0);                                                      \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_maxpool_spatial_arg(OWL_AMD_POOL2D_ARG_ARGS(T)) {         \
    return pool_dropin_arg<T>(input, output, argmax, batches, input_cols, input_rows,            \
                              in_channel, kernel_cols, kernel_rows, output_cols, output_rows,    \
                              col_stride, row_stride, pad_cols, pad_rows);                       \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_maxpool_cuboid(OWL_AMD_POOL3D_FWD_ARGS(T)) {              \
    return pool_dropin_fwd<T>(true, input, output, batches, input_cols, input_rows, input_dpts,  \
                              in_channel, kernel_cols, kernel_rows, kernel_dpts, output_cols,    \
                              output_rows, output_dpts, col_stride, row_stride, dpt_stride,      \
                              padding);                                                          \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_avgpool_cuboid(OWL_AMD_POOL3D_FWD_ARGS(T)) {              \
    return pool_dropin_fwd<T>(false, input, output, batches, input_cols, input_rows, input_dpts, \
                              in_channel, kernel_cols, kernel_rows, kernel_dpts, output_cols,    \
                              output_rows, output_dpts, col_stride, row_stride, dpt_stride,      \
                              padding);                                                          \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_maxpool_cuboid_backward(OWL_AMD_POOL3D_BWD_ARGS(T)) {     \
    return pool3d_bwd_dropin<T>(true, input, output_back, input_back, batches, input_cols,       \
                                input_rows, input_dpts, in_channel, kernel_cols, kernel_rows,    \
                                kernel_dpts, output_cols, output_rows, output_dpts, col_stride,  \
                                row_stride, dpt_stride, padding);                                \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_avgpool_cuboid_backward(OWL_AMD_POOL3D_BWD_ARGS(T)) {     \
    return pool3d_bwd_dropin<T>(false, input, output_back, input_back, batches, input_cols,      \
                                input_rows, input_dpts, in_channel, kernel_cols, kernel_rows,    \
                                kernel_dpts, output_cols, output_rows, output_dpts, col_stride,  \
                                row_stride, dpt_stride, padding);                                \
  }

#define DEFINE_BWD_ALIAS(N, T, OP, SUF)                                                           \
  extern "C" int owl_amd_##N##_ndarray_conv_##OP##_##SUF(OWL_AMD_CONV2D_BWD_ARGS(T)) {           \
    return owl_amd_##N##_ndarray_conv_##OP(input, kernel, output, batches, input_cols,           \
        input_rows, in_channel, kernel_cols, kernel_rows, output_cols, output_rows, out_channel, \
        row_stride, col_stride, row_in_stride, col_in_stride);                                   \
  }

#define C3_CALL_FWD                                                                               \
  (input, kernel, output, batches, input_cols, input_rows, input_dpts, in_channel, kernel_cols,  \
   kernel_rows, kernel_dpts, output_cols, output_rows, output_dpts, out_channel, dpt_stride,     \
   row_stride, col_stride, padding)
#define C3_CALL_BWD                                                                               \
  (input, kernel, output, batches, input_cols, input_rows, input_dpts, in_channel, kernel_cols,  \
   kernel_rows, kernel_dpts, output_cols, output_rows, output_dpts, out_channel, dpt_stride,     \
   row_stride, col_stride)
#define C3_ARGS_FWD(T) OWL_AMD_CONV3D_FWD_ARGS(T)
#define C3_ARGS_BWD(T) OWL_AMD_CONV3D_BWD_ARGS(T)
#define DEFINE_C3_ALIAS(N, T, OP, BASE, SUF, KIND)                                                \
  extern "C" int owl_amd_##N##_ndarray_conv_##OP##_##SUF(C3_ARGS_##KIND(T)) {                    \
    return owl_amd_##N##_ndarray_conv_##OP##_##BASE C3_CALL_##KIND;                              \
  }

namespace {
// 3-D pooling backward derives its pad from `padding` like the forward
// (owl_ndarray_pool_impl.h:363-373).
template <typename T>
int pool3d_bwd_dropin(bool is_max, T *input, T *output_back, T *input_back, int64_t B, int64_t H,
                      int64_t W, int64_t D, int64_t C, int64_t kh, int64_t kw, int64_t kd,
                      int64_t Ho, int64_t Wo, int64_t Do, int64_t sh, int64_t sw, int64_t sd,
                      int64_t padding) {
  int64_t ph = 0, pw = 0, pd = 0;
  if (padding != 1) {
    ph = owl_pad(sh, Ho, kh, H);
    pw = owl_pad(sw, Wo, kw, W);
    pd = owl_pad(sd, Do, kd, D);
  }
  return pool_dropin_bwd<T>(is_max, input, output_back, input_back, B, H, W, D, C, kh, kw, kd, Ho,
                            Wo, Do, sh, sw, sd, ph, pw, pd);
}
}  // namespace

DEFINE_DROPINS(float32, float)
DEFINE_DROPINS(float64, double)
DEFINE_DROPINS(complex32, owl_amd_complex32)
DEFINE_DROPINS(complex64, owl_amd_complex64)

// Upsampling drop-ins, Owl's stub argument order (owl_ndarray_upsampling_impl.h:9-14):
// cols = H (scaled by col_scale), rows = W (scaled by row_scale).
namespace {
template <typename T>
int upsample_dropin(bool backward, T *input, T *output, int64_t B, int64_t H, int64_t W, int64_t C, int64_t Ho,
                    int64_t Wo, int64_t sh, int64_t sw) {
  if (B < 0 || H < 0 || W < 0 || C < 0 || Ho < 0 || Wo < 0 || sh < 1 || sw < 1) ARG_FAIL("bad upsampling sizes");
  const size_t nx = (size_t)(B * H * W * C) * sizeof(T), ny = (size_t)(B * Ho * Wo * C) * sizeof(T);
  if ((backward ? nx : ny) == 0) return OWL_AMD_OK;
  if (!input || !output) ARG_FAIL("null pointer");
  const std::string key = dropin_key(backward ? 'U' : 'u', dt<T>(), B, H, W, C, Ho, Wo, sh, sw);
  if (backward)
    return run_dropin({{output, ny}, {nullptr, nx}}, 1, input, nx, [&](void **dv, hipStream_t st, DropinCtx &c) {
      return owl_amd_upsampling2d_bwd(dt<T>(), B, H, W, C, Ho, Wo, sh, sw, dv[0], dv[1], st);
    }, -1, nullptr, 0, false, &key);
  return run_dropin({{input, nx}, {nullptr, ny}}, 1, output, ny, [&](void **dv, hipStream_t st, DropinCtx &c) {
    return owl_amd_upsampling2d_fwd(dt<T>(), B, H, W, C, Ho, Wo, sh, sw, dv[0], dv[1], st);
  }, -1, nullptr, 0, false, &key);
}
}  // namespace

#define DEFINE_UPSAMPLE_DROPINS(N, T)                                                             \
  extern "C" int owl_amd_##N##_ndarray_upsampling_spatial_backward(OWL_AMD_UPSAMPLE_ARGS(T)) {    \
    return upsample_dropin<T>(true, input, output, batches, input_cols, input_rows, in_channel,   \
                              output_cols, output_rows, col_scale, row_scale);                    \
  }                                                                                               \
  extern "C" int owl_amd_##N##_ndarray_upsampling_spatial(OWL_AMD_UPSAMPLE_ARGS(T)) {             \
    return upsample_dropin<T>(false, input, output, batches, input_cols, input_rows, in_channel,  \
                              output_cols, output_rows, col_scale, row_scale);                    \
  }
DEFINE_UPSAMPLE_DROPINS(float32, float)
DEFINE_UPSAMPLE_DROPINS(float64, double)
DEFINE_UPSAMPLE_DROPINS(complex32, owl_amd_complex32)
DEFINE_UPSAMPLE_DROPINS(complex64, owl_amd_complex64)

// ------------------------------------------------------------------ ragged groups
// A plan for running many forward convolutions (different shapes) as one
// persistent launch.  Each problem's plan -- tile edge, split-K factor, or a
// launch of its own -- comes from the shape-class cache the autotuner fills
// (or a size heuristic until it has run).
namespace {

// in_t, out_t, aligned, log2 M, log2 N, log2 K, unit strides, taps
typedef std::tuple<int, int, int, int, int, int, int, int> ClassKey;
// plan kinds: the tile edge in the persistent launch (0 full, 1 half), or a
// launch of the single-problem kernels of its own (owl_amd_conv_fwd picks the
// halo, direct-A or generic kernel) after the persistent one
constexpr int GROUP_OWN_LAUNCH = 2;
constexpr int GROUP_MAX_SPLIT = 16;
// split-K candidates the autotuner tries keep at least this many K-steps per
// split, and at most this many bytes of FP32 partials per shape class
constexpr int64_t kGroupMinStepsPerSplit = 16;
constexpr size_t kGroupSplitCap = (size_t)256 << 20;

struct GroupPlan {
  int kind;   // 0 / 1 / GROUP_OWN_LAUNCH
  int split;  // requested K splits in the persistent launch (1: none)
};
std::map<ClassKey, GroupPlan> g_tune_cache;

int ilog2(int64_t v) {
  int r = 0;
  while (v > 1) {
    v >>= 1;
    ++r;
  }
  return r;
}

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

}  // namespace

struct owl_amd_conv_group {
  int n = 0, in_t = 0, out_t = 0;
  std::vector<owl_amd_conv_desc> descs;
  std::vector<GroupEntry> tab;
  std::vector<ClassKey> keys;
  std::vector<size_t> ws_off;     // byte offset of each split problem's partials in ws
  std::vector<void *> ys;         // outputs of the last run / capture / autotune
  GroupEntry *dev_tab = nullptr;
  void *ws = nullptr;             // FP32 split-K partials of the plan
  size_t ws_bytes = 0;
  int64_t total = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int group_edge(int in_t, int cfg) {
  if (in_t == OWL_AMD_F32) return group_tile_edge<float>(cfg);
  if (in_t == OWL_AMD_F16) return group_tile_edge<f16_t>(cfg);
  return group_tile_edge<bf16_t>(cfg);
}

static int group_kstep(int in_t) {
  if (in_t == OWL_AMD_F32) return group_k_step<float>();
  if (in_t == OWL_AMD_F16) return group_k_step<f16_t>();
  return group_k_step<bf16_t>();
}

static int64_t group_tiles_mn(int in_t, const GroupEntry &e) {
  const int edge = group_edge(in_t, e.cfg);
  return ((e.p.M + edge - 1) / edge) * (int64_t)((e.p.N + edge - 1) / edge);
}

// Set a problem's plan: tile edge, and the K splits as whole K-steps (the
// effective count can come out below the requested one; 1 when the problem
// is empty or has its own launch).
static void group_apply(int in_t, GroupEntry &e, GroupPlan plan) {
  e.cfg = plan.kind;
  e.p.split = 1;
  e.p.k_per_split = 0;
  if (plan.kind == GROUP_OWN_LAUNCH || plan.split <= 1 || e.p.M == 0 || e.p.N == 0 || e.p.K == 0) return;
  const int64_t bk = group_kstep(in_t);
  const int64_t steps = (e.p.K + bk - 1) / bk;
  const int64_t kps = ((steps + plan.split - 1) / plan.split) * bk;
  const int64_t s = (e.p.K + kps - 1) / kps;
  if (s < 2) return;
  e.p.split = (int)s;
  e.p.k_per_split = kps;
}

static size_t group_partial_bytes(const GroupEntry &e) {
  return e.p.split > 1 ? (size_t)e.p.split * (size_t)e.p.M * (size_t)e.p.N * sizeof(float) : 0;
}

static void group_drop_graph(owl_amd_conv_group *g) {
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  g->exec = nullptr;
  g->graph = nullptr;
}

// Tile ranges of the persistent launch (problems with their own launch own
// no tiles: the kernel's search over tile_begin skips them) and the
// partials workspace of the split problems, grown when the plan needs more.
static int group_layout(owl_amd_conv_group *g) {
  int64_t t = 0;
  size_t off = 0;
  for (int i = 0; i < g->n; ++i) {
    GroupEntry &e = g->tab[i];
    e.tile_begin = t;
    e.tiles_mn = 0;
    g->ws_off[i] = off;
    if (e.cfg == GROUP_OWN_LAUNCH) continue;
    e.tiles_mn = group_tiles_mn(g->in_t, e);
    t += e.tiles_mn * e.p.split;
    off += align256(group_partial_bytes(e));
  }
  g->total = t;
  if (off > g->ws_bytes) {
    if (g->ws) {
      (void)hipDeviceSynchronize();
      (void)hipFree(g->ws);
      g->ws = nullptr;
      g->ws_bytes = 0;
    }
    hipError_t err = hipMalloc(&g->ws, off);
    if (err != hipSuccess) {
      owl_amd_set_error("hipMalloc group split-K partials (%zu bytes): %s", off, hipGetErrorString(err));
      return OWL_AMD_ERR_NOMEM;
    }
    g->ws_bytes = off;
  }
  return OWL_AMD_OK;
}

static int group_launch(owl_amd_conv_group *g, const GroupEntry *tab, int nprob, int64_t total, hipStream_t st) {
  if (total == 0) return OWL_AMD_OK;
  int grid = (int)(total < 2048 ? total : 2048);  // persistent: ~8 workgroups per CU
  if (g->in_t == OWL_AMD_BF16 && g->out_t == OWL_AMD_BF16) run_group_fwd<bf16_t, bf16_t>(tab, nprob, total, grid, st);
  else if (g->in_t == OWL_AMD_BF16 && g->out_t == OWL_AMD_F32) run_group_fwd<bf16_t, float>(tab, nprob, total, grid, st);
  else if (g->in_t == OWL_AMD_F16 && g->out_t == OWL_AMD_F16) run_group_fwd<f16_t, f16_t>(tab, nprob, total, grid, st);
  else run_group_fwd<float, float>(tab, nprob, total, grid, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

// Sum a split problem's partials (at p.c) into its output, in split order.
static int group_reduce(owl_amd_conv_group *g, const ConvParams &p, void *y, hipStream_t st) {
  const float *part = (const float *)p.c;
  const int64_t n = p.M * (int64_t)p.N;
  if (g->out_t == OWL_AMD_F32) run_splitk_reduce<float, float>(part, (float *)y, n, p.split, st);
  else if (g->out_t == OWL_AMD_BF16) run_splitk_reduce<float, bf16_t>(part, (bf16_t *)y, n, p.split, st);
  else run_splitk_reduce<float, f16_t>(part, (f16_t *)y, n, p.split, st);
  OWL_CHECK_HIP(hipGetLastError());
  return OWL_AMD_OK;
}

static bool in_set(const std::vector<int> *idx, int i) {
  return !idx || std::find(idx->begin(), idx->end(), i) != idx->end();
}

// After a persistent launch: the split-K reductions of its split problems,
// then the problems with their own launch, each as a single-problem forward
// on st -- those of idx, or all when idx is null.
static int group_finish(owl_amd_conv_group *g, const GroupEntry *tab, const std::vector<int> *idx, hipStream_t st) {
  for (int i = 0; i < g->n; ++i) {
    if (!in_set(idx, i) || tab[i].cfg == GROUP_OWN_LAUNCH || tab[i].p.split <= 1) continue;
    int rc = group_reduce(g, tab[i].p, g->ys[i], st);
    if (rc) return rc;
  }
  for (int i = 0; i < g->n; ++i) {
    if (!in_set(idx, i) || tab[i].cfg != GROUP_OWN_LAUNCH) continue;
    const ConvParams &p = tab[i].p;
    int rc = owl_amd_conv_fwd(&g->descs[i], g->in_t, g->out_t, p.a, p.b, g->ys[i], st);
    if (rc) return rc;
  }
  return OWL_AMD_OK;
}

// One run: the persistent launch, the reductions, the own launches.
static int group_dispatch(owl_amd_conv_group *g, hipStream_t st) {
  int rc = group_launch(g, g->dev_tab, g->n, g->total, st);
  return rc ? rc : group_finish(g, g->tab.data(), nullptr, st);
}

static void group_set_ptrs(owl_amd_conv_group *g, const void *const *xs, const void *const *ws, void *const *ys) {
  for (int i = 0; i < g->n; ++i) {
    GroupEntry &e = g->tab[i];
    e.p.a = xs[i];
    e.p.b = ws[i];
    g->ys[i] = ys[i];
    e.p.c = e.p.split > 1 ? (void *)((char *)g->ws + g->ws_off[i]) : ys[i];
    const int v = vec_of(g->in_t);
    e.aligned = e.p.Ci % v == 0 && e.p.Co % v == 0 && aligned16(xs[i], ws[i], nullptr);
  }
}

static int group_check_ptrs(const owl_amd_conv_group *g, const void *const *xs, const void *const *ws,
                            void *const *ys) {
  if (g->n > 0 && (!xs || !ws || !ys)) ARG_FAIL("bad group arguments");
  for (int i = 0; i < g->n; ++i)
    if (!xs[i] || !ws[i] || !ys[i]) ARG_FAIL("null pointer for problem %d", i);
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_group_create(owl_amd_conv_group **out, int n, const owl_amd_conv_desc *descs,
                                         int in_t, int out_t) {
  if (!out || n < 0 || (n > 0 && !descs)) ARG_FAIL("bad group arguments");
  const bool ok_t = (in_t == OWL_AMD_BF16 && (out_t == OWL_AMD_BF16 || out_t == OWL_AMD_F32)) ||
                    (in_t == OWL_AMD_F16 && out_t == OWL_AMD_F16) || (in_t == OWL_AMD_F32 && out_t == OWL_AMD_F32);
  if (!ok_t) {
    owl_amd_set_error("group: unsupported dtype pair in=%d out=%d", in_t, out_t);
    return OWL_AMD_ERR_UNSUPPORTED;
  }
  for (int i = 0; i < n; ++i) {
    int rc = check_conv(&descs[i]);
    if (rc) return rc;
  }
  owl_amd_conv_group *g = new owl_amd_conv_group();
  g->n = n;
  g->in_t = in_t;
  g->out_t = out_t;
  g->descs.assign(descs, descs + n);
  g->tab.resize(n);
  g->keys.resize(n);
  g->ws_off.assign(n, 0);
  g->ys.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    GroupEntry &e = g->tab[i];
    memset(&e, 0, sizeof(e));
    e.p = make_params(&descs[i]);
    e.p.M = descs[i].B * descs[i].Ho * descs[i].Wo * descs[i].Do;
    e.p.N = e.p.Co;
    e.p.K = descs[i].kh * descs[i].kw * descs[i].kd * descs[i].Ci;
    e.p.ldc = e.p.Co;
    const int v = vec_of(in_t);
    const int al = e.p.Ci % v == 0 && e.p.Co % v == 0;
    g->keys[i] = ClassKey(in_t, out_t, al, ilog2(e.p.M), ilog2(e.p.N), ilog2(e.p.K),
                          e.p.sh == 1 && e.p.sw == 1 && e.p.sd == 1, e.p.kh * e.p.kw * e.p.kd);
    GroupPlan plan;
    {
      std::lock_guard<std::recursive_mutex> lk(g_mu);
      auto it = g_tune_cache.find(g->keys[i]);
      if (it != g_tune_cache.end()) {
        plan = it->second;
      } else {
        // heuristic until tuned: half-size tiles when the full tile would
        // leave most of the chip idle or mostly empty; no splits
        const int full = group_edge(in_t, 0);
        const int64_t full_tiles = ((e.p.M + full - 1) / full) * (int64_t)((e.p.N + full - 1) / full);
        plan = GroupPlan{full_tiles < 256 || e.p.N <= full / 2 ? 1 : 0, 1};
      }
    }
    group_apply(in_t, e, plan);
  }
  int rc = group_layout(g);
  if (rc) {
    owl_amd_conv_group_destroy(g);
    return rc;
  }
  hipError_t err = hipMalloc(&g->dev_tab, sizeof(GroupEntry) * (n > 0 ? n : 1));
  if (err != hipSuccess) {
    owl_amd_conv_group_destroy(g);
    owl_amd_set_error("hipMalloc group table: %s", hipGetErrorString(err));
    return OWL_AMD_ERR_NOMEM;
  }
  *out = g;
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_group_info(const owl_amd_conv_group *g, int *tile_edges, int64_t *total_tiles) {
  if (!g) ARG_FAIL("null group");
  if (tile_edges)
    for (int i = 0; i < g->n; ++i)
      tile_edges[i] = g->tab[i].cfg == GROUP_OWN_LAUNCH ? 0 : group_edge(g->in_t, g->tab[i].cfg);
  if (total_tiles) *total_tiles = g->total;
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_group_plan(const owl_amd_conv_group *g, int *kinds, int *splits) {
  if (!g) ARG_FAIL("null group");
  for (int i = 0; i < g->n; ++i) {
    if (kinds) kinds[i] = g->tab[i].cfg;
    if (splits) splits[i] = g->tab[i].p.split;
  }
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_group_set_plan(owl_amd_conv_group *g, const int *kinds, const int *splits) {
  if (!g || (g->n > 0 && (!kinds || !splits))) ARG_FAIL("bad group plan arguments");
  for (int i = 0; i < g->n; ++i) {
    if (kinds[i] < 0 || kinds[i] > GROUP_OWN_LAUNCH) ARG_FAIL("problem %d: plan kind %d not 0, 1 or 2", i, kinds[i]);
    if (splits[i] < 1 || splits[i] > GROUP_MAX_SPLIT)
      ARG_FAIL("problem %d: split %d outside 1..%d", i, splits[i], GROUP_MAX_SPLIT);
    if (kinds[i] == GROUP_OWN_LAUNCH && splits[i] != 1) ARG_FAIL("problem %d: an own launch takes split 1", i);
  }
  for (int i = 0; i < g->n; ++i) group_apply(g->in_t, g->tab[i], GroupPlan{kinds[i], splits[i]});
  group_drop_graph(g);
  return group_layout(g);
}

extern "C" int owl_amd_conv_group_run(owl_amd_conv_group *g, const void *const *xs, const void *const *ws,
                                      void *const *ys, void *stream) {
  if (!g) ARG_FAIL("bad group run arguments");
  int rc = group_check_ptrs(g, xs, ws, ys);
  if (rc) return rc;
  group_set_ptrs(g, xs, ws, ys);
  hipStream_t st = (hipStream_t)stream;
  // the table copy is ordered before the launch on the same stream; the host
  // vector is not touched again until the next run, which first waits here
  OWL_CHECK_HIP(hipMemcpyAsync(g->dev_tab, g->tab.data(), sizeof(GroupEntry) * g->n, hipMemcpyHostToDevice, st));
  rc = group_dispatch(g, st);
  if (rc) return rc;
  OWL_CHECK_HIP(hipStreamSynchronize(st));
  return OWL_AMD_OK;
}

// Time each shape class under every candidate plan -- a grouped launch of
// just that class's problems with either tile edge, unsplit or split-K in
// two to sixteen splits (when every problem of the class keeps >= 16 K-steps
// per split and its partials stay within 256 MiB), and the class's problems
// as single-problem launches -- and keep the fastest in the process-wide
// cache.  A captured graph is dropped (capture again after tuning).
extern "C" int owl_amd_conv_group_autotune(owl_amd_conv_group *g, const void *const *xs, const void *const *ws,
                                           void *const *ys, void *stream) {
  if (!g) ARG_FAIL("null group");
  int rc = group_check_ptrs(g, xs, ws, ys);
  if (rc) return rc;
  group_drop_graph(g);
  group_set_ptrs(g, xs, ws, ys);
  hipStream_t st = (hipStream_t)stream;
  std::map<ClassKey, std::vector<int>> classes;
  for (int i = 0; i < g->n; ++i) classes[g->keys[i]].push_back(i);
  const int64_t bk = group_kstep(g->in_t);
  // candidate plans per class, and the partials the largest one needs
  std::map<ClassKey, std::vector<GroupPlan>> cands;
  size_t tmp_ws_bytes = 0;
  for (auto &kv : classes) {
    std::vector<GroupPlan> &c = cands[kv.first];
    c.push_back(GroupPlan{0, 1});
    c.push_back(GroupPlan{1, 1});
    int64_t min_steps = INT64_MAX;
    size_t mn = 0;
    for (int i : kv.second) {
      min_steps = std::min(min_steps, (g->tab[i].p.K + bk - 1) / bk);
      mn += (size_t)g->tab[i].p.M * (size_t)g->tab[i].p.N;
    }
    for (int s = 2; s <= GROUP_MAX_SPLIT; s *= 2) {
      const size_t bytes = (size_t)s * mn * sizeof(float) + kv.second.size() * 256;
      if (min_steps / s < kGroupMinStepsPerSplit || bytes > kGroupSplitCap) break;
      c.push_back(GroupPlan{0, s});
      c.push_back(GroupPlan{1, s});
      tmp_ws_bytes = std::max(tmp_ws_bytes, bytes);
    }
    c.push_back(GroupPlan{GROUP_OWN_LAUNCH, 1});
  }
  GroupEntry *tmp = nullptr;
  void *tmp_ws = nullptr;
  OWL_CHECK_HIP(hipMalloc(&tmp, sizeof(GroupEntry) * (g->n > 0 ? g->n : 1)));
  if (tmp_ws_bytes && hipMalloc(&tmp_ws, tmp_ws_bytes) != hipSuccess) {
    (void)hipFree(tmp);
    owl_amd_set_error("hipMalloc autotune split-K partials (%zu bytes)", tmp_ws_bytes);
    return OWL_AMD_ERR_NOMEM;
  }
  hipEvent_t e0, e1;
  OWL_CHECK_HIP(hipEventCreate(&e0));
  OWL_CHECK_HIP(hipEventCreate(&e1));
  std::vector<GroupEntry> sub_tab(g->tab);  // the class's entries at their problem index
  for (auto &kv : classes) {
    float best = 1e30f;
    GroupPlan best_plan{0, 1};
    for (const GroupPlan &plan : cands[kv.first]) {
      // a table of just this class's problems under the candidate plan
      std::vector<GroupEntry> sub;
      int64_t t = 0;
      size_t off = 0;
      for (int i : kv.second) {
        GroupEntry e = g->tab[i];
        group_apply(g->in_t, e, plan);
        e.tile_begin = t;
        e.tiles_mn = plan.kind == GROUP_OWN_LAUNCH ? 0 : group_tiles_mn(g->in_t, e);
        t += e.tiles_mn * e.p.split;
        e.p.c = g->ys[i];
        if (e.p.split > 1) {
          e.p.c = (char *)tmp_ws + off;
          off += align256(group_partial_bytes(e));
        }
        sub.push_back(e);
        sub_tab[i] = e;
      }
      if (t > 0 && hipMemcpy(tmp, sub.data(), sizeof(GroupEntry) * sub.size(), hipMemcpyHostToDevice) != hipSuccess) {
        rc = OWL_AMD_ERR_HIP;
        break;
      }
      auto once = [&]() {
        int r = group_launch(g, tmp, (int)sub.size(), t, st);
        return r ? r : group_finish(g, sub_tab.data(), &kv.second, st);
      };
      rc = once();  // warm-up
      if (rc) break;
      (void)hipEventRecord(e0, st);
      for (int r = 0; r < 3 && rc == OWL_AMD_OK; ++r) rc = once();
      (void)hipEventRecord(e1, st);
      (void)hipEventSynchronize(e1);
      if (rc) break;
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) {
        best = ms;
        best_plan = plan;
      }
    }
    if (rc) break;
    {
      std::lock_guard<std::recursive_mutex> lk(g_mu);
      g_tune_cache[kv.first] = best_plan;
    }
    for (int i : kv.second) group_apply(g->in_t, g->tab[i], best_plan);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamSynchronize(st);
  (void)hipFree(tmp);
  if (tmp_ws) (void)hipFree(tmp_ws);
  int lrc = group_layout(g);
  return rc ? rc : lrc;
}

// Capture one run into a hipGraph (the problem table is uploaded before
// capture, so the graph holds the grouped kernel node, then the split-K
// reductions and the kernel nodes of the problems with their own launch).
extern "C" int owl_amd_conv_group_capture(owl_amd_conv_group *g, const void *const *xs, const void *const *ws,
                                          void *const *ys, void *stream) {
  if (!g || !stream) ARG_FAIL("group capture needs a group and a non-default stream");
  int rc = group_check_ptrs(g, xs, ws, ys);
  if (rc) return rc;
  group_set_ptrs(g, xs, ws, ys);
  OWL_CHECK_HIP(hipMemcpy(g->dev_tab, g->tab.data(), sizeof(GroupEntry) * g->n, hipMemcpyHostToDevice));
  hipStream_t st = (hipStream_t)stream;
  group_drop_graph(g);
  OWL_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  rc = group_dispatch(g, st);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(st, &graph);
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) {
    owl_amd_set_error("hipStreamEndCapture: %s", hipGetErrorString(e));
    return OWL_AMD_ERR_HIP;
  }
  g->graph = graph;
  OWL_CHECK_HIP(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
  return OWL_AMD_OK;
}

extern "C" int owl_amd_conv_group_replay(owl_amd_conv_group *g, void *stream) {
  if (!g || !g->exec) ARG_FAIL("group has no captured graph");
  OWL_CHECK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
  return OWL_AMD_OK;
}

extern "C" void owl_amd_conv_group_destroy(owl_amd_conv_group *g) {
  if (!g) return;
  group_drop_graph(g);
  if (g->dev_tab || g->ws) (void)hipDeviceSynchronize();
  if (g->dev_tab) (void)hipFree(g->dev_tab);
  if (g->ws) (void)hipFree(g->ws);
  delete g;
}

/*
 * owl_oracle.c -- CPU restatement of Owl's convolution / pooling hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it.  The
 * product path (owl_amd/) never links, loads or calls it.
 *
 * Conventions (Owl's, owl_ndarray_conv_impl.h:246-253 and :1211-1220):
 *   input  x  [B][H][W][D][Ci]      (Owl names H "cols", W "rows", D "dpts")
 *   kernel w  [kh][kw][kd][Ci][Co]  (HWIO / DHWIO, row-major)
 *   output y  [B][Ho][Wo][Do][Co]
 * A 2-D problem is the 3-D one with D = Do = kd = 1, which is exactly how
 * the 2-D and 3-D templates index memory (the depth term vanishes).
 * Strides s*, dilation rates r* (Owl "*_in_stride") and leading pads p* are
 * per axis; the padding is computed by the caller from Owl's formula
 * (owl_ndarray_conv_impl.h:288-294, 502-505, 3141-3150):
 *     p = max(0, (s*(o-1) + k_eff - n) / 2)  with C truncating division,
 *     k_eff = k + (k-1)*(r-1).
 *
 * Parity pins (tests/): Owl's golden vectors (test/unit_*conv*.ml,
 * test/unit_pool*.ml), brute-force adjoint identities, 1x1 conv == GEMM,
 * FP64 central finite differences.
 */
#include <cpuid.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
  int64_t B, H, W, D, Ci;
  int64_t kh, kw, kd, Co;
  int64_t Ho, Wo, Do;
  int64_t sh, sw, sd;
  int64_t rh, rw, rd;
  int64_t ph, pw, pd;
} oracle_conv_geom;

typedef struct {
  int64_t B, H, W, D, C;
  int64_t kh, kw, kd;
  int64_t Ho, Wo, Do;
  int64_t sh, sw, sd;
  int64_t ph, pw, pd;
} oracle_pool_geom;

#define IN_IDX(g, b, h, w, d, c) \
  ((((((b) * (g)->H + (h)) * (g)->W + (w)) * (g)->D + (d)) * (g)->Ci) + (c))
#define OUT_IDX(g, b, h, w, d, o) \
  ((((((b) * (g)->Ho + (h)) * (g)->Wo + (w)) * (g)->Do + (d)) * (g)->Co) + (o))
#define K_IDX(g, a, c, e, q, o) \
  ((((((a) * (g)->kw + (c)) * (g)->kd + (e)) * (g)->Ci + (q)) * (g)->Co) + (o))

/* ------------------------------------------------------------------------
 * FP64 direct loops (the oracle of record).
 * Forward: owl_ndarray_conv_impl.h:2440-2535 (spatial_naive),
 * :2735-2850 (cuboid_naive), dilation as in :3097-3190.  Sums are carried in
 * long double so the oracle sits well below FP64 rounding of the result.
 * ---------------------------------------------------------------------- */
void oracle_conv_fwd_f64(const oracle_conv_geom *g, const double *x,
                         const double *w, double *y) {
  for (int64_t b = 0; b < g->B; ++b)
    for (int64_t i = 0; i < g->Ho; ++i)
      for (int64_t j = 0; j < g->Wo; ++j)
        for (int64_t l = 0; l < g->Do; ++l)
          for (int64_t o = 0; o < g->Co; ++o) {
            long double s = 0.0L;
            for (int64_t a = 0; a < g->kh; ++a) {
              int64_t h = i * g->sh - g->ph + a * g->rh;
              if (h < 0 || h >= g->H) continue;
              for (int64_t c = 0; c < g->kw; ++c) {
                int64_t ww = j * g->sw - g->pw + c * g->rw;
                if (ww < 0 || ww >= g->W) continue;
                for (int64_t e = 0; e < g->kd; ++e) {
                  int64_t d = l * g->sd - g->pd + e * g->rd;
                  if (d < 0 || d >= g->D) continue;
                  const double *xp = x + IN_IDX(g, b, h, ww, d, 0);
                  const double *wp = w + K_IDX(g, a, c, e, 0, o);
                  for (int64_t q = 0; q < g->Ci; ++q)
                    s += (long double)xp[q] * (long double)wp[q * g->Co];
                }
              }
            }
            y[OUT_IDX(g, b, i, j, l, o)] = (double)s;
          }
}

/* Backward input: dx[b,h,w,d,q] = sum over (i,j,l,a,c,e,o) with
 * h = i*sh - ph + a*rh (etc.) of dy[b,i,j,l,o] * w[a,c,e,q,o]
 * (owl_ndarray_conv_impl.h:2642-2725, :2977-3082).  Restated as a gather
 * over output positions so every dx element is one long-double sum. */
void oracle_conv_bwd_input_f64(const oracle_conv_geom *g, const double *w,
                               const double *dy, double *dx) {
  for (int64_t b = 0; b < g->B; ++b)
    for (int64_t h = 0; h < g->H; ++h)
      for (int64_t ww = 0; ww < g->W; ++ww)
        for (int64_t d = 0; d < g->D; ++d)
          for (int64_t q = 0; q < g->Ci; ++q) {
            long double s = 0.0L;
            for (int64_t a = 0; a < g->kh; ++a) {
              int64_t th = h + g->ph - a * g->rh;
              if (th < 0 || th % g->sh) continue;
              int64_t i = th / g->sh;
              if (i >= g->Ho) continue;
              for (int64_t c = 0; c < g->kw; ++c) {
                int64_t tw = ww + g->pw - c * g->rw;
                if (tw < 0 || tw % g->sw) continue;
                int64_t j = tw / g->sw;
                if (j >= g->Wo) continue;
                for (int64_t e = 0; e < g->kd; ++e) {
                  int64_t td = d + g->pd - e * g->rd;
                  if (td < 0 || td % g->sd) continue;
                  int64_t l = td / g->sd;
                  if (l >= g->Do) continue;
                  const double *gp = dy + OUT_IDX(g, b, i, j, l, 0);
                  const double *wp = w + K_IDX(g, a, c, e, q, 0);
                  for (int64_t o = 0; o < g->Co; ++o)
                    s += (long double)gp[o] * (long double)wp[o];
                }
              }
            }
            dx[IN_IDX(g, b, h, ww, d, q)] = (double)s;
          }
}

/* Backward kernel: dw[a,c,e,q,o] = sum_{b,i,j,l} x[b,h(i,a),w(j,c),d(l,e),q]
 * * dy[b,i,j,l,o]  (owl_ndarray_conv_impl.h:2546-2632, :2860-2967). */
void oracle_conv_bwd_kernel_f64(const oracle_conv_geom *g, const double *x,
                                const double *dy, double *dw) {
  int64_t nk = g->kh * g->kw * g->kd * g->Ci * g->Co;
  long double *acc = (long double *)calloc((size_t)nk, sizeof(long double));
  for (int64_t b = 0; b < g->B; ++b)
    for (int64_t i = 0; i < g->Ho; ++i)
      for (int64_t j = 0; j < g->Wo; ++j)
        for (int64_t l = 0; l < g->Do; ++l) {
          const double *gp = dy + OUT_IDX(g, b, i, j, l, 0);
          for (int64_t a = 0; a < g->kh; ++a) {
            int64_t h = i * g->sh - g->ph + a * g->rh;
            if (h < 0 || h >= g->H) continue;
            for (int64_t c = 0; c < g->kw; ++c) {
              int64_t ww = j * g->sw - g->pw + c * g->rw;
              if (ww < 0 || ww >= g->W) continue;
              for (int64_t e = 0; e < g->kd; ++e) {
                int64_t d = l * g->sd - g->pd + e * g->rd;
                if (d < 0 || d >= g->D) continue;
                const double *xp = x + IN_IDX(g, b, h, ww, d, 0);
                long double *ap = acc + K_IDX(g, a, c, e, 0, 0);
                for (int64_t q = 0; q < g->Ci; ++q) {
                  long double xv = xp[q];
                  for (int64_t o = 0; o < g->Co; ++o)
                    ap[q * g->Co + o] += xv * (long double)gp[o];
                }
              }
            }
          }
        }
  for (int64_t t = 0; t < nk; ++t) dw[t] = (double)acc[t];
  free(acc);
}

/* ------------------------------------------------------------------------
 * FP32 restatement of Owl's CPU algorithm, used for the CPU baseline timing
 * and the FP32 comparison.  Owl's dispatch: a 2-D convolution whose im2col
 * matrix has fewer than 512 Ki elements (owl_ndarray_conv_impl.h:14,
 * :297-298) is im2col + one SGEMM (:302-333 forward; :515-546 backward
 * input: GEMM then a serial col2im +=; :728-766 backward kernel: GEMM into
 * [Co][K] then a transpose into HWIO); a larger one runs the cache-blocked
 * (GEBP) loops of :340-452, :553-663 and :774-887 with the block sizes of
 * compute_block_sizes (:21-87) from the host's CPUID cache sizes
 * (owl_core_utils.c:211-300).  3-D and dilated convolutions always take the
 * im2col path (:1211-1554, :3097-3784).  The SGEMM is the host's OpenBLAS,
 * passed in as a function pointer with the CBLAS signature.  The GEBP
 * backward passes' cumulative offsets (idx_kn_base / idx_mn_base += n,
 * :589-590, :854-855, wrong once Co is split into three or more blocks) are
 * not reproduced: the offsets here are k * Co + n and m * Co + n.
 * ---------------------------------------------------------------------- */
typedef void (*cblas_sgemm_fn)(int order, int ta, int tb, int m, int n, int k,
                               float alpha, const float *a, int lda,
                               const float *b, int ldb, float beta, float *c,
                               int ldc);
enum { CblasRowMajor_ = 101, CblasNoTrans_ = 111, CblasTrans_ = 112 };
#define OWL_IM2COL_THRESHOLD (512 * 1024)

/* Host cache sizes as Owl reads them: AMD from CPUID 0x80000005 / 0x80000006,
 * otherwise Intel's deterministic cache leaf 4 (32 / 256 / 2048 KiB when a
 * level is missing or the leaf is absent). */
void oracle_host_cache_sizes(int *l1p, int *l2p, int *l3p) {
  unsigned a, b, c, d;
  int l1 = 32 * 1024, l2 = 256 * 1024, l3 = 2048 * 1024;
  if (__get_cpuid(0, &a, &b, &c, &d)) {
    const int amd = (b == 0x68747541 && d == 0x69746e65 && c == 0x444d4163) ||
                    (b == 0x69444d41 && d == 0x74656273 && c == 0x21726574);
    if (amd) {
      __cpuid(0x80000005, a, b, c, d);
      l1 = (int)(c >> 24) * 1024;
      __cpuid(0x80000006, a, b, c, d);
      l2 = (int)(c >> 16) * 1024;
      l3 = (int)((d & 0xFFFC000) >> 18) * 512 * 1024;
    } else if (a >= 4) {
      int f1 = 0, f2 = 0, f3 = 0;
      for (unsigned id = 0; id < 16; ++id) {
        __cpuid_count(4, id, a, b, c, d);
        const unsigned type = a & 0x0F;
        if (type == 0) break;
        if (type == 1 || type == 3) {
          const int size = (int)(((b >> 22) + 1) * (((b >> 12) & 0x3FF) + 1) * ((b & 0xFFF) + 1) * (c + 1));
          const unsigned level = (a >> 5) & 7;
          if (level == 1) f1 = size;
          if (level == 2) f2 = size;
          if (level == 3) f3 = size;
        }
      }
      if (f1) l1 = f1;
      if (f2) l2 = f2;
      if (f3) l3 = f3;
    }
  }
  *l1p = l1;
  *l2p = l2;
  *l3p = l3;
}

/* Owl's block sizes for one GEMM whose sizes are passed as (k, m, n): the
 * contraction k is cut to fit L1 in strips of 8, then n (or, when n fits and
 * k was not cut, m) to fit a fixed 1.5 MiB "L2"; each cut spreads the sizes
 * evenly over the blocks.  The callers pass (K, Co, M) for the forward and
 * (M, K, Co) for both backward passes, as Owl does. */
void oracle_owl_block_sizes(int *kp, int *mp, int *np, int ts, int l1, int l2, int l3) {
  (void)l2;
  int k = *kp, m = *mp, n = *np;
  if (k < 50 && m < 50 && n < 50) return;
  const int nr = 4, mr = 16 / (2 * nr) * ts, k_strip = 8;
  const int k_div = (mr + nr) * ts, k_sub = mr * nr * ts;
  int max_kc = ((l1 - k_sub) / k_div) & ~(k_strip - 1);
  if (max_kc < 1) max_kc = 1;
  const int k0 = k;
  if (k > max_kc)
    k = k % max_kc == 0 ? max_kc : max_kc - k_strip * ((max_kc - 1 - k % max_kc) / (k_strip * (k / max_kc + 1)));
  const int l2_fixed = 1572864;
  const int rest_l1 = l1 - k_sub - m * k * ts;
  const int max_nc = rest_l1 >= nr * k * ts ? rest_l1 / (k * ts) : (3 * l2_fixed) / (4 * max_kc * ts);
  int nc = l2_fixed / (2 * k * ts);
  if (nc > max_nc) nc = max_nc;
  nc &= ~(nr - 1);
  if (n > nc) {
    n = n % nc == 0 ? nc : nc - nr * ((nc - n % nc) / (nr * (n / nc + 1)));
  } else if (k0 == k) {
    const int kn_bytes = k * n * ts;
    int lm = l2_fixed, max_mc = m;
    if (kn_bytes < 1024) {
      lm = l1;
    } else if (l3 != 0 && kn_bytes <= 32768) {
      lm = l2;
      if (max_mc > 576) max_mc = 576;
    }
    int mc = lm / (3 * k * ts);
    if (mc > max_mc) mc = max_mc;
    if (mc > mr) {
      mc -= mc % mr;
    } else if (mc == 0) {
      *kp = k; *mp = m; *np = n;
      return;
    }
    m = m % mc == 0 ? mc : mc - mr * ((mc - m % mc) / (mr * (m / mc + 1)));
  }
  *kp = k; *mp = m; *np = n;
}

static void im2col_f32(const oracle_conv_geom *g, const float *x, float *col) {
  const int64_t K = g->kh * g->kw * g->kd * g->Ci;
  const int64_t M = g->B * g->Ho * g->Wo * g->Do;
  for (int64_t m = 0; m < M; ++m) {
    int64_t l = m % g->Do, t = m / g->Do;
    int64_t j = t % g->Wo;
    t /= g->Wo;
    int64_t i = t % g->Ho, b = t / g->Ho;
    float *cp = col + m * K;
    int64_t cnt = 0;
    for (int64_t a = 0; a < g->kh; ++a) {
      int64_t h = i * g->sh - g->ph + a * g->rh;
      for (int64_t c = 0; c < g->kw; ++c) {
        int64_t ww = j * g->sw - g->pw + c * g->rw;
        for (int64_t e = 0; e < g->kd; ++e) {
          int64_t d = l * g->sd - g->pd + e * g->rd;
          int inb = h >= 0 && h < g->H && ww >= 0 && ww < g->W && d >= 0 && d < g->D;
          if (inb) {
            const float *xp = x + IN_IDX(g, b, h, ww, d, 0);
            for (int64_t q = 0; q < g->Ci; ++q) cp[cnt + q] = xp[q];
          } else {
            for (int64_t q = 0; q < g->Ci; ++q) cp[cnt + q] = 0.0f;
          }
          cnt += g->Ci;
        }
      }
    }
  }
}

/* The GEBP branch applies to 2-D undilated layers whose im2col matrix has at
 * least 512 Ki elements, or more than an int counts (Owl's overflow check
 * sends those there too). */
static int owl_gebp(const oracle_conv_geom *g) {
  const int64_t K = g->kh * g->kw * g->Ci, M = g->B * g->Ho * g->Wo;
  return g->D == 1 && g->kd == 1 && g->Do == 1 && g->rh == 1 && g->rw == 1 && K * M >= OWL_IM2COL_THRESHOLD;
}

/* One block of the im2col matrix: rows m0 .. m0 + am, columns k0 .. k0 + ak
 * (column k = (a kw + c) Ci + q), row-major am x ak, zero outside the image.
 * acc = 0 packs x into blk; acc = 1 adds blk into x (the backward input's
 * scatter).  Serial, as Owl's packing is; each tap's channels are one
 * contiguous run, which Owl's AVX loaders (:101-237) also copy as vectors. */
static void gebp_block(const oracle_conv_geom *g, float *x, float *blk, int64_t m0, int am, int64_t k0, int ak,
                       int acc) {
  const int64_t kwc = g->kw * g->Ci;
  const int64_t a0 = k0 / kwc, c0 = (k0 - a0 * kwc) / g->Ci, q0 = k0 - a0 * kwc - c0 * g->Ci;
  for (int r = 0; r < am; ++r) {
    const int64_t m = m0 + r;
    const int64_t j = m % g->Wo, t = m / g->Wo;
    const int64_t i = t % g->Ho, b = t / g->Ho;
    float *row = blk + (int64_t)r * ak;
    int64_t a = a0, c = c0, q = q0;
    for (int e = 0; e < ak;) {
      int64_t run = g->Ci - q;
      if (run > ak - e) run = ak - e;
      const int64_t h = i * g->sh - g->ph + a, w = j * g->sw - g->pw + c;
      const int in = h >= 0 && h < g->H && w >= 0 && w < g->W;
      float *xp = in ? x + IN_IDX(g, b, h, w, 0, q) : NULL;
      if (!acc) {
        if (in) memcpy(row + e, xp, (size_t)run * sizeof(float));
        else memset(row + e, 0, (size_t)run * sizeof(float));
      } else if (in) {
        for (int64_t u = 0; u < run; ++u) xp[u] += row[e + u];
      }
      e += (int)run;
      q = 0;
      if (++c == g->kw) {
        c = 0;
        ++a;
      }
    }
  }
}

static int gebp_fwd(const oracle_conv_geom *g, const float *x, const float *w, float *y, cblas_sgemm_fn gemm,
                    const int *cache) {
  const int K = (int)(g->kh * g->kw * g->Ci), M = (int)(g->B * g->Ho * g->Wo), Co = (int)g->Co;
  int kc = K, nc = Co, mc = M;
  oracle_owl_block_sizes(&kc, &nc, &mc, 4, cache[0], cache[1], cache[2]);
  float *mk = (float *)malloc((size_t)mc * kc * sizeof(float));
  float *kn = (float *)malloc((size_t)kc * nc * sizeof(float));
  float *mn = (float *)malloc((size_t)mc * nc * sizeof(float));
  if (!mk || !kn || !mn) { free(mk); free(kn); free(mn); return -1; }
  memset(y, 0, (size_t)M * Co * sizeof(float));
  for (int m0 = 0; m0 < M; m0 += mc) {
    const int am = M - m0 < mc ? M - m0 : mc;
    for (int k0 = 0; k0 < K; k0 += kc) {
      const int ak = K - k0 < kc ? K - k0 : kc;
      gebp_block(g, (float *)x, mk, m0, am, k0, ak, 0);
      for (int n0 = 0; n0 < Co; n0 += nc) {
        const int an = Co - n0 < nc ? Co - n0 : nc;
        for (int e = 0; e < ak; ++e) memcpy(kn + (int64_t)e * an, w + (int64_t)(k0 + e) * Co + n0, an * sizeof(float));
        gemm(CblasRowMajor_, CblasNoTrans_, CblasNoTrans_, am, an, ak, 1.0f, mk, ak, kn, an, 0.0f, mn, an);
        for (int r = 0; r < am; ++r)
          for (int c = 0; c < an; ++c) y[(int64_t)(m0 + r) * Co + n0 + c] += mn[(int64_t)r * an + c];
      }
    }
  }
  free(mk); free(kn); free(mn);
  return 0;
}

static int gebp_bwd_input(const oracle_conv_geom *g, const float *w, const float *dy, float *dx,
                          cblas_sgemm_fn gemm, const int *cache) {
  const int K = (int)(g->kh * g->kw * g->Ci), M = (int)(g->B * g->Ho * g->Wo), Co = (int)g->Co;
  int mc = M, kc = K, nc = Co;
  oracle_owl_block_sizes(&mc, &kc, &nc, 4, cache[0], cache[1], cache[2]);
  float *mk = (float *)malloc((size_t)mc * kc * sizeof(float));
  float *kn = (float *)malloc((size_t)kc * nc * sizeof(float));
  float *mn = (float *)malloc((size_t)mc * nc * sizeof(float));
  if (!mk || !kn || !mn) { free(mk); free(kn); free(mn); return -1; }
  memset(dx, 0, (size_t)(g->B * g->H * g->W * g->Ci) * sizeof(float));
  for (int m0 = 0; m0 < M; m0 += mc) {
    const int am = M - m0 < mc ? M - m0 : mc;
    for (int k0 = 0; k0 < K; k0 += kc) {
      const int ak = K - k0 < kc ? K - k0 : kc;
      for (int n0 = 0; n0 < Co; n0 += nc) {
        const int an = Co - n0 < nc ? Co - n0 : nc;
        for (int e = 0; e < ak; ++e) memcpy(kn + (int64_t)e * an, w + (int64_t)(k0 + e) * Co + n0, an * sizeof(float));
        for (int r = 0; r < am; ++r) memcpy(mn + (int64_t)r * an, dy + (int64_t)(m0 + r) * Co + n0, an * sizeof(float));
        gemm(CblasRowMajor_, CblasNoTrans_, CblasTrans_, am, ak, an, 1.0f, mn, an, kn, an, 0.0f, mk, ak);
        gebp_block(g, dx, mk, m0, am, k0, ak, 1);
      }
    }
  }
  free(mk); free(kn); free(mn);
  return 0;
}

static int gebp_bwd_kernel(const oracle_conv_geom *g, const float *x, const float *dy, float *dw,
                           cblas_sgemm_fn gemm, const int *cache) {
  const int K = (int)(g->kh * g->kw * g->Ci), M = (int)(g->B * g->Ho * g->Wo), Co = (int)g->Co;
  int mc = M, kc = K, nc = Co;
  oracle_owl_block_sizes(&mc, &kc, &nc, 4, cache[0], cache[1], cache[2]);
  float *mk = (float *)malloc((size_t)mc * kc * sizeof(float));
  float *kn = (float *)malloc((size_t)kc * nc * sizeof(float));
  float *mn = (float *)malloc((size_t)mc * nc * sizeof(float));
  if (!mk || !kn || !mn) { free(mk); free(kn); free(mn); return -1; }
  memset(dw, 0, (size_t)K * Co * sizeof(float));
  for (int m0 = 0; m0 < M; m0 += mc) {
    const int am = M - m0 < mc ? M - m0 : mc;
    for (int k0 = 0; k0 < K; k0 += kc) {
      const int ak = K - k0 < kc ? K - k0 : kc;
      gebp_block(g, (float *)x, mk, m0, am, k0, ak, 0);
      for (int n0 = 0; n0 < Co; n0 += nc) {
        const int an = Co - n0 < nc ? Co - n0 : nc;
        for (int r = 0; r < am; ++r) memcpy(mn + (int64_t)r * an, dy + (int64_t)(m0 + r) * Co + n0, an * sizeof(float));
        gemm(CblasRowMajor_, CblasTrans_, CblasNoTrans_, an, ak, am, 1.0f, mn, an, mk, ak, 0.0f, kn, ak);
        for (int c = 0; c < an; ++c)
          for (int e = 0; e < ak; ++e) dw[(int64_t)(k0 + e) * Co + n0 + c] += kn[(int64_t)c * ak + e];
      }
    }
  }
  free(mk); free(kn); free(mn);
  return 0;
}

/* cache: {l1, l2, l3} bytes for the GEBP block sizes, or NULL for the host's.
 * Returns 0, 1 when the GEBP branch ran, or -1 when out of memory. */
static const int *cache_or_host(const int *cache, int *buf) {
  if (cache) return cache;
  oracle_host_cache_sizes(&buf[0], &buf[1], &buf[2]);
  return buf;
}

int oracle_conv_fwd_owl_f32_cache(const oracle_conv_geom *g, const float *x, const float *w, float *y,
                                  cblas_sgemm_fn gemm, const int *cache) {
  int cb[3];
  if (owl_gebp(g)) return gebp_fwd(g, x, w, y, gemm, cache_or_host(cache, cb)) ? -1 : 1;
  const int64_t K = g->kh * g->kw * g->kd * g->Ci;
  const int64_t M = g->B * g->Ho * g->Wo * g->Do;
  float *col = (float *)malloc((size_t)(M * K) * sizeof(float));
  if (!col) return -1;
  im2col_f32(g, x, col);
  gemm(CblasRowMajor_, CblasNoTrans_, CblasNoTrans_, (int)M, (int)g->Co, (int)K,
       1.0f, col, (int)K, w, (int)g->Co, 0.0f, y, (int)g->Co);
  free(col);
  return 0;
}

int oracle_conv_bwd_input_owl_f32_cache(const oracle_conv_geom *g, const float *w, const float *dy, float *dx,
                                        cblas_sgemm_fn gemm, const int *cache) {
  int cb[3];
  if (owl_gebp(g)) return gebp_bwd_input(g, w, dy, dx, gemm, cache_or_host(cache, cb)) ? -1 : 1;
  const int64_t K = g->kh * g->kw * g->kd * g->Ci;
  const int64_t M = g->B * g->Ho * g->Wo * g->Do;
  float *col = (float *)malloc((size_t)(M * K) * sizeof(float));
  if (!col) return -1;
  gemm(CblasRowMajor_, CblasNoTrans_, CblasTrans_, (int)M, (int)K, (int)g->Co,
       1.0f, dy, (int)g->Co, w, (int)g->Co, 0.0f, col, (int)K);
  memset(dx, 0, (size_t)(g->B * g->H * g->W * g->D * g->Ci) * sizeof(float));
  for (int64_t m = 0; m < M; ++m) {
    int64_t l = m % g->Do, t = m / g->Do;
    int64_t j = t % g->Wo;
    t /= g->Wo;
    int64_t i = t % g->Ho, b = t / g->Ho;
    const float *cp = col + m * K;
    int64_t cnt = 0;
    for (int64_t a = 0; a < g->kh; ++a) {
      int64_t h = i * g->sh - g->ph + a * g->rh;
      for (int64_t c = 0; c < g->kw; ++c) {
        int64_t ww = j * g->sw - g->pw + c * g->rw;
        for (int64_t e = 0; e < g->kd; ++e) {
          int64_t d = l * g->sd - g->pd + e * g->rd;
          if (h >= 0 && h < g->H && ww >= 0 && ww < g->W && d >= 0 && d < g->D) {
            float *xp = dx + IN_IDX(g, b, h, ww, d, 0);
            for (int64_t q = 0; q < g->Ci; ++q) xp[q] += cp[cnt + q];
          }
          cnt += g->Ci;
        }
      }
    }
  }
  free(col);
  return 0;
}

int oracle_conv_bwd_kernel_owl_f32_cache(const oracle_conv_geom *g, const float *x, const float *dy, float *dw,
                                         cblas_sgemm_fn gemm, const int *cache) {
  int cb[3];
  if (owl_gebp(g)) return gebp_bwd_kernel(g, x, dy, dw, gemm, cache_or_host(cache, cb)) ? -1 : 1;
  const int64_t K = g->kh * g->kw * g->kd * g->Ci;
  const int64_t M = g->B * g->Ho * g->Wo * g->Do;
  float *col = (float *)malloc((size_t)(M * K) * sizeof(float));
  float *k2 = (float *)malloc((size_t)(K * g->Co) * sizeof(float));
  if (!col || !k2) { free(col); free(k2); return -1; }
  im2col_f32(g, x, col);
  gemm(CblasRowMajor_, CblasTrans_, CblasNoTrans_, (int)g->Co, (int)K, (int)M,
       1.0f, dy, (int)g->Co, col, (int)K, 0.0f, k2, (int)K);
  int64_t cnt = 0;
  for (int64_t kk = 0; kk < K; ++kk)
    for (int64_t o = 0; o < g->Co; ++o) dw[cnt++] = k2[o * K + kk];
  free(col);
  free(k2);
  return 0;
}

int oracle_conv_fwd_owl_f32(const oracle_conv_geom *g, const float *x, const float *w, float *y,
                            cblas_sgemm_fn gemm) {
  return oracle_conv_fwd_owl_f32_cache(g, x, w, y, gemm, NULL) < 0 ? -1 : 0;
}

int oracle_conv_bwd_input_owl_f32(const oracle_conv_geom *g, const float *w, const float *dy, float *dx,
                                  cblas_sgemm_fn gemm) {
  return oracle_conv_bwd_input_owl_f32_cache(g, w, dy, dx, gemm, NULL) < 0 ? -1 : 0;
}

int oracle_conv_bwd_kernel_owl_f32(const oracle_conv_geom *g, const float *x, const float *dy, float *dw,
                                   cblas_sgemm_fn gemm) {
  return oracle_conv_bwd_kernel_owl_f32_cache(g, x, dy, dw, gemm, NULL) < 0 ? -1 : 0;
}

/* ------------------------------------------------------------------------
 * Pooling, bit-exact restatement of owl_ndarray_pool_impl.h with the macro
 * semantics of owl_ndarray_pool_stub.c:16-51:
 *   max: acc = -INF; if (acc < t) acc = t          (first strict max wins,
 *        NaN never wins)
 *   avg: acc = 0; acc += t; result acc / count_in_bounds
 * Window scan order: H ("cols") outer, W ("rows"), D ("dpts") inner, as in
 * :66-77 and :285-300.  A 2-D call is the 3-D one with D = kd = 1.
 * ---------------------------------------------------------------------- */
#define PIN(g, b, h, w, d, c) \
  ((((((b) * (g)->H + (h)) * (g)->W + (w)) * (g)->D + (d)) * (g)->C) + (c))
#define POUT(g, b, h, w, d, c) \
  ((((((b) * (g)->Ho + (h)) * (g)->Wo + (w)) * (g)->Do + (d)) * (g)->C) + (c))

#define DEFINE_POOL(SUF, T)                                                      \
  void oracle_pool_fwd_##SUF(const oracle_pool_geom *g, int is_max,             \
                             const T *x, T *y) {                                 \
    for (int64_t b = 0; b < g->B; ++b)                                           \
      for (int64_t i = 0; i < g->Ho; ++i)                                        \
        for (int64_t j = 0; j < g->Wo; ++j)                                      \
          for (int64_t l = 0; l < g->Do; ++l)                                    \
            for (int64_t ch = 0; ch < g->C; ++ch) {                              \
              T acc = is_max ? (T)(-INFINITY) : (T)0;                            \
              int cnt = 0;                                                       \
              for (int64_t h = i * g->sh - g->ph; h < i * g->sh - g->ph + g->kh; ++h) \
                for (int64_t w = j * g->sw - g->pw; w < j * g->sw - g->pw + g->kw; ++w) \
                  for (int64_t d = l * g->sd - g->pd; d < l * g->sd - g->pd + g->kd; ++d) \
                    if (h >= 0 && h < g->H && w >= 0 && w < g->W && d >= 0 && d < g->D) { \
                      T t = x[PIN(g, b, h, w, d, ch)];                           \
                      if (is_max) { if (acc < t) acc = t; }                      \
                      else acc += t;                                             \
                      cnt++;                                                     \
                    }                                                            \
              y[POUT(g, b, i, j, l, ch)] = is_max ? acc : acc / (T)cnt;          \
            }                                                                    \
  }                                                                              \
  /* owl_ndarray_pool_impl.h:459-544: value + flat int64 argmax, -1 if none */   \
  void oracle_pool_argmax_##SUF(const oracle_pool_geom *g, const T *x, T *y,    \
                                int64_t *arg) {                                  \
    for (int64_t b = 0; b < g->B; ++b)                                           \
      for (int64_t i = 0; i < g->Ho; ++i)                                        \
        for (int64_t j = 0; j < g->Wo; ++j)                                      \
          for (int64_t l = 0; l < g->Do; ++l)                                    \
            for (int64_t ch = 0; ch < g->C; ++ch) {                              \
              T acc = (T)(-INFINITY);                                            \
              int64_t mi = -1;                                                   \
              for (int64_t h = i * g->sh - g->ph; h < i * g->sh - g->ph + g->kh; ++h) \
                for (int64_t w = j * g->sw - g->pw; w < j * g->sw - g->pw + g->kw; ++w) \
                  for (int64_t d = l * g->sd - g->pd; d < l * g->sd - g->pd + g->kd; ++d) \
                    if (h >= 0 && h < g->H && w >= 0 && w < g->W && d >= 0 && d < g->D) { \
                      int64_t id = PIN(g, b, h, w, d, ch);                       \
                      T t = x[id];                                               \
                      if (acc < t) { acc = t; mi = id; }                         \
                    }                                                            \
              y[POUT(g, b, i, j, l, ch)] = acc;                                  \
              arg[POUT(g, b, i, j, l, ch)] = mi;                                 \
            }                                                                    \
  }                                                                              \
  /* owl_ndarray_pool_impl.h:102-202 and :324-447: dx zeroed, then for each   \
   * window in (b,i,j,l,ch) order: max adds dy to the recomputed first strict  \
   * max (flat index 0 if none), avg adds dy/count to every in-bounds cell. */  \
  void oracle_pool_bwd_##SUF(const oracle_pool_geom *g, int is_max,             \
                             const T *x, const T *dy, T *dx) {                   \
    memset(dx, 0, (size_t)(g->B * g->H * g->W * g->D * g->C) * sizeof(T));       \
    for (int64_t b = 0; b < g->B; ++b)                                           \
      for (int64_t i = 0; i < g->Ho; ++i)                                        \
        for (int64_t j = 0; j < g->Wo; ++j)                                      \
          for (int64_t l = 0; l < g->Do; ++l)                                    \
            for (int64_t ch = 0; ch < g->C; ++ch) {                              \
              T m = dy[POUT(g, b, i, j, l, ch)];                                 \
              T acc = (T)(-INFINITY);                                            \
              int64_t mi = 0;                                                    \
              int cnt = 0;                                                       \
              for (int64_t h = i * g->sh - g->ph; h < i * g->sh - g->ph + g->kh; ++h) \
                for (int64_t w = j * g->sw - g->pw; w < j * g->sw - g->pw + g->kw; ++w) \
                  for (int64_t d = l * g->sd - g->pd; d < l * g->sd - g->pd + g->kd; ++d) \
                    if (h >= 0 && h < g->H && w >= 0 && w < g->W && d >= 0 && d < g->D) { \
                      int64_t id = PIN(g, b, h, w, d, ch);                       \
                      if (is_max) { T t = x[id]; if (acc < t) { acc = t; mi = id; } } \
                      cnt++;                                                     \
                    }                                                            \
              if (is_max) { dx[mi] += m; continue; }                             \
              T v = m / (T)cnt;                                                  \
              for (int64_t h = i * g->sh - g->ph; h < i * g->sh - g->ph + g->kh; ++h) \
                for (int64_t w = j * g->sw - g->pw; w < j * g->sw - g->pw + g->kw; ++w) \
                  for (int64_t d = l * g->sd - g->pd; d < l * g->sd - g->pd + g->kd; ++d) \
                    if (h >= 0 && h < g->H && w >= 0 && w < g->W && d >= 0 && d < g->D) \
                      dx[PIN(g, b, h, w, d, ch)] += v;                           \
            }                                                                    \
  }

DEFINE_POOL(f32, float)
DEFINE_POOL(f64, double)

/* ---------------------------------------------------------------- upsampling
 * Backward: owl_ndarray_upsampling_impl.h:9-60 (spatial_backward), restated
 * loop for loop -- dx zeroed, then for b, output col c, output row r, channel
 * i: dx[b, min(c / col_scale, in_cols - 1), min(r / row_scale, in_rows - 1), i]
 * += dy[b, c, r, i].  Forward: repeat x [|1; col_scale; row_scale; 1|]
 * (owl_dense_ndarray_generic.ml:8223-8234) -- each input cell copied to its
 * col_scale x row_scale block; cells past the scaled extent take the last
 * input row / column, mirroring the backward's clamp. */
#define DEFINE_UPSAMPLE(S, T)                                                          \
  void oracle_upsampling_bwd_##S(T *dx, const T *dy, int64_t B, int64_t in_cols,       \
                                 int64_t in_rows, int64_t C, int64_t out_cols,         \
                                 int64_t out_rows, int64_t col_scale, int64_t row_scale) { \
    memset(dx, 0, (size_t)(B * in_cols * in_rows * C) * sizeof(T));                    \
    for (int64_t b = 0; b < B; ++b)                                                    \
      for (int64_t c = 0; c < out_cols; ++c) {                                         \
        int64_t ic = c / col_scale;                                                    \
        if (ic > in_cols - 1) ic = in_cols - 1;                                        \
        for (int64_t r = 0; r < out_rows; ++r) {                                       \
          int64_t ir = r / row_scale;                                                  \
          if (ir > in_rows - 1) ir = in_rows - 1;                                      \
          for (int64_t i = 0; i < C; ++i)                                              \
            dx[((b * in_cols + ic) * in_rows + ir) * C + i] +=                         \
                dy[((b * out_cols + c) * out_rows + r) * C + i];                       \
        }                                                                              \
      }                                                                                \
  }                                                                                    \
  void oracle_upsampling_fwd_##S(const T *x, T *y, int64_t B, int64_t in_cols,         \
                                 int64_t in_rows, int64_t C, int64_t out_cols,         \
                                 int64_t out_rows, int64_t col_scale, int64_t row_scale) { \
    for (int64_t b = 0; b < B; ++b)                                                    \
      for (int64_t c = 0; c < out_cols; ++c) {                                         \
        int64_t ic = c / col_scale;                                                    \
        if (ic > in_cols - 1) ic = in_cols - 1;                                        \
        for (int64_t r = 0; r < out_rows; ++r) {                                       \
          int64_t ir = r / row_scale;                                                  \
          if (ir > in_rows - 1) ir = in_rows - 1;                                      \
          for (int64_t i = 0; i < C; ++i)                                              \
            y[((b * out_cols + c) * out_rows + r) * C + i] =                           \
                x[((b * in_cols + ic) * in_rows + ir) * C + i];                        \
        }                                                                              \
      }                                                                                \
  }

DEFINE_UPSAMPLE(f32, float)
DEFINE_UPSAMPLE(f64, double)

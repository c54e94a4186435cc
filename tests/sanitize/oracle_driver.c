/* Drives every entry point of the CPU oracle (oracle/owl_oracle.c, included
 * as one translation unit) over random small geometries, for a build under
 * AddressSanitizer / UndefinedBehaviorSanitizer (tests/test_oracle_sanitize.py):
 * 2-D / 3-D, strides, rates, SAME / VALID pads, the GEBP branch with small
 * cache sizes (several m, k and n blocks), pooling with NaN and -inf windows,
 * argmax, and upsampling with clamped extents.  A naive row-major SGEMM stands
 * in for OpenBLAS.  Exits non-zero on a failed consistency check. */
#include "../../oracle/owl_oracle.c"

#include <stdio.h>

static void naive_sgemm(int order, int ta, int tb, int m, int n, int k, float alpha, const float *a, int lda,
                        const float *b, int ldb, float beta, float *c, int ldc) {
  (void)order;
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int p = 0; p < k; ++p) {
        const float av = ta == CblasTrans_ ? a[(int64_t)p * lda + i] : a[(int64_t)i * lda + p];
        const float bv = tb == CblasTrans_ ? b[(int64_t)j * ldb + p] : b[(int64_t)p * ldb + j];
        s += (double)av * bv;
      }
      c[(int64_t)i * ldc + j] = (float)(alpha * s + (beta != 0.0f ? beta * c[(int64_t)i * ldc + j] : 0.0));
    }
}

static uint64_t rs = 88172645463325252ull;
static uint64_t rnd(void) {
  rs ^= rs << 13;
  rs ^= rs >> 7;
  rs ^= rs << 17;
  return rs;
}
static int64_t ri(int64_t lo, int64_t hi) { return lo + (int64_t)(rnd() % (uint64_t)(hi - lo + 1)); }
static double rd(void) { return (double)(rnd() >> 11) / 9007199254740992.0 * 2.0 - 1.0; }

static int64_t out_dim(int same, int64_t n, int64_t ke, int64_t s) {
  return same ? (n + s - 1) / s : (n - ke + s) / s;
}
static int64_t lead(int64_t s, int64_t o, int64_t ke, int64_t n) {
  int64_t t = s * (o - 1) + ke - n;
  return t > 0 ? t / 2 : 0;
}

static int conv_case(int nd, int gebp) {
  oracle_conv_geom g;
  memset(&g, 0, sizeof g);
  g.B = ri(1, 3);
  g.Ci = ri(1, 9);
  g.Co = ri(1, 11);
  int64_t *in[3] = {&g.H, &g.W, &g.D}, *k[3] = {&g.kh, &g.kw, &g.kd}, *o[3] = {&g.Ho, &g.Wo, &g.Do};
  int64_t *s[3] = {&g.sh, &g.sw, &g.sd}, *r[3] = {&g.rh, &g.rw, &g.rd}, *p[3] = {&g.ph, &g.pw, &g.pd};
  const int same = (int)ri(0, 1);
  for (int a = 0; a < 3; ++a) {
    const int used = a < nd;
    *in[a] = used ? (gebp ? ri(24, 32) : ri(2, 9)) : 1;
    *k[a] = used ? (gebp ? 3 : ri(1, 3)) : 1;
    *s[a] = used && !gebp ? ri(1, 2) : 1;
    *r[a] = used && !gebp ? ri(1, 2) : 1;
    const int64_t ke = (*k[a] - 1) * *r[a] + 1;
    if (!same && *in[a] < ke) *in[a] = ke;
    *o[a] = out_dim(same, *in[a], ke, *s[a]);
    *p[a] = same ? lead(*s[a], *o[a], ke, *in[a]) : 0;
  }
  if (gebp) { /* M K >= 8 * 24 * 24 * 9 * 16 > 512 Ki: the GEBP branch */
    g.B = 8;
    g.Ci = ri(16, 32);
    g.Co = ri(9, 40);
  }
  const int64_t nx = g.B * g.H * g.W * g.D * g.Ci, nw = g.kh * g.kw * g.kd * g.Ci * g.Co;
  const int64_t ny = g.B * g.Ho * g.Wo * g.Do * g.Co;
  double *x = malloc(nx * 8), *w = malloc(nw * 8), *dy = malloc(ny * 8), *y = malloc(ny * 8), *dx = malloc(nx * 8),
         *dw = malloc(nw * 8);
  float *xf = malloc(nx * 4), *wf = malloc(nw * 4), *dyf = malloc(ny * 4), *yf = malloc(ny * 4), *dxf = malloc(nx * 4),
        *dwf = malloc(nw * 4);
  for (int64_t i = 0; i < nx; ++i) xf[i] = (float)(x[i] = rd());
  for (int64_t i = 0; i < nw; ++i) wf[i] = (float)(w[i] = rd());
  for (int64_t i = 0; i < ny; ++i) dyf[i] = (float)(dy[i] = rd());
  oracle_conv_fwd_f64(&g, x, w, y);
  oracle_conv_bwd_input_f64(&g, w, dy, dx);
  oracle_conv_bwd_kernel_f64(&g, x, dy, dw);
  /* the FP32 Owl algorithm, small caches so GEBP cuts every dimension */
  const int cache[3] = {(int)ri(1024, 8192), 65536, 1 << 20};
  const int r0 = oracle_conv_fwd_owl_f32_cache(&g, xf, wf, yf, naive_sgemm, gebp ? cache : NULL);
  const int r1 = oracle_conv_bwd_input_owl_f32_cache(&g, wf, dyf, dxf, naive_sgemm, gebp ? cache : NULL);
  const int r2 = oracle_conv_bwd_kernel_owl_f32_cache(&g, xf, dyf, dwf, naive_sgemm, gebp ? cache : NULL);
  const int rc = (r0 < 0 || r1 < 0 || r2 < 0) ? -1 : ((r0 == 1 && r1 == 1 && r2 == 1) != gebp ? -2 : 0);
  double e = 0, n = 0;
  for (int64_t i = 0; i < ny; ++i) { e += (yf[i] - y[i]) * (yf[i] - y[i]); n += y[i] * y[i]; }
  for (int64_t i = 0; i < nx; ++i) { e += (dxf[i] - dx[i]) * (dxf[i] - dx[i]); n += dx[i] * dx[i]; }
  for (int64_t i = 0; i < nw; ++i) { e += (dwf[i] - dw[i]) * (dwf[i] - dw[i]); n += dw[i] * dw[i]; }
  free(x); free(w); free(dy); free(y); free(dx); free(dw);
  free(xf); free(wf); free(dyf); free(yf); free(dxf); free(dwf);
  if (rc < 0 || e > 1e-10 * n) {
    fprintf(stderr, "conv case nd=%d gebp=%d: rc %d rel %g\n", nd, gebp, rc, n > 0 ? e / n : e);
    return 1;
  }
  return 0;
}

static int pool_case(int nd) {
  oracle_pool_geom g;
  memset(&g, 0, sizeof g);
  g.B = ri(1, 3);
  g.C = ri(1, 7);
  int64_t *in[3] = {&g.H, &g.W, &g.D}, *k[3] = {&g.kh, &g.kw, &g.kd}, *o[3] = {&g.Ho, &g.Wo, &g.Do};
  int64_t *s[3] = {&g.sh, &g.sw, &g.sd}, *p[3] = {&g.ph, &g.pw, &g.pd};
  const int same = (int)ri(0, 1);
  for (int a = 0; a < 3; ++a) {
    const int used = a < nd;
    *in[a] = used ? ri(1, 9) : 1;
    *k[a] = used ? ri(1, 4) : 1;
    *s[a] = used ? ri(1, 3) : 1;
    if (!same && *in[a] < *k[a]) *in[a] = *k[a];
    *o[a] = out_dim(same, *in[a], *k[a], *s[a]);
    *p[a] = same ? lead(*s[a], *o[a], *k[a], *in[a]) : 0;
  }
  const int64_t nx = g.B * g.H * g.W * g.D * g.C, ny = g.B * g.Ho * g.Wo * g.Do * g.C;
  float *x = malloc(nx * 4), *y = malloc(ny * 4), *dy = malloc(ny * 4), *dx = malloc(nx * 4), *ya = malloc(ny * 4);
  double *xd = malloc(nx * 8), *yd = malloc(ny * 8), *dyd = malloc(ny * 8), *dxd = malloc(nx * 8);
  int64_t *arg = malloc(ny * 8);
  for (int64_t i = 0; i < nx; ++i) {
    const uint64_t u = rnd() % 16;
    xd[i] = x[i] = u == 0 ? NAN : (u == 1 ? -INFINITY : (float)rd());
  }
  for (int64_t i = 0; i < ny; ++i) dyd[i] = dy[i] = (float)rd();
  int bad = 0;
  for (int is_max = 0; is_max < 2; ++is_max) {
    oracle_pool_fwd_f32(&g, is_max, x, y);
    oracle_pool_bwd_f32(&g, is_max, x, dy, dx);
    oracle_pool_fwd_f64(&g, is_max, xd, yd);
    oracle_pool_bwd_f64(&g, is_max, xd, dyd, dxd);
  }
  if (nd == 2) {
    oracle_pool_argmax_f32(&g, x, ya, arg);
    for (int64_t i = 0; i < ny; ++i) {
      if (arg[i] < -1 || arg[i] >= nx) bad = 1;
      if (arg[i] >= 0 && !(x[arg[i]] == ya[i])) bad = 1;
    }
  }
  free(x); free(y); free(dy); free(dx); free(ya); free(xd); free(yd); free(dyd); free(dxd); free(arg);
  if (bad) fprintf(stderr, "pool case nd=%d: argmax inconsistent\n", nd);
  return bad;
}

static int upsample_case(void) {
  const int64_t B = ri(1, 2), H = ri(1, 5), W = ri(1, 5), C = ri(1, 4), sh = ri(1, 3), sw = ri(1, 3);
  const int64_t Ho = H * sh + ri(-1, 1), Wo = W * sw + ri(-1, 1);
  if (Ho < 1 || Wo < 1) return 0;
  float *x = malloc(B * H * W * C * 4), *y = malloc(B * Ho * Wo * C * 4), *dx = malloc(B * H * W * C * 4);
  for (int64_t i = 0; i < B * H * W * C; ++i) x[i] = (float)rd();
  oracle_upsampling_fwd_f32(x, y, B, H, W, C, Ho, Wo, sh, sw);
  oracle_upsampling_bwd_f32(dx, y, B, H, W, C, Ho, Wo, sh, sw);
  free(x); free(y); free(dx);
  return 0;
}

int main(void) {
  int bad = 0;
  int l1, l2, l3;
  oracle_host_cache_sizes(&l1, &l2, &l3);
  for (int t = 0; t < 40; ++t) bad |= conv_case(2, 0);
  for (int t = 0; t < 20; ++t) bad |= conv_case(3, 0);
  for (int t = 0; t < 4; ++t) bad |= conv_case(2, 1);
  for (int t = 0; t < 60; ++t) bad |= pool_case(2) | pool_case(3);
  for (int t = 0; t < 40; ++t) bad |= upsample_case();
  printf("oracle driver: %s (host caches %d %d %d)\n", bad ? "FAILED" : "ok", l1, l2, l3);
  return bad;
}

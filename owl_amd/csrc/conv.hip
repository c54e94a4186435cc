// Implicit-GEMM convolution kernels for gfx950 (forward, backward-input,
// backward-kernel) over Owl's NHWC / NDHWC layouts.
//
// Every pass is one GEMM C[M][N] = sum_k A[m][k] * B[k][n] whose operands are
// gathered straight from the activation / kernel tensors (no im2col buffer):
//
//   pass     M            N    K            A[m][k]                 B[k][n]
//   forward  B*Ho*Wo*Do   Co   T*Ci         x at tap t, channel ci  w[t][ci][n]
//   bwd in   B*H*W*D      Ci   T*Co         dy at the output pixel  w[t][n][co]
//                                           that tap t of pixel m
//                                           reads (if any)
//   bwd ker  T*Ci         Co   B*Ho*Wo*Do   x at pixel k, tap/ci m  dy[k][n]
//
// (T = kh*kw*kd taps.)  This is the arithmetic Owl's C performs with im2col +
// cblas_?gemm (owl_ndarray_conv_impl.h:297-338 forward, :509-546 backward
// input, :719-766 backward kernel; 3-D :1211-1554; dilated :3097-3784); the
// backward input pass is a gather, so no two lanes write one element.
//
// Tiles are staged in LDS with k contiguous in each row, so every MFMA
// fragment is a 32-byte read per lane whatever the element type.  k within a
// K-step is permuted so that each lane's k values are contiguous; A and B use
// the same permutation, which leaves the sum unchanged.
//
// MFMA per element type (FP32 accumulate except FP64):
//   bf16 / f16  v_mfma_f32_32x32x16_{bf16,f16}   BK = 32
//   f32         v_mfma_f32_32x32x2f32            BK = 16
//   f64         v_mfma_f64_16x16x4f64            BK = 16
#include "owl_amd_common.h"
#include "conv_kernels.h"

namespace owl_amd {

template <typename TI> struct Mfma;

// E8M0 exponent that the scaled FP8 MFMA reads as 2^0 (the per-tensor scales
// are applied in the epilogue instead).
#ifndef OWL_FP8_MFMA_SCALE_ONE
#define OWL_FP8_MFMA_SCALE_ONE 127
#endif

template <> struct Mfma<fp8e4m3_t> {
  static constexpr int TM = 32, KPL = 32, BK = 64, BM = 128, BN = 128;
  typedef float acc_elem;
  typedef f32x16 acc_t;
  static constexpr int NACC = 16;
  __device__ static __forceinline__ acc_t mma(const fp8e4m3_t *a, const fp8e4m3_t *b, acc_t c) {
    i32x8 av, bv;
    __builtin_memcpy(&av, a, 32);
    __builtin_memcpy(&bv, b, 32);
    // cbsz = blgp = 0: both operands OCP FP8 e4m3
    return __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(av, bv, c, 0, 0, 0, OWL_FP8_MFMA_SCALE_ONE, 0,
                                                           OWL_FP8_MFMA_SCALE_ONE);
  }
};

template <> struct Mfma<bf16_t> {
  static constexpr int TM = 32, KPL = 8, BK = 32, BM = 128, BN = 128;
  typedef float acc_elem;
  typedef f32x16 acc_t;
  static constexpr int NACC = 16;
  __device__ static __forceinline__ acc_t mma(const bf16_t *a, const bf16_t *b, acc_t c) {
    bf16x8 av, bv;
    __builtin_memcpy(&av, a, 16);
    __builtin_memcpy(&bv, b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
  }
};

template <> struct Mfma<f16_t> {
  static constexpr int TM = 32, KPL = 8, BK = 32, BM = 128, BN = 128;
  typedef float acc_elem;
  typedef f32x16 acc_t;
  static constexpr int NACC = 16;
  __device__ static __forceinline__ acc_t mma(const f16_t *a, const f16_t *b, acc_t c) {
    f16x8 av, bv;
    __builtin_memcpy(&av, a, 16);
    __builtin_memcpy(&bv, b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, c, 0, 0, 0);
  }
};

template <> struct Mfma<float> {
  static constexpr int TM = 32, KPL = 1, BK = 16, BM = 128, BN = 128;
  typedef float acc_elem;
  typedef f32x16 acc_t;
  static constexpr int NACC = 16;
  __device__ static __forceinline__ acc_t mma(const float *a, const float *b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x2f32(*a, *b, c, 0, 0, 0);
  }
};

template <> struct Mfma<double> {
  static constexpr int TM = 16, KPL = 1, BK = 16, BM = 64, BN = 64;
  typedef double acc_elem;
  typedef f64x4 acc_t;
  static constexpr int NACC = 4;
  __device__ static __forceinline__ acc_t mma(const double *a, const double *b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(*a, *b, c, 0, 0, 0);
  }
};

// Row index of accumulator element r held by lane l (column is l % TM).
template <int TM> __device__ __forceinline__ int acc_row(int lane, int r);
template <> __device__ __forceinline__ int acc_row<32>(int lane, int r) {
  return 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
}
// v_mfma_f64_16x16x4f64 (measured on gfx950): lane l, register r holds
// D[l / 16 + 4 r][l % 16].
template <> __device__ __forceinline__ int acc_row<16>(int lane, int r) {
  return (lane >> 4) + 4 * r;
}

constexpr int NT = 256;  // threads per workgroup (4 waves, 2 x 2)

template <typename TI, int BMX = Mfma<TI>::BM> struct Tile {
  typedef Mfma<TI> MF;
  static constexpr int BM = BMX, BN = BMX, BK = MF::BK;
  static constexpr int VEC = 16 / (int)sizeof(TI);     // elements per 16-byte vector
  static constexpr int PAD = 16 / (int)sizeof(TI);     // 16 bytes of row padding
  static constexpr int LDR = BK + PAD;                 // LDS row length (elements)
  static constexpr int WM = BM / 2, WN = BN / 2;       // wave tile
  static constexpr int FM = WM / MF::TM, FN = WN / MF::TM;
  static constexpr int KG = 64 / MF::TM;               // lane k-groups per MFMA
  static constexpr int LK = BK / KG;                   // k values per lane per K-step
  static constexpr int NINS = LK / MF::KPL;            // MFMAs per fragment pair per K-step
  // K-contiguous loader: vectors of VEC along k.
  static constexpr int KC_VPR = BK / VEC;              // vectors per row
  static constexpr int KC_ROWS_PER_PASS = NT / KC_VPR;
  // MN-contiguous (transposing) loader: VEC rows x KV k-values per thread.
  static constexpr int MN_RG = BM / VEC;               // row groups (BM == BN)
  static constexpr int MN_KGR = NT / MN_RG;            // k groups
  static constexpr int MN_KV = BK / MN_KGR;            // k values per thread
  static constexpr int ELEMS = BM * BK / NT;           // elements per thread per operand
  static_assert(BM == BN, "square tiles assumed by the loaders");
  static_assert(KC_ROWS_PER_PASS * (ELEMS / VEC) == BM, "K-contiguous loader covers the tile");
  static_assert(MN_KGR * MN_KV == BK, "MN loader covers BK");
  static_assert(MN_KV * VEC == ELEMS, "MN loader element count");
};

template <typename T> __device__ __forceinline__ T zero_of() { return (T)0.0f; }
template <> __device__ __forceinline__ fp8e4m3_t zero_of<fp8e4m3_t>() { return (fp8e4m3_t)0; }

// Decode a tap index t into (a, c, e) for a [kh][kw][kd] kernel.
__device__ __forceinline__ void tap_decode(int t, int kw, int kd, int &a, int &c, int &e) {
  int kwd = kw * kd;
  a = t / kwd;
  int r = t - a * kwd;
  c = r / kd;
  e = r - c * kd;
}

// A mixed-radix counter: digits d0 (fastest), d1, d2 with radices r0..r2
// and an unbounded top digit.  step() adds a fixed amount, given as digits
// of the same radices, with one carry per digit -- the loaders advance their
// k position by a K-step this way instead of dividing every step.
struct Radix3 {
  int d0, d1, d2, r0, r1, r2, s0, s1, s2;
  int64_t top, stop;
  __device__ __forceinline__ void init(int64_t v, int64_t step, int ra, int rb, int rc) {
    r0 = ra; r1 = rb; r2 = rc;
    d0 = (int)(v % r0); v /= r0;
    d1 = (int)(v % r1); v /= r1;
    d2 = (int)(v % r2); top = v / r2;
    s0 = (int)(step % r0); step /= r0;
    s1 = (int)(step % r1); step /= r1;
    s2 = (int)(step % r2); stop = step / r2;
  }
  __device__ __forceinline__ void step() {
    int c;
    d0 += s0;
    c = d0 >= r0;
    if (c) d0 -= r0;
    d1 += s1 + c;
    c = d1 >= r1;
    if (c) d1 -= r1;
    d2 += s2 + c;
    c = d2 >= r2;
    if (c) d2 -= r2;
    top += stop + c;
  }
};

// ---------------------------------------------------------------------------
// The kernel.  PASS: 0 forward, 1 backward input, 2 backward kernel.
// ALIGNED: channel counts are multiples of VEC and pointers 16-byte aligned,
// so operands move as 16-byte vectors; otherwise element by element.
// ---------------------------------------------------------------------------
// LDS bytes one tile of the body below uses (two stages of A and B).
template <typename TI, int BMX> constexpr int igemm_lds_bytes() {
  return 2 * (2 * BMX) * Tile<TI, BMX>::LDR * (int)sizeof(TI);
}

// One output tile (index `tile` in M-major, N-minor order) of one split of
// the GEMM, computed by the calling workgroup.  `lds_raw` must hold
// igemm_lds_bytes<TI, BMX>() bytes.
template <int PASS, typename TI, typename TO, bool ALIGNED, int BMX>
__device__ __forceinline__ void igemm_tile(const ConvParams &p, int64_t tile, int split_idx, char *lds_raw) {
  typedef Tile<TI, BMX> TL;
  typedef Mfma<TI> MF;
  constexpr int BM = TL::BM, BN = TL::BN, BK = TL::BK, VEC = TL::VEC, LDR = TL::LDR;
  constexpr int ELEMS = TL::ELEMS;

  typedef TI stage_t[(BM + BN) * LDR];
  stage_t *lds = (stage_t *)lds_raw;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;

  const int64_t M = p.M;
  const int N = p.N;
  const int64_t Kfull = p.K;

  const int n_tiles = (N + BN - 1) / BN;
  const int64_t m0 = (tile / n_tiles) * BM;
  const int n0 = (int)(tile % n_tiles) * BN;

  // split-K: the backward kernel's pixel splits, or a grouped forward's
  // (tap, channel) splits
  const bool splitk = PASS == 2 || (PASS == 0 && p.split > 1);
  int64_t kbeg = 0, kend = Kfull;
  if (splitk) {
    kbeg = (int64_t)split_idx * p.k_per_split;
    kend = kbeg + p.k_per_split;
    if (kend > Kfull) kend = Kfull;
  }

  const TI *__restrict__ Ag = (const TI *)p.a;
  const TI *__restrict__ Bg = (const TI *)p.b;

  // ---------------- per-thread loader state ----------------
  // A operand, passes 0/1 (K-contiguous gather): each thread owns ELEMS/VEC
  // rows, all at the same k offset kvA within the K-step.
  constexpr int AR = ELEMS / VEC;
  const int kvA = (tid % TL::KC_VPR) * VEC;
  int64_t a_base[AR];
  int a_q0[AR], a_q1[AR], a_q2[AR];  // spatial origin of the row's window
  bool a_ok[AR];
  // A operand, pass 2 (MN-contiguous over m = (tap, ci)); B operand pass 0/2.
  // k-groups vary fastest across lanes: the transposed LDS writes of one
  // instruction then fill consecutive bytes of a row (no bank conflicts).
  const int kgp = tid % TL::MN_KGR;
  const int rg = tid / TL::MN_KGR;

  if (PASS == 0 || PASS == 1) {
#pragma unroll
    for (int i = 0; i < AR; ++i) {
      int row = tid / TL::KC_VPR + i * TL::KC_ROWS_PER_PASS;
      int64_t m = m0 + row;
      a_ok[i] = m < M;
      if (!a_ok[i]) m = 0;
      if (PASS == 0) {
        int64_t t = m;
        int l = (int)(t % p.Do); t /= p.Do;
        int j = (int)(t % p.Wo); t /= p.Wo;
        int ii = (int)(t % p.Ho);
        int64_t b = t / p.Ho;
        a_base[i] = b * (int64_t)p.H * p.W * p.D * p.Ci;
        a_q0[i] = ii * p.sh - p.ph;
        a_q1[i] = j * p.sw - p.pw;
        a_q2[i] = l * p.sd - p.pd;
      } else {
        int64_t t = m;
        int d = (int)(t % p.D); t /= p.D;
        int w = (int)(t % p.W); t /= p.W;
        int h = (int)(t % p.H);
        int64_t b = t / p.H;
        a_base[i] = b * (int64_t)p.Ho * p.Wo * p.Do * p.Co;
        a_q0[i] = h + p.ph;
        a_q1[i] = w + p.pw;
        a_q2[i] = d + p.pd;
      }
    }
  }
  // Pass 2: the thread's VEC rows m are fixed; precompute tap / channel.
  int w_tap_a = 0, w_tap_c = 0, w_tap_e = 0, w_ci = 0;
  bool w_row_ok = false;
  if (PASS == 2 && ALIGNED) {
    int64_t m = m0 + rg * VEC;
    w_row_ok = m < M;
    int mm = w_row_ok ? (int)m : 0;
    int t = mm / p.Ci;
    w_ci = mm - t * p.Ci;
    tap_decode(t, p.kw, p.kd, w_tap_a, w_tap_c, w_tap_e);
  }

  // k positions of this thread's loads, stepped by BK per K-step:
  // passes 0/1: k = ((ta * kw + tc) * kd + te) * C + c  -> (c, te, tc; ta)
  // pass 2:     k = ((b * Ho + i) * Wo + j) * Do + l    -> (l, j, i; b)
  Radix3 kc;
  if (PASS == 0 || PASS == 1) kc.init(kbeg + kvA, BK, PASS == 0 ? p.Ci : p.Co, p.kd, p.kw);
  else kc.init(kbeg + kgp * TL::MN_KV, BK, p.Do, p.Wo, p.Ho);

  TI ra[ELEMS], rb[ELEMS];

  // ---------------- global -> registers ----------------
  auto load_a = [&](int64_t k0) __attribute__((always_inline)) {
    if (PASS == 0 || PASS == 1) {
      const int C = PASS == 0 ? p.Ci : p.Co;
      if (ALIGNED) {
        int64_t k = k0 + kvA;
        bool kok = k < kend;
        const int c = kc.d0, te = kc.d1, tc = kc.d2, ta = (int)kc.top;
#pragma unroll
        for (int i = 0; i < AR; ++i) {
          bool ok = kok && a_ok[i];
          int64_t off = 0;
          if (PASS == 0) {
            int h = a_q0[i] + ta * p.rh, w = a_q1[i] + tc * p.rw, d = a_q2[i] + te * p.rd;
            ok = ok && h >= 0 && h < p.H && w >= 0 && w < p.W && d >= 0 && d < p.D;
            off = a_base[i] + (((int64_t)h * p.W + w) * p.D + d) * p.Ci + c;
          } else {
            int th = a_q0[i] - ta * p.rh, tw = a_q1[i] - tc * p.rw, td = a_q2[i] - te * p.rd;
            ok = ok && th >= 0 && tw >= 0 && td >= 0;
            int oi = th / p.sh, oj = tw / p.sw, ol = td / p.sd;
            ok = ok && oi * p.sh == th && oj * p.sw == tw && ol * p.sd == td &&
                 oi < p.Ho && oj < p.Wo && ol < p.Do;
            off = a_base[i] + (((int64_t)oi * p.Wo + oj) * p.Do + ol) * p.Co + c;
          }
          uint4 v = make_uint4(0, 0, 0, 0);
          if (ok) v = *(const uint4 *)(Ag + off);
          __builtin_memcpy(&ra[i * VEC], &v, 16);
        }
      } else {
#pragma unroll
        for (int i = 0; i < AR; ++i) {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            int64_t k = k0 + kvA + v;
            bool ok = a_ok[i] && k < kend;
            int t = ok ? (int)(k / C) : 0;
            int c = (int)(k - (int64_t)t * C);
            int ta, tc, te;
            tap_decode(t, p.kw, p.kd, ta, tc, te);
            int64_t off = 0;
            if (PASS == 0) {
              int h = a_q0[i] + ta * p.rh, w = a_q1[i] + tc * p.rw, d = a_q2[i] + te * p.rd;
              ok = ok && h >= 0 && h < p.H && w >= 0 && w < p.W && d >= 0 && d < p.D;
              off = a_base[i] + (((int64_t)h * p.W + w) * p.D + d) * p.Ci + c;
            } else {
              int th = a_q0[i] - ta * p.rh, tw = a_q1[i] - tc * p.rw, td = a_q2[i] - te * p.rd;
              ok = ok && th >= 0 && tw >= 0 && td >= 0;
              int oi = th / p.sh, oj = tw / p.sw, ol = td / p.sd;
              ok = ok && oi * p.sh == th && oj * p.sw == tw && ol * p.sd == td &&
                   oi < p.Ho && oj < p.Wo && ol < p.Do;
              off = a_base[i] + (((int64_t)oi * p.Wo + oj) * p.Do + ol) * p.Co + c;
            }
            ra[i * VEC + v] = ok ? Ag[off] : zero_of<TI>();
          }
        }
      }
    } else {
      // Pass 2: A[m = (tap, ci)][k = output pixel] = x at that pixel's window
      // position for the tap; thread loads VEC rows x MN_KV pixels.
      constexpr int KV = TL::MN_KV;
      int64_t k = k0 + kgp * KV;
      // the first pixel's coordinates, then step pixel by pixel
      int l = kc.d0, j = kc.d1, ii = kc.d2;
      int64_t b = kc.top;
#pragma unroll
      for (int kv = 0; kv < KV; ++kv) {
        bool kok = (k + kv) < kend;
        if (ALIGNED) {
          int h = ii * p.sh - p.ph + w_tap_a * p.rh;
          int w = j * p.sw - p.pw + w_tap_c * p.rw;
          int d = l * p.sd - p.pd + w_tap_e * p.rd;
          bool ok = kok && w_row_ok && h >= 0 && h < p.H && w >= 0 && w < p.W && d >= 0 && d < p.D;
          TI tmp[VEC];
          if (ok) {
            int64_t off = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.Ci + w_ci;
            uint4 v = *(const uint4 *)(Ag + off);
            __builtin_memcpy(tmp, &v, 16);
          } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) tmp[v] = zero_of<TI>();
          }
#pragma unroll
          for (int v = 0; v < VEC; ++v) ra[v * KV + kv] = tmp[v];
        } else {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            int64_t m = m0 + rg * VEC + v;
            bool ok = kok && m < M;
            int mm = ok ? (int)m : 0;
            int tt = mm / p.Ci;
            int ci = mm - tt * p.Ci;
            int ta, tc, te;
            tap_decode(tt, p.kw, p.kd, ta, tc, te);
            int h = ii * p.sh - p.ph + ta * p.rh;
            int w = j * p.sw - p.pw + tc * p.rw;
            int d = l * p.sd - p.pd + te * p.rd;
            ok = ok && h >= 0 && h < p.H && w >= 0 && w < p.W && d >= 0 && d < p.D;
            int64_t off = (((b * p.H + h) * (int64_t)p.W + w) * p.D + d) * (int64_t)p.Ci + ci;
            ra[v * KV + kv] = ok ? Ag[off] : zero_of<TI>();
          }
        }
        // step to the next output pixel
        if (++l == p.Do) {
          l = 0;
          if (++j == p.Wo) {
            j = 0;
            if (++ii == p.Ho) { ii = 0; ++b; }
          }
        }
      }
    }
  };

  auto load_b = [&](int64_t k0) __attribute__((always_inline)) {
    if (PASS == 1) {
      // B[k = (tap, co)][n = ci] = w[tap][n][co]: K-contiguous.
      constexpr int BR = ELEMS / VEC;
#pragma unroll
      for (int i = 0; i < BR; ++i) {
        int row = tid / TL::KC_VPR + i * TL::KC_ROWS_PER_PASS;
        int n = n0 + row;
        if (ALIGNED) {
          int64_t k = k0 + kvA;
          bool ok = k < kend && n < N;
          uint4 v = make_uint4(0, 0, 0, 0);
          if (ok) {
            // same k as the A load: tap and channel from its counter
            const int t = ((int)kc.top * p.kw + kc.d2) * p.kd + kc.d1;
            v = *(const uint4 *)(Bg + ((int64_t)t * p.Ci + n) * p.Co + kc.d0);
          }
          __builtin_memcpy(&rb[i * VEC], &v, 16);
        } else {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            int64_t k = k0 + kvA + v;
            bool ok = k < kend && n < N;
            TI val = zero_of<TI>();
            if (ok) {
              int t = (int)(k / p.Co);
              int co = (int)(k - (int64_t)t * p.Co);
              val = Bg[((int64_t)t * p.Ci + n) * p.Co + co];
            }
            rb[i * VEC + v] = val;
          }
        }
      }
    } else {
      // Pass 0: B[k][n] = w[k*Co + n].  Pass 2: B[k][n] = dy[k*Co + n].
      // Both MN-contiguous along n.
      constexpr int KV = TL::MN_KV;
      int n = n0 + rg * VEC;
#pragma unroll
      for (int kv = 0; kv < KV; ++kv) {
        int64_t k = k0 + kgp * KV + kv;
        bool kok = k < kend;
        if (ALIGNED) {
          TI tmp[VEC];
          if (kok && n < N) {
            uint4 v = *(const uint4 *)(Bg + k * p.Co + n);
            __builtin_memcpy(tmp, &v, 16);
          } else {
#pragma unroll
            for (int v = 0; v < VEC; ++v) tmp[v] = zero_of<TI>();
          }
#pragma unroll
          for (int v = 0; v < VEC; ++v) rb[v * KV + kv] = tmp[v];
        } else {
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            bool ok = kok && (n + v) < N;
            rb[v * KV + kv] = ok ? Bg[k * p.Co + n + v] : zero_of<TI>();
          }
        }
      }
    }
  };

  // ---------------- registers -> LDS ----------------
  auto store_tiles = [&](int buf) __attribute__((always_inline)) {
    TI *As = lds[buf];
    TI *Bs = lds[buf] + BM * LDR;
    if (PASS == 0 || PASS == 1) {
#pragma unroll
      for (int i = 0; i < AR; ++i) {
        int row = tid / TL::KC_VPR + i * TL::KC_ROWS_PER_PASS;
        __builtin_memcpy(&As[row * LDR + kvA], &ra[i * VEC], 16);
      }
    } else {
      constexpr int KV = TL::MN_KV;
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        __builtin_memcpy(&As[(rg * VEC + v) * LDR + kgp * KV], &ra[v * KV], KV * sizeof(TI));
    }
    if (PASS == 1) {
      constexpr int BR = ELEMS / VEC;
#pragma unroll
      for (int i = 0; i < BR; ++i) {
        int row = tid / TL::KC_VPR + i * TL::KC_ROWS_PER_PASS;
        __builtin_memcpy(&Bs[row * LDR + kvA], &rb[i * VEC], 16);
      }
    } else {
      constexpr int KV = TL::MN_KV;
#pragma unroll
      for (int v = 0; v < VEC; ++v)
        __builtin_memcpy(&Bs[(rg * VEC + v) * LDR + kgp * KV], &rb[v * KV], KV * sizeof(TI));
    }
  };

  // ---------------- MFMA main loop ----------------
  typename MF::acc_t acc[TL::FM][TL::FN];
#pragma unroll
  for (int i = 0; i < TL::FM; ++i)
#pragma unroll
    for (int j = 0; j < TL::FN; ++j)
#pragma unroll
      for (int r = 0; r < MF::NACC; ++r) acc[i][j][r] = 0;

  const int frow = lane % MF::TM;
  const int fkg = lane / MF::TM;

  auto compute = [&](int buf) __attribute__((always_inline)) {
    const TI *As = lds[buf];
    const TI *Bs = lds[buf] + BM * LDR;
    TI fa[TL::FM][TL::LK], fb[TL::FN][TL::LK];
#pragma unroll
    for (int i = 0; i < TL::FM; ++i) {
      const TI *src = &As[(wm * TL::WM + i * MF::TM + frow) * LDR + fkg * TL::LK];
      __builtin_memcpy(&fa[i][0], src, TL::LK * sizeof(TI));
    }
#pragma unroll
    for (int j = 0; j < TL::FN; ++j) {
      const TI *src = &Bs[(wn * TL::WN + j * MF::TM + frow) * LDR + fkg * TL::LK];
      __builtin_memcpy(&fb[j][0], src, TL::LK * sizeof(TI));
    }
#pragma unroll
    for (int s = 0; s < TL::NINS; ++s)
#pragma unroll
      for (int i = 0; i < TL::FM; ++i)
#pragma unroll
        for (int j = 0; j < TL::FN; ++j)
          acc[i][j] = MF::mma(&fa[i][s * MF::KPL], &fb[j][s * MF::KPL], acc[i][j]);
  };

  const int64_t nk = (kend - kbeg + BK - 1) / BK;
  if (nk > 0) {
    load_a(kbeg);
    load_b(kbeg);
    kc.step();
    store_tiles(0);
    __syncthreads();
    for (int64_t kt = 0; kt < nk; ++kt) {
      const int cur = (int)(kt & 1);
      const bool more = kt + 1 < nk;
      if (more) {
        load_a(kbeg + (kt + 1) * BK);
        load_b(kbeg + (kt + 1) * BK);
        kc.step();
      }
      compute(cur);
      if (more) store_tiles(cur ^ 1);
      __syncthreads();
    }
  }

  // ---------------- epilogue ----------------
  const int col = lane % MF::TM;
  // FP8: undo the per-tensor quantisation scales of both operands
  const float oscale = p.scale_a ? (*p.scale_a) * (*p.scale_b) : 1.0f;
#pragma unroll
  for (int i = 0; i < TL::FM; ++i)
#pragma unroll
    for (int j = 0; j < TL::FN; ++j) {
      int n = n0 + wn * TL::WN + j * MF::TM + col;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < MF::NACC; ++r) {
        int64_t m = m0 + wm * TL::WM + i * MF::TM + acc_row<MF::TM>(lane, r);
        if (m >= M) continue;
        if (splitk && p.split > 1) {
          typedef typename MF::acc_elem TA;
          TA *ws = (TA *)p.c + (int64_t)split_idx * M * N;
          ws[m * N + n] = (TA)acc[i][j][r];
        } else {
          typedef typename MF::acc_elem TA;
          TA v = p.scale_a ? (TA)((float)acc[i][j][r] * oscale) : (TA)acc[i][j][r];
          if (PASS == 0 && (p.bias || p.act)) {
            // fused bias and activation (Owl's relu is fmax(x, 0): NaN -> 0)
            if (p.bias) v += ((const TA *)p.bias)[n];
            if (p.pre) ((TO *)p.pre)[m * p.ldc + n] = (TO)v;
            if (p.act == OWL_AMD_ACT_RELU) v = fmax(v, (TA)0);
          }
          ((TO *)p.c)[m * p.ldc + n] = (TO)v;
        }
      }
    }
}

// 1-D grid of tiles x splits, split-major in launch order.
template <int PASS, typename TI, typename TO, bool ALIGNED>
__global__ __launch_bounds__(NT) void conv_igemm_kernel(ConvParams p, int tiles) {
  constexpr int BMX = Mfma<TI>::BM;
  __shared__ __attribute__((aligned(16))) char lds[igemm_lds_bytes<TI, BMX>()];
  const int lid = blockIdx.x;
  igemm_tile<PASS, TI, TO, ALIGNED, BMX>(p, lid % tiles, lid / tiles, lds);
}

// Deterministic split-K reduction: dw[e] = sum over splits in split order,
// carried with Neumaier compensation in the accumulator type (FP32, or FP64
// for the FP64 path).  Each thread owns 4 consecutive elements and keeps
// several splits' loads in flight.
template <typename TA>
__device__ __forceinline__ void neumaier(TA &s, TA &comp, TA v) {
  TA t = s + v;
  if (fabs(s) >= fabs(v))
    comp += (s - t) + v;
  else
    comp += (v - t) + s;
  s = t;
}

template <typename TA, typename TO>
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const TA *__restrict__ ws, TO *__restrict__ out,
                                                            int64_t n, int splits) {
  int64_t e0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
  if (e0 >= n) return;
  TA s[4] = {0, 0, 0, 0}, comp[4] = {0, 0, 0, 0};
  if (e0 + 4 <= n && (n % 4) == 0) {
    int k = 0;
    for (; k + 4 <= splits; k += 4) {
      TA v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) v[u][q] = ws[(int64_t)(k + u) * n + e0 + q];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) neumaier(s[q], comp[q], v[u][q]);
    }
    for (; k < splits; ++k)
#pragma unroll
      for (int q = 0; q < 4; ++q) neumaier(s[q], comp[q], ws[(int64_t)k * n + e0 + q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) out[e0 + q] = (TO)(s[q] + comp[q]);
  } else {
    for (int q = 0; q < 4 && e0 + q < n; ++q) {
      for (int k = 0; k < splits; ++k) neumaier(s[q], comp[q], ws[(int64_t)k * n + e0 + q]);
      out[e0 + q] = (TO)(s[q] + comp[q]);
    }
  }
}

// ---------------------------------------------------------------------------
// Ragged batches: many forward convolutions of different shapes in one
// persistent launch.  The host builds a table of problems (each with its
// own tile edge and split-K factor, chosen per shape class) and the prefix
// sum of their tile counts; workgroups stride over the concatenated tile
// list, find their problem by binary search and run the same tile body as
// the single-problem implicit-GEMM kernel, so an unsplit grouped result
// equals that kernel's bit for bit.  A split problem's tiles write FP32
// partials (one slice per split of its (tap, channel) range), which the
// deterministic split-K reduction then sums in split order.
// ---------------------------------------------------------------------------
template <typename TI, typename TO>
__global__ __launch_bounds__(NT) void conv_group_fwd_kernel(const GroupEntry *__restrict__ tab, int nprob,
                                                            int64_t total) {
  constexpr int L0 = igemm_lds_bytes<TI, Mfma<TI>::BM>();
  constexpr int L1 = igemm_lds_bytes<TI, Mfma<TI>::BM / 2>();
  __shared__ __attribute__((aligned(16))) char lds[L0 > L1 ? L0 : L1];
  for (int64_t t = blockIdx.x; t < total; t += gridDim.x) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (tab[mid].tile_begin <= t) lo = mid;
      else hi = mid - 1;
    }
    const ConvParams p = tab[lo].p;
    int64_t local = t - tab[lo].tile_begin;
    // split-major within a problem: split s owns tiles [s tiles_mn, (s+1) tiles_mn)
    int sidx = 0;
    if (p.split > 1) {
      const int64_t tmn = tab[lo].tiles_mn;
      sidx = (int)(local / tmn);
      local -= sidx * tmn;
    }
    const int cfg = tab[lo].cfg, al = tab[lo].aligned;
    if (cfg == 0) {
      if (al) igemm_tile<0, TI, TO, true, Mfma<TI>::BM>(p, local, sidx, lds);
      else igemm_tile<0, TI, TO, false, Mfma<TI>::BM>(p, local, sidx, lds);
    } else {
      if (al) igemm_tile<0, TI, TO, true, Mfma<TI>::BM / 2>(p, local, sidx, lds);
      else igemm_tile<0, TI, TO, false, Mfma<TI>::BM / 2>(p, local, sidx, lds);
    }
    __syncthreads();
  }
}

template <typename TI, typename TO>
void run_group_fwd(const GroupEntry *dev_tab, int nprob, int64_t total, int grid, hipStream_t st) {
  launch_k((conv_group_fwd_kernel<TI, TO>), dim3((unsigned)grid), dim3(NT), st, dev_tab, nprob,
                     total);
}

template <typename TI> int group_tile_edge(int cfg) { return cfg == 0 ? Mfma<TI>::BM : Mfma<TI>::BM / 2; }
template <typename TI> int group_k_step() { return Mfma<TI>::BK; }

template void run_group_fwd<bf16_t, bf16_t>(const GroupEntry *, int, int64_t, int, hipStream_t);
template void run_group_fwd<bf16_t, float>(const GroupEntry *, int, int64_t, int, hipStream_t);
template void run_group_fwd<f16_t, f16_t>(const GroupEntry *, int, int64_t, int, hipStream_t);
template void run_group_fwd<float, float>(const GroupEntry *, int, int64_t, int, hipStream_t);
template int group_tile_edge<bf16_t>(int);
template int group_tile_edge<f16_t>(int);
template int group_tile_edge<float>(int);
template int group_k_step<bf16_t>();
template int group_k_step<f16_t>();
template int group_k_step<float>();

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------
template <int PASS, typename TI, typename TO>
static void launch_igemm(const ConvParams &p, bool aligned, hipStream_t st) {
  typedef Tile<TI> TL;
  int64_t tiles = ((p.M + TL::BM - 1) / TL::BM) * (int64_t)((p.N + TL::BN - 1) / TL::BN);
  const int64_t blocks = tiles * (PASS == 2 || (PASS == 0 && p.split > 1) ? p.split : 1);
  dim3 grid((unsigned)blocks);
  if (aligned)
    launch_k((conv_igemm_kernel<PASS, TI, TO, true>), grid, dim3(NT), st, p, (int)tiles);
  else
    launch_k((conv_igemm_kernel<PASS, TI, TO, false>), grid, dim3(NT), st, p, (int)tiles);
}

template <typename TI> int tile_bm() { return Tile<TI>::BM; }
template <typename TI> int tile_bk() { return Tile<TI>::BK; }

template <int PASS, typename TI, typename TO>
void run_igemm(const ConvParams &p, bool aligned, hipStream_t st) {
  launch_igemm<PASS, TI, TO>(p, aligned, st);
}

// Wide form for n % 4 == 0: a block covers 64 outputs (16 groups of 4) and
// spreads the splits over 16 thread groups (group g sums splits g, g + 16,
// ...); the 16 group sums are then combined in group order, so the result
// is fixed for a given (n, splits) and every launch gives the same bits.
template <typename TA, typename TO>
__global__ __launch_bounds__(256) void splitk_reduce_wide_kernel(const TA *__restrict__ ws, TO *__restrict__ out,
                                                                 int64_t n, int splits) {
  __shared__ TA red[2][16][16 * 4];
  const int c = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int64_t e0 = ((int64_t)blockIdx.x * 16 + c) * 4;
  const bool live = e0 < n;
  TA s[4] = {0, 0, 0, 0}, comp[4] = {0, 0, 0, 0};
  if (live) {
    int k = g;
    for (; k + 48 < splits; k += 64) {
      TA v[4][4];
#pragma unroll
      for (int u = 0; u < 4; ++u) __builtin_memcpy(v[u], ws + (int64_t)(k + 16 * u) * n + e0, 4 * sizeof(TA));
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int q = 0; q < 4; ++q) neumaier(s[q], comp[q], v[u][q]);
    }
    for (; k < splits; k += 16) {
      TA v[4];
      __builtin_memcpy(v, ws + (int64_t)k * n + e0, 4 * sizeof(TA));
#pragma unroll
      for (int q = 0; q < 4; ++q) neumaier(s[q], comp[q], v[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    red[0][g][c * 4 + q] = s[q];
    red[1][g][c * 4 + q] = comp[q];
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const int cc = threadIdx.x >> 2, q = threadIdx.x & 3;
    const int64_t e = ((int64_t)blockIdx.x * 16 + cc) * 4 + q;
    if (e < n) {
      TA t = 0, tc = 0;
      for (int gg = 0; gg < 16; ++gg) {
        neumaier(t, tc, red[0][gg][threadIdx.x]);
        tc += red[1][gg][threadIdx.x];
      }
      out[e] = (TO)(t + tc);
    }
  }
}

template <typename TA, typename TO>
void run_splitk_reduce(const TA *ws, TO *out, int64_t n, int splits, hipStream_t st) {
  if (n % 4 == 0) {
    const int64_t blocks = (n + 63) / 64;
    launch_k((splitk_reduce_wide_kernel<TA, TO>), dim3((unsigned)blocks), dim3(256), st, ws, out, n,
                       splits);
    return;
  }
  int64_t blocks = (n + 1023) / 1024;
  launch_k((splitk_reduce_kernel<TA, TO>), dim3((unsigned)blocks), dim3(256), st,
                     ws, out, n, splits);
}

// Explicit instantiations used by the dispatcher (dispatch.cpp).
#define OWL_INST(PASS, TI, TO) template void run_igemm<PASS, TI, TO>(const ConvParams &, bool, hipStream_t);
OWL_INST(0, float, float)
OWL_INST(1, float, float)
OWL_INST(2, float, float)
OWL_INST(0, double, double)
OWL_INST(1, double, double)
OWL_INST(2, double, double)
OWL_INST(0, bf16_t, bf16_t)
OWL_INST(1, bf16_t, bf16_t)
OWL_INST(2, bf16_t, bf16_t)
OWL_INST(0, bf16_t, float)
OWL_INST(1, bf16_t, float)
OWL_INST(2, bf16_t, float)
OWL_INST(0, f16_t, f16_t)
OWL_INST(1, f16_t, f16_t)
OWL_INST(2, f16_t, f16_t)
OWL_INST(0, f16_t, float)
OWL_INST(1, f16_t, float)
OWL_INST(2, f16_t, float)
OWL_INST(0, fp8e4m3_t, float)
OWL_INST(0, fp8e4m3_t, bf16_t)
#undef OWL_INST
template int tile_bm<fp8e4m3_t>();
template int tile_bk<fp8e4m3_t>();
template void run_splitk_reduce<float, float>(const float *, float *, int64_t, int, hipStream_t);
template void run_splitk_reduce<float, bf16_t>(const float *, bf16_t *, int64_t, int, hipStream_t);
template void run_splitk_reduce<float, f16_t>(const float *, f16_t *, int64_t, int, hipStream_t);
template void run_splitk_reduce<double, double>(const double *, double *, int64_t, int, hipStream_t);
template int tile_bm<float>();
template int tile_bm<double>();
template int tile_bm<bf16_t>();
template int tile_bm<f16_t>();
template int tile_bk<float>();
template int tile_bk<double>();
template int tile_bk<bf16_t>();
template int tile_bk<f16_t>();

}  // namespace owl_amd

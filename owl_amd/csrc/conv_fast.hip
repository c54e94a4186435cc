// Forward and unit-stride backward-input convolution for gfx950, the
// throughput path for channel counts that are multiples of the K-step.
//
// Same GEMM view as conv.hip (forward: M = output pixels, N = Co, K = taps x
// Ci; backward input: M = input pixels, N = Ci, K = taps x Co) but a
// different data flow, chosen for CDNA4's LDS / MFMA balance:
//
//   * A (activations) never touches LDS.  In NHWC the 16 or 32 channels a
//     lane feeds to one MFMA are contiguous, so each lane loads its fragment
//     straight from global memory (32 bytes = two dwordx4 per fragment) in
//     the MFMA operand layout.  Neighbouring taps re-read the same pixels
//     from L1 / L2, never from HBM twice.
//   * B (the kernel tile, BK x BN) is shared by the four waves and staged in
//     LDS, double-buffered, k-contiguous per n row.
//   * The K loop walks taps and channel blocks incrementally: no integer
//     division inside the loop.
//   * The next K-step's A fragments and B tile are in flight while the
//     current step's MFMAs run.
//   * Workgroups are renumbered so that consecutive M tiles (which share
//     input halo rows) run on the same XCD and hit its L2.
#include "conv_kernels.h"
#include "owl_amd_common.h"

namespace owl_amd {

template <typename TI> struct FastMfma;
template <> struct FastMfma<bf16_t> {
  static constexpr int BK = 32, KPL = 8;
  __device__ static __forceinline__ f32x16 mma(const uint4 &a, const uint4 &b, f32x16 c) {
    bf16x8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, c, 0, 0, 0);
  }
};
template <> struct FastMfma<f16_t> {
  static constexpr int BK = 32, KPL = 8;
  __device__ static __forceinline__ f32x16 mma(const uint4 &a, const uint4 &b, f32x16 c) {
    f16x8 av, bv;
    __builtin_memcpy(&av, &a, 16);
    __builtin_memcpy(&bv, &b, 16);
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, c, 0, 0, 0);
  }
};
template <> struct FastMfma<float> {
  static constexpr int BK = 16, KPL = 1;
};

__device__ __forceinline__ int xcd_swizzle(int bid, int nblocks) {
  // Hardware dispatches workgroup i to XCD i % 8.  Give XCD x the contiguous
  // range of tiles [x * nblocks / 8, (x + 1) * nblocks / 8).
  if (nblocks % 8 != 0) return bid;
  return (bid % 8) * (nblocks / 8) + bid / 8;
}

// PASS 0: forward.  PASS 1: backward input with unit strides.
template <int PASS, typename TI, typename TO, int FM, int FN, int WGM, int WGN>
__global__ __launch_bounds__(256, 2) void conv_fast_kernel(ConvParams p) {
  typedef FastMfma<TI> MF;
  constexpr int BK = MF::BK;
  constexpr int LK = BK / 2;                   // k values per lane per fragment per K-step (32 bytes)
  constexpr int BM = WGM * FM * 32, BN = WGN * FN * 32;
  constexpr int LDR = BK + 16 / (int)sizeof(TI);  // LDS row, padded by 16 bytes
  constexpr int BELEMS = BK * BN / 256;        // B elements per thread per K-step
  static_assert(WGM * WGN == 4, "four waves");
  static_assert(LK * sizeof(TI) == 32, "32-byte fragments");

  __shared__ __attribute__((aligned(16))) TI Bs[2][BN * LDR];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int N = p.N;
  const int64_t M = p.M;
  const int n_tiles = (N + BN - 1) / BN;
  const int nblocks = gridDim.x;
  const int bid = xcd_swizzle(blockIdx.x, nblocks);
  const int64_t m0 = (int64_t)(bid / n_tiles) * BM;
  const int n0 = (bid % n_tiles) * BN;

  const int C = PASS == 0 ? p.Ci : p.Co;  // contracted channels per tap
  const int T = p.kh * p.kw * p.kd;
  const int csteps = C / BK;
  const int nk = T * csteps;

  const TI *__restrict__ Ag = (const TI *)p.a;
  const TI *__restrict__ Bg = (const TI *)p.b;

  // ---- per-lane A rows (one per M fragment) ----
  const int frow = lane & 31, fkg = lane >> 5;
  int64_t a_base[FM];
  int a_q0[FM], a_q1[FM], a_q2[FM];
  bool a_ok[FM];
#pragma unroll
  for (int f = 0; f < FM; ++f) {
    int64_t m = m0 + wm * FM * 32 + f * 32 + frow;
    a_ok[f] = m < M;
    if (!a_ok[f]) m = 0;
    int64_t t = m;
    if (PASS == 0) {
      int l = (int)(t % p.Do); t /= p.Do;
      int j = (int)(t % p.Wo); t /= p.Wo;
      int i = (int)(t % p.Ho);
      int64_t b = t / p.Ho;
      a_base[f] = b * (int64_t)p.H * p.W * p.D * p.Ci + fkg * LK;
      a_q0[f] = i * p.sh - p.ph;
      a_q1[f] = j * p.sw - p.pw;
      a_q2[f] = l * p.sd - p.pd;
    } else {
      int d = (int)(t % p.D); t /= p.D;
      int w = (int)(t % p.W); t /= p.W;
      int h = (int)(t % p.H);
      int64_t b = t / p.H;
      a_base[f] = b * (int64_t)p.Ho * p.Wo * p.Do * p.Co + fkg * LK;
      a_q0[f] = h + p.ph;
      a_q1[f] = w + p.pw;
      a_q2[f] = d + p.pd;
    }
  }

  // ---- B staging thread layout ----
  // forward: B[k][n] = w[(t*Ci + c0 + k)*Co + n] is n-contiguous: each thread
  //   loads BKV k-rows of 4 consecutive n and writes 4 n-rows of BKV k.
  // backward input: B[k][n] = w[(t*Ci + n)*Co + c0 + k] is k-contiguous:
  //   each thread loads 16-byte vectors along k.
  constexpr int BKV = BELEMS / 4;                 // forward: k values per thread
  constexpr int VEC = 16 / (int)sizeof(TI);
  constexpr int BNV = BELEMS / VEC;               // backward input: vectors per thread
  const int f_ng = tid % (BN / 4), f_kg = tid / (BN / 4);
  const int b_kv = (tid % (BK / VEC)) * VEC, b_row = tid / (BK / VEC);
  constexpr int B_ROWS_PER_PASS = 256 / (BK / VEC);

  TI rb[BELEMS];
  uint4 ra[2][FM][2];

  // incremental (tap, channel block) walk
  int ta = 0, tc = 0, te = 0, c0 = 0, tap = 0;

  auto load_a = [&](int buf, int ta_, int tc_, int te_, int c0_) {
#pragma unroll
    for (int f = 0; f < FM; ++f) {
      bool ok = a_ok[f];
      int64_t off;
      if (PASS == 0) {
        int h = a_q0[f] + ta_ * p.rh, w = a_q1[f] + tc_ * p.rw, d = a_q2[f] + te_ * p.rd;
        ok = ok && (unsigned)h < (unsigned)p.H && (unsigned)w < (unsigned)p.W && (unsigned)d < (unsigned)p.D;
        off = a_base[f] + (((int64_t)h * p.W + w) * p.D + d) * p.Ci + c0_;
      } else {
        int oi = a_q0[f] - ta_ * p.rh, oj = a_q1[f] - tc_ * p.rw, ol = a_q2[f] - te_ * p.rd;
        ok = ok && (unsigned)oi < (unsigned)p.Ho && (unsigned)oj < (unsigned)p.Wo && (unsigned)ol < (unsigned)p.Do;
        off = a_base[f] + (((int64_t)oi * p.Wo + oj) * p.Do + ol) * p.Co + c0_;
      }
      if (ok) {
        const uint4 *src = (const uint4 *)(Ag + off);
        ra[buf][f][0] = src[0];
        ra[buf][f][1] = src[1];
      } else {
        ra[buf][f][0] = make_uint4(0, 0, 0, 0);
        ra[buf][f][1] = make_uint4(0, 0, 0, 0);
      }
    }
  };

  auto load_b = [&](int tap_, int c0_) {
    if (PASS == 0) {
      const int n = n0 + f_ng * 4;
      const bool nok = n < N;
#pragma unroll
      for (int kv = 0; kv < BKV; ++kv) {
        int k = f_kg * BKV + kv;
        TI tmp[4];
        if (nok) {
          const TI *src = Bg + ((int64_t)tap_ * p.Ci + c0_ + k) * p.Co + n;
          __builtin_memcpy(tmp, src, 4 * sizeof(TI));
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v) tmp[v] = (TI)0.0f;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) rb[v * BKV + kv] = tmp[v];
      }
    } else {
#pragma unroll
      for (int i = 0; i < BNV; ++i) {
        int n = n0 + b_row + i * B_ROWS_PER_PASS;
        uint4 v = make_uint4(0, 0, 0, 0);
        if (n < N) v = *(const uint4 *)(Bg + ((int64_t)tap_ * p.Ci + n) * p.Co + c0_ + b_kv);
        __builtin_memcpy(&rb[i * VEC], &v, 16);
      }
    }
  };

  auto store_b = [&](int buf) {
    TI *dst = Bs[buf];
    if (PASS == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        __builtin_memcpy(&dst[(f_ng * 4 + v) * LDR + f_kg * BKV], &rb[v * BKV], BKV * sizeof(TI));
    } else {
#pragma unroll
      for (int i = 0; i < BNV; ++i)
        __builtin_memcpy(&dst[(b_row + i * B_ROWS_PER_PASS) * LDR + b_kv], &rb[i * VEC], 16);
    }
  };

  auto advance = [&]() {
    c0 += BK;
    if (c0 == C) {
      c0 = 0;
      ++tap;
      if (++te == p.kd) {
        te = 0;
        if (++tc == p.kw) {
          tc = 0;
          ++ta;
        }
      }
    }
  };

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  // prologue: step 0 into buffers 0
  load_a(0, 0, 0, 0, 0);
  load_b(0, 0);
  store_b(0);
  __syncthreads();

  for (int ks = 0; ks < nk; ++ks) {
    const int cur = ks & 1;
    const bool more = ks + 1 < nk;
    if (more) {
      advance();
      load_a(cur ^ 1, ta, tc, te, c0);
      load_b(tap, c0);
    }
    // MFMAs on the current step
    const TI *bs = Bs[cur];
    uint4 fb[FN][2];
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const uint4 *src = (const uint4 *)&bs[(wn * FN * 32 + j * 32 + frow) * LDR + fkg * LK];
      fb[j][0] = src[0];
      fb[j][1] = src[1];
    }
    if constexpr (sizeof(TI) == 2) {
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(ra[cur][i][s], fb[j][s], acc[i][j]);
    } else {
      float af[FM][8], bf[FN][8];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        __builtin_memcpy(&af[i][0], &ra[cur][i][0], 16);
        __builtin_memcpy(&af[i][4], &ra[cur][i][1], 16);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        __builtin_memcpy(&bf[j][0], &fb[j][0], 16);
        __builtin_memcpy(&bf[j][4], &fb[j][1], 16);
      }
#pragma unroll
      for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[i][s], bf[j][s], acc[i][j], 0, 0, 0);
    }
    if (more) store_b(cur ^ 1);
    __syncthreads();
  }

  // ---- epilogue: lane holds column (lane % 32) of each fragment ----
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * FN * 32 + j * 32 + frow;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int64_t m = m0 + wm * FM * 32 + i * 32 + 8 * (r >> 2) + 4 * fkg + (r & 3);
        if (m < M) ((TO *)p.c)[m * p.ldc + n] = (TO)acc[i][j][r];
      }
    }
}

template <int PASS, typename TI, typename TO, int FM, int FN, int WGM, int WGN>
static void launch_fast(const ConvParams &p, hipStream_t st) {
  constexpr int BM = WGM * FM * 32, BN = WGN * FN * 32;
  int64_t tiles = ((p.M + BM - 1) / BM) * (int64_t)((p.N + BN - 1) / BN);
  hipLaunchKernelGGL((conv_fast_kernel<PASS, TI, TO, FM, FN, WGM, WGN>), dim3((unsigned)tiles), dim3(256), 0,
                     st, p);
}

// Tile shapes: 0 -> 256 x 64, 1 -> 256 x 128, 2 -> 128 x 256, 3 -> 128 x 128.
template <int PASS, typename TI, typename TO>
void run_fast(const ConvParams &p, int shape, hipStream_t st) {
  switch (shape) {
    case 0: launch_fast<PASS, TI, TO, 2, 2, 4, 1>(p, st); break;
    case 1: launch_fast<PASS, TI, TO, 2, 4, 4, 1>(p, st); break;
    case 2: launch_fast<PASS, TI, TO, 2, 4, 2, 2>(p, st); break;
    default: launch_fast<PASS, TI, TO, 2, 2, 2, 2>(p, st); break;
  }
}

int fast_bk(int elem_bytes) { return elem_bytes == 4 ? 16 : 32; }

#define OWL_FAST(PASS, TI, TO) template void run_fast<PASS, TI, TO>(const ConvParams &, int, hipStream_t);
OWL_FAST(0, bf16_t, bf16_t)
OWL_FAST(0, bf16_t, float)
OWL_FAST(1, bf16_t, bf16_t)
OWL_FAST(1, bf16_t, float)
OWL_FAST(0, f16_t, f16_t)
OWL_FAST(0, f16_t, float)
OWL_FAST(1, f16_t, f16_t)
OWL_FAST(1, f16_t, float)
OWL_FAST(0, float, float)
OWL_FAST(1, float, float)
#undef OWL_FAST

}  // namespace owl_amd

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
#include <cstdlib>

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


// Raw buffer resource over [base, base + bytes): loads at offsets >= bytes
// return zeros, which is how padding taps read their zeros.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void *base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), (short)0, (int)bytes, 0x00020000);
}
constexpr uint32_t OOB = 0x80000000u;  // >= every num_records used here (tensors < 2 GiB)

typedef unsigned int v4u32 __attribute__((ext_vector_type(4)));
typedef unsigned int v2u32 __attribute__((ext_vector_type(2)));
// Cache policy of the 16-bit kernels' result stores: non-temporal (nt).  A
// layer's output is written once and read by the next pass long after; the
// streaming hint keeps the end-of-tile write bursts from allocating in L2
// (halo forward 40.8 -> 38.4 us, backward input 39.4 -> 37.5 us on a
// 64x56x56x64 -> 128 layer, same bits).
constexpr int ST_NT = 2;
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint4 bload(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  return __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, 0, 0));
}

// PASS 0: forward.  PASS 1: backward input with unit strides.  PD: how many
// K-steps of A fragments are in flight ahead of the MFMAs (register ring of
// PD + 1 slots; the loop is unrolled by PD + 1 so slots are static).
template <int PASS, typename TI, typename TO, int FM, int FN, int WGM, int WGN, int PD>
__global__ __launch_bounds__(256, 2) void conv_fast_kernel(ConvParams p) {
  typedef FastMfma<TI> MF;
  constexpr int BK = MF::BK;
  constexpr int LK = BK / 2;                      // k values per lane per fragment per K-step (32 bytes)
  constexpr int BM = WGM * FM * 32, BN = WGN * FN * 32;
  constexpr int LDR = BK + 16 / (int)sizeof(TI);  // LDS row, padded by 16 bytes
  constexpr int BELEMS = BK * BN / 256;           // B elements per thread per K-step
  constexpr int ES = (int)sizeof(TI);
  static_assert(WGM * WGN == 4, "four waves");
  static_assert(LK * sizeof(TI) == 32, "32-byte fragments");

  __shared__ __attribute__((aligned(16))) TI Bs[2][BN * LDR];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  const int N = p.N;
  const int M = (int)p.M;
  const int n_tiles = (N + BN - 1) / BN;
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int m0 = (bid / n_tiles) * BM;
  const int n0 = (bid % n_tiles) * BN;

  const int C = PASS == 0 ? p.Ci : p.Co;  // contracted channels per tap
  const int T = p.kh * p.kw * p.kd;
  const int csteps = C / BK;
  const int nk = T * csteps;

  // A source: x (forward) or dy (backward input), addressed in bytes
  const uint32_t a_bytes = PASS == 0 ? (uint32_t)((int64_t)p.B * p.H * p.W * p.D * p.Ci * ES)
                                     : (uint32_t)((int64_t)p.B * p.Ho * p.Wo * p.Do * p.Co * ES);
  const __amdgpu_buffer_rsrc_t arsrc = make_rsrc(p.a, a_bytes);
  // the kernel tile through buffer loads: columns past N and taps past the
  // last one read zeros from out-of-range offsets, with no per-lane branch
  const __amdgpu_buffer_rsrc_t brsrc = make_rsrc(p.b, (uint32_t)((int64_t)T * p.Ci * p.Co * ES));

  // ---- per-lane A rows (one per M fragment) ----
  // row_off: element offset of the row's window origin (may be negative for
  // padded origins; only in-range taps are ever dereferenced).
  const int frow = lane & 31, fkg = lane >> 5;
  int row_off[FM], q0[FM], q1[FM], q2[FM];
#pragma unroll
  for (int f = 0; f < FM; ++f) {
    int m = m0 + wm * FM * 32 + f * 32 + frow;
    const bool ok = m < M;
    if (!ok) m = 0;
    int t = m;
    if (PASS == 0) {
      int l = t % p.Do; t /= p.Do;
      int j = t % p.Wo; t /= p.Wo;
      int i = t % p.Ho;
      int b = t / p.Ho;
      q0[f] = i * p.sh - p.ph;
      q1[f] = j * p.sw - p.pw;
      q2[f] = l * p.sd - p.pd;
      row_off[f] = ((b * p.H + q0[f]) * p.W + q1[f]) * p.D * p.Ci + q2[f] * p.Ci + fkg * LK;
    } else {
      int d = t % p.D; t /= p.D;
      int w = t % p.W; t /= p.W;
      int h = t % p.H;
      int b = t / p.H;
      q0[f] = h + p.ph;
      q1[f] = w + p.pw;
      q2[f] = d + p.pd;
      row_off[f] = ((b * p.Ho + q0[f]) * p.Wo + q1[f]) * p.Do * p.Co + q2[f] * p.Co + fkg * LK;
    }
    if (!ok) q0[f] = -(1 << 28);  // never in range
  }

  // ---- B staging thread layout ----
  // forward: B[k][n] = w[(t*Ci + c0 + k)*Co + n] is n-contiguous: each thread
  //   loads BKV k-rows of 4 consecutive n and writes 4 n-rows of BKV k.
  // backward input: B[k][n] = w[(t*Ci + n)*Co + c0 + k] is k-contiguous:
  //   each thread loads 16-byte vectors along k.
  constexpr int BKV = BELEMS / 4;
  constexpr int VEC = 16 / ES;
  constexpr int BNV = BELEMS / VEC;
  const int f_ng = tid % (BN / 4), f_kg = tid / (BN / 4);
  const int b_kv = (tid % (BK / VEC)) * VEC, b_row = tid / (BK / VEC);
  constexpr int B_ROWS_PER_PASS = 256 / (BK / VEC);

  constexpr int NS = PD + 1;
  TI rb[NS][BELEMS];
  uint4 ra[NS][FM][2];

  // B walk (one step ahead) and A walk (PD steps ahead)
  int tap = 0, c0 = 0;
  int la = 0, lc = 0, le = 0, lc0 = 0;
  uint32_t voff[FM];  // byte offset of the current A tap for each row, or OOB

  auto set_tap = [&]() __attribute__((always_inline)) {
    const int dh = la * p.rh, dw = lc * p.rw, dd = le * p.rd;
#pragma unroll
    for (int f = 0; f < FM; ++f) {
      int h, w, d, off;
      bool ok;
      if (PASS == 0) {
        h = q0[f] + dh; w = q1[f] + dw; d = q2[f] + dd;
        ok = (unsigned)h < (unsigned)p.H && (unsigned)w < (unsigned)p.W && (unsigned)d < (unsigned)p.D;
        off = row_off[f] + ((dh * p.W + dw) * p.D + dd) * p.Ci;
      } else {
        h = q0[f] - dh; w = q1[f] - dw; d = q2[f] - dd;
        ok = (unsigned)h < (unsigned)p.Ho && (unsigned)w < (unsigned)p.Wo && (unsigned)d < (unsigned)p.Do;
        off = row_off[f] - ((dh * p.Wo + dw) * p.Do + dd) * p.Co;
      }
      voff[f] = ok ? (uint32_t)off * ES : OOB;
    }
  };

  auto load_a = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int f = 0; f < FM; ++f) {
      const uint32_t o = voff[f] + (uint32_t)(lc0 * ES);
      ra[buf][f][0] = bload(arsrc, o);
      ra[buf][f][1] = bload(arsrc, o + 16);
    }
  };

  auto advance_a = [&]() __attribute__((always_inline)) {
    lc0 += BK;
    if (lc0 == C) {
      lc0 = 0;
      if (++le == p.kd) {
        le = 0;
        if (++lc == p.kw) {
          lc = 0;
          ++la;
        }
      }
      if (la < p.kh) set_tap();
    }
  };

  auto load_b = [&](int slot, int tap_, int c0_) __attribute__((always_inline)) {
    TI *rbs = rb[slot];
    const bool tok = tap_ < T;
    if (PASS == 0) {
      const int n = n0 + f_ng * 4;
      const bool nok = tok && n < N;
      const uint32_t src = (uint32_t)((tap_ * p.Ci + c0_ + f_kg * BKV) * p.Co + n) * ES;
#pragma unroll
      for (int kv = 0; kv < BKV; ++kv) {
        TI tmp[4];
        if constexpr (sizeof(TI) == 4) {
          const uint4 v = bload(brsrc, nok ? src + (uint32_t)(kv * p.Co * ES) : OOB);
          __builtin_memcpy(tmp, &v, 16);
        } else {
          const uint2 v = __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(
                                                        brsrc, (int)(nok ? src + (uint32_t)(kv * p.Co * ES) : OOB), 0, 0));
          __builtin_memcpy(tmp, &v, 8);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) rbs[v * BKV + kv] = tmp[v];
      }
    } else {
#pragma unroll
      for (int i = 0; i < BNV; ++i) {
        const int n = n0 + b_row + i * B_ROWS_PER_PASS;
        const uint4 v = bload(brsrc, tok && n < N ? (uint32_t)((tap_ * p.Ci + n) * p.Co + c0_ + b_kv) * ES : OOB);
        __builtin_memcpy(&rbs[i * VEC], &v, 16);
      }
    }
  };

  auto store_b = [&](int buf, int slot) __attribute__((always_inline)) {
    TI *dst = Bs[buf];
    const TI *rbs = rb[slot];
    if (PASS == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v)
        __builtin_memcpy(&dst[(f_ng * 4 + v) * LDR + f_kg * BKV], &rbs[v * BKV], BKV * sizeof(TI));
    } else {
#pragma unroll
      for (int i = 0; i < BNV; ++i)
        __builtin_memcpy(&dst[(b_row + i * B_ROWS_PER_PASS) * LDR + b_kv], &rbs[i * VEC], 16);
    }
  };

  auto advance_b = [&]() __attribute__((always_inline)) {
    c0 += BK;
    if (c0 == C) {
      c0 = 0;
      ++tap;
    }
  };

  f32x16 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  // prologue: A and B for steps 0 .. PD-1 into ring slots 0 .. PD-1; the B
  // tile of step 0 goes to LDS now, step s+1's at the end of iteration s.
  // Loads past the last step are issued all the same (A from the last tap's
  // rows, B from out-of-range offsets), so every step issues the same loads
  // and the compiler's waits keep PD steps of them in flight: with the loads
  // and the kernel-tile store conditional, it waited for every load in
  // flight at each step.
  set_tap();
#pragma unroll
  for (int u = 0; u < PD; ++u) {
    load_a(u);
    advance_a();
    load_b(u, tap, c0);
    advance_b();
  }
  store_b(0, 0);
  __syncthreads();

  const int fb_off = (wn * FN * 32 + frow) * LDR + fkg * LK;
  // step ks from ring slot u = ks % NS (a constant once unrolled)
  auto step = [&](int u, int ks) __attribute__((always_inline)) {
        const int bcur = ks & 1;
        load_a((u + PD) % NS);
        advance_a();
        load_b((u + PD) % NS, tap, c0);
        advance_b();
        const TI *bs = Bs[bcur] + fb_off;
        uint4 fb[FN][2];
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const uint4 *src = (const uint4 *)&bs[j * 32 * LDR];
          fb[j][0] = src[0];
          fb[j][1] = src[1];
        }
        if constexpr (sizeof(TI) == 2) {
#pragma unroll
          for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(ra[u][i][s2], fb[j][s2], acc[i][j]);
        } else {
          float af[FM][8], bf[FN][8];
#pragma unroll
          for (int i = 0; i < FM; ++i) {
            __builtin_memcpy(&af[i][0], &ra[u][i][0], 16);
            __builtin_memcpy(&af[i][4], &ra[u][i][1], 16);
          }
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            __builtin_memcpy(&bf[j][0], &fb[j][0], 16);
            __builtin_memcpy(&bf[j][4], &fb[j][1], 16);
          }
#pragma unroll
          for (int s2 = 0; s2 < 8; ++s2)
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
              for (int j = 0; j < FN; ++j)
                acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(af[i][s2], bf[j][s2], acc[i][j], 0, 0, 0);
        }
        // after the last step this fills a buffer nobody reads again
        store_b(bcur ^ 1, (u + 1) % NS);
        __syncthreads();
  };
  int kb = 0;
  for (; kb + NS <= nk; kb += NS) {
#pragma unroll
    for (int u = 0; u < NS; ++u) step(u, kb + u);
  }
#pragma unroll
  for (int u = 0; u < NS - 1; ++u)
    if (kb + u < nk) step(u, kb + u);

  // ---- epilogue: lane holds column (lane % 32) of each fragment ----
  // Raw buffer stores: the uniform row offset goes in soffset, the lane's
  // column (and its 4-row group) in a 32-bit voffset; rows past M get an
  // out-of-range voffset and the hardware drops the store.
  const uint32_t c_bytes = (uint32_t)((int64_t)M * p.ldc * sizeof(TO));
  const __amdgpu_buffer_rsrc_t crsrc = make_rsrc(p.c, c_bytes);
  const __amdgpu_buffer_rsrc_t prsrc = make_rsrc(p.pre ? p.pre : p.c, c_bytes);
  const bool full = m0 + BM <= M;
  // fused bias and activation of the forward (none for backward input)
  const bool epi = PASS == 0 && (p.bias || p.act);
  auto put = [&](__amdgpu_buffer_rsrc_t r, float v, uint32_t vo, int so) __attribute__((always_inline)) {
    if constexpr (sizeof(TO) == 2) {
      const TO h = (TO)v;
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), r, (int)vo, so, 0);
    } else {
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)vo, so, 0);
    }
  };
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * FN * 32 + j * 32 + frow;
      const int mrow = m0 + wm * FM * 32 + i * 32 + 4 * fkg;  // + 8*(r>>2) + (r&3)
      const uint32_t lane_off = (uint32_t)(mrow * p.ldc + n) * (uint32_t)sizeof(TO);
      const float bn = (epi && p.bias && n < N) ? ((const float *)p.bias)[n] : 0.0f;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int dr = 8 * (r >> 2) + (r & 3);
        const int so = dr * p.ldc * (int)sizeof(TO);
        uint32_t vo = lane_off;
        if (!full && mrow + dr >= M) vo = OOB;
        if (n >= N) vo = OOB;
        float v = acc[i][j][r];
        if (epi) {
          if (p.bias) v += bn;
          if (p.pre) put(prsrc, v, vo, so);
          if (p.act == OWL_AMD_ACT_RELU) v = fmaxf(v, 0.0f);
        }
        put(crsrc, v, vo, so);
      }
    }
}

// ---------------------------------------------------------------------------
// Backward kernel (weight gradient) for 16-bit storage:
//   dw[m = (tap, ci)][n = co] = sum over output pixels k of
//       x[pixel k seen through tap][ci] * dy[k][co]
// The contraction runs over pixels, which are strided in NHWC for both
// operands, so the tiles are copied to LDS in their natural layout
// ([pixel][channel], 16-byte vectors, no register transposes) and the MFMA
// operands are read with gfx950's transposing LDS load: in each group of 16
// lanes, ds_read_b64_tr_b16 gives lane i halfword (i % 4) of the 8 bytes
// addressed by lane 4e + i/4 as its element e.  Pointing lane j at
// (pixel k0 + j/4, channels c0 + 4 (j % 4) ..) therefore hands lane i the 4
// consecutive pixels k0..k0+3 of channel c0 + i -- one MFMA operand quarter.
// Rows are padded by 64 bytes so the 32 lanes of a read phase (4 pixels x 64
// bytes) touch every bank once.
// Tile: 128 (tap, ci) x 128 co x 32 pixels, 4 waves of 64 x 64; split-K
// over pixels as in the LDS-staged kernel.
// ---------------------------------------------------------------------------
typedef short v4s16 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) v4s16 lds_v4s16;

template <typename TI> __device__ __forceinline__ f32x16 mfma16(const v4s16 &a0, const v4s16 &a1, const v4s16 &b0,
                                                               const v4s16 &b1, f32x16 c);
template <> __device__ __forceinline__ f32x16 mfma16<bf16_t>(const v4s16 &a0, const v4s16 &a1, const v4s16 &b0,
                                                             const v4s16 &b1, f32x16 c) {
  typedef short v8s16 __attribute__((ext_vector_type(8)));
  v8s16 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  v8s16 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
template <> __device__ __forceinline__ f32x16 mfma16<f16_t>(const v4s16 &a0, const v4s16 &a1, const v4s16 &b0,
                                                            const v4s16 &b1, f32x16 c) {
  typedef short v8s16 __attribute__((ext_vector_type(8)));
  v8s16 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
  v8s16 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                0);
}

// FMW: 32-row fragments per wave along M; the tile is BM = 64 FMW rows
// (2 x 2 waves of 32 FMW x 64).  TI float: the FP32 weight gradient.  A
// 32x32x2 FP32 MFMA takes one k value per lane and operand, which a plain
// 4-byte LDS read of the natural [pixel][channel] rows supplies (lanes 0-31
// one pixel's 32 channels, lanes 32-63 the next pixel's), so nothing is
// transposed; a K-step is 16 pixels (the same bytes per row as 32 16-bit
// pixels) and rows are padded to 128 bytes past a multiple of 256, which
// puts the two half-waves' reads on disjoint banks.  D1: a 2-D problem (D = Do = kd = 1, no
// depth pad), whose depth digit is always 0: the loader skips its bounds
// check and carry.  KG: groups of four waves per workgroup.  With KG = 2 the
// two groups sum the two halves of the split's pixels into the same tile,
// each with its own LDS stages, and add their sums through LDS at the end
// (first half + second half): half as many FP32 partials for the reduce, at
// one 512-thread workgroup per CU instead of two of 256.
// FNW: 32-column fragments per wave along N (2: tiles 128 output channels
// wide; 1: 64 wide, for layers with at most 64, which would leave half of a
// 128-wide tile's MFMA work on padding).
template <typename TI, typename TO, int FMW, bool D1, int KG, int FNW = 2>
__global__ __launch_bounds__(256 * KG, KG == 1 ? 2 : 1) void conv_wgrad_kernel(ConvParams p) {
  // BK 64 (80 KB of LDS) ran slower than BK 32: one workgroup per CU
  constexpr bool F32 = sizeof(TI) == 4;
  constexpr int ES = (int)sizeof(TI), VEC = 16 / ES;  // bytes per value, values per 16-byte chunk
  constexpr int BM = 64 * FMW, BN = 64 * FNW, BK = F32 ? 16 : 32, KP = 1;  // KP pixels per loader thread
  constexpr int G = 256 / BK;       // loader threads per pixel
  constexpr int QA = BM / (VEC * G);  // A chunks per loader thread and pixel
  constexpr int QB = BN / (VEC * G);  // B chunks per loader thread and pixel
  // LDS row strides in bytes: 16-bit rows + 64 bytes pad; FP32 rows + 128
  constexpr int RSA = BM * ES + (F32 ? 128 : 64);
  constexpr int RSB = BN * ES + (F32 ? 128 : 64);
  constexpr int STAGE = BK * (RSA + RSB);  // A rows, then B rows
  static_assert(QA * VEC * G == BM && QB * VEC * G == BN && QB == FNW, "loader covers the tile");
  // 40 KB (FMW 2), 48 KB (FMW 3) per group
  __shared__ __attribute__((aligned(16))) char lds_all[KG][2][STAGE];
  // tid / wave: within the group
  const int tid = threadIdx.x & 255, grp = threadIdx.x >> 8, lane = tid & 63, wave = tid >> 6;
  char(*lds)[STAGE] = lds_all[grp];
  const int wm = wave >> 1, wn = wave & 1;
  const int M = (int)p.M, N = p.N;
  const int n_tiles = (N + BN - 1) / BN;
  const int mn_tiles = ((M + BM - 1) / BM) * n_tiles;
  // 1-D grid of mn_tiles x split: the tiles of one split read the same
  // pixels, so consecutive tile indices go to one XCD (xcd_swizzle) and
  // share its L2 (64x56x56x64 -> 128 BF16 layer 39.4 -> 37.5 us against
  // the tiles of a split spread over the XCDs in launch order)
  const int lid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tile = lid % mn_tiles, split_idx = lid / mn_tiles;
  const int m0 = (tile / n_tiles) * BM, n0 = (tile % n_tiles) * BN;
  // this group's pixels: the whole split (KG 1) or one half of it (KG 2,
  // k_per_split a multiple of 2 BK); both groups run the first half's step
  // count, the second half's loads past its end read zeros
  const int64_t sbeg = (int64_t)split_idx * p.k_per_split;
  const int64_t send = sbeg + p.k_per_split < p.K ? sbeg + p.k_per_split : p.K;
  const int64_t half = p.k_per_split / KG;
  const int64_t kbeg = sbeg + grp * half;
  int64_t kend = kbeg + half < send ? kbeg + half : send;
  if (kend < kbeg) kend = kbeg;
  const int64_t len0 = half < send - sbeg ? half : send - sbeg;
  const int nk = (int)((len0 + BK - 1) / BK);

  const uint32_t x_bytes = (uint32_t)((int64_t)p.B * p.H * p.W * p.D * p.Ci * ES);
  const uint32_t g_bytes = (uint32_t)((int64_t)p.K * p.Co * ES);
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(p.a, x_bytes);
  const __amdgpu_buffer_rsrc_t gr = make_rsrc(p.b, g_bytes);

  // ---- loader assignment: per K-step the A tile (BK pixels x BM rows m)
  // is G QA 16-byte chunks per pixel and the B tile (BK pixels x 128
  // columns n) 2 G.  Thread t loads pixel kp = t / G of the step, and of it
  // the A chunks c8 + G q (q < QA) and the B chunks c8, c8 + G.
  const int c8 = tid % G, kp = tid / G;
  // A chunk q covers rows m0 + VEC (c8 + G q) .. + VEC - 1, all in one tap (Ci % VEC == 0):
  // tap window offset (tdh, tdw, tdd) and the chunk's element offset from
  // the pixel's window origin
  int tdh[QA], tdw[QA], tdd[QA], toff[QA];
  bool aok[QA];
#pragma unroll
  for (int q = 0; q < QA; ++q) {
    const int am = m0 + VEC * (c8 + G * q);
    aok[q] = am < M;
    const int tap = aok[q] ? am / p.Ci : 0;
    const int ci = aok[q] ? am - tap * p.Ci : 0;
    const int kwd = p.kw * p.kd;
    const int ta = tap / kwd, r = tap - (tap / kwd) * kwd;
    const int tc = r / p.kd, te = r - (r / p.kd) * p.kd;
    tdh[q] = ta * p.rh - p.ph;
    tdw[q] = tc * p.rw - p.pw;
    tdd[q] = te * p.rd - p.pd;
    toff[q] = ((tdh[q] * p.W + tdw[q]) * p.D + tdd[q]) * p.Ci + ci;
  }
  const int bn = n0 + VEC * c8;
  const bool bok0 = bn < N, bok1 = FNW == 2 && bn + 64 < N;

  // Pixel positions, advanced by BK per K-step without division: the step is
  // split once into digits of the radix (Do, Wo, Ho); the scaled digits
  // pld = l sd, pjw = j sw, pih = i sh carry into each other, and the
  // element offset pbase of the pixel's window origin follows with a fixed
  // correction per carry.
  const int kend32 = (int)kend;
  int pk[KP], pld[KP], pjw[KP], pih[KP], pbase[KP], gbase[KP];
#pragma unroll
  for (int h = 0; h < KP; ++h) {
    int t = (int)kbeg + kp + BK * h;
    pk[h] = t;
    const int l = t % p.Do; t /= p.Do;
    const int j = t % p.Wo; t /= p.Wo;
    const int i = t % p.Ho, b = t / p.Ho;
    pld[h] = l * p.sd; pjw[h] = j * p.sw; pih[h] = i * p.sh;
    pbase[h] = (((b * p.H + pih[h]) * p.W + pjw[h]) * p.D + pld[h]) * p.Ci;
    gbase[h] = pk[h] * p.Co + bn;  // dy element offset of chunk 0
  }
  int inc_l, inc_j, inc_i, base_step;
  const int dx_l = p.sd * p.Ci, dx_j = p.sw * p.D * p.Ci, dx_i = p.sh * p.W * p.D * p.Ci;
  const int dx_b = p.H * p.W * p.D * p.Ci;
  {
    int t = BK;
    const int sl = t % p.Do; t /= p.Do;
    const int sj = t % p.Wo; t /= p.Wo;
    const int si = t % p.Ho, sb = t / p.Ho;
    inc_l = sl * p.sd; inc_j = sj * p.sw; inc_i = si * p.sh;
    base_step = sl * dx_l + sj * dx_j + si * dx_i + sb * dx_b;
  }
  const int lim_l = p.Do * p.sd, lim_j = p.Wo * p.sw, lim_i = p.Ho * p.sh;
  const int corr_l = dx_j - p.Do * dx_l, corr_j = dx_i - p.Wo * dx_j, corr_i = dx_b - p.Ho * dx_i;
  const int gstep = BK * p.Co;

  // register ring: loads run two K-steps ahead of the LDS stage they fill
  typedef uint4 Stage[KP][QA + QB];  // per pixel: A chunks 0 .. QA-1, then B chunks 0 .. QB-1
  Stage rs0, rs1;
  auto load = [&](Stage &rg) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < KP; ++h) {
      // branch-free (bitwise predicates, one select per load)
      const bool kok = pk[h] < kend32;
#pragma unroll
      for (int q = 0; q < QA; ++q) {
        const int hh = pih[h] + tdh[q], ww = pjw[h] + tdw[q], dd = pld[h] + tdd[q];
        const bool ok = kok & aok[q] & ((unsigned)hh < (unsigned)p.H) & ((unsigned)ww < (unsigned)p.W) &
                        (D1 || ((unsigned)dd < (unsigned)p.D));
        const uint32_t off = (uint32_t)(pbase[h] + toff[q]) * (uint32_t)ES;
        rg[h][q] = bload(xr, ok ? off : OOB);
      }
      const uint32_t goff = (uint32_t)gbase[h] * (uint32_t)ES;
      rg[h][QA] = bload(gr, (kok & bok0) ? goff : OOB);
      if constexpr (QB == 2) rg[h][QA + 1] = bload(gr, (kok & bok1) ? goff + 64u * ES : OOB);
      // advance to the pixel BK further on
      pk[h] += BK;
      gbase[h] += gstep;
      pbase[h] += base_step;
      int c = 0;
      if constexpr (!D1) {
        pld[h] += inc_l;
        c = pld[h] >= lim_l;
        if (c) {
          pld[h] -= lim_l;
          pbase[h] += corr_l;
        }
      }
      pjw[h] += inc_j + (c ? p.sw : 0);
      c = pjw[h] >= lim_j;
      if (c) {
        pjw[h] -= lim_j;
        pbase[h] += corr_j;
      }
      pih[h] += inc_i + (c ? p.sh : 0);
      c = pih[h] >= lim_i;
      if (c) {
        pih[h] -= lim_i;
        pbase[h] += corr_i;
      }
    }
  };
  auto store = [&](const Stage &rg, int st) __attribute__((always_inline)) {
#pragma unroll
    for (int h = 0; h < KP; ++h) {
      const int row = kp + BK * h;
#pragma unroll
      for (int q = 0; q < QA; ++q) *(uint4 *)(&lds[st][row * RSA + (c8 + G * q) * 16]) = rg[h][q];
#pragma unroll
      for (int q = 0; q < QB; ++q) *(uint4 *)(&lds[st][BK * RSA + row * RSB + (c8 + G * q) * 16]) = rg[h][QA + q];
    }
  };

  f32x16 acc[FMW][FNW];
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < FNW; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  // per-lane byte offset of its transposed-read source within a tile:
  // pixel row (lane%16)/4 (+8 for the upper k-group, +4 for the second half),
  // column base (lane/16 %2)*16 + 4*(lane%4) within the fragment's 32 columns
  const int j16 = lane & 15;
  const int tr_row = (j16 >> 2) + 8 * (lane >> 5), tr_col = 2 * (16 * ((lane >> 4) & 1) + 4 * (j16 & 3));
  const int tr_a = tr_row * RSA + tr_col, tr_b = BK * RSA + tr_row * RSB + tr_col;

  // one K-step: MFMAs from LDS stage ks & 1; rn holds step ks + 1 (stored
  // to the other stage) and is refilled with step ks + 3.  Loads past the
  // split's end read zeros, so every step issues the same loads.
  auto kstep = [&](int ks, Stage &rn) __attribute__((always_inline)) {
    const int cur = ks & 1;
    const char *Ls = lds[cur];
    if constexpr (F32) {
      // pixel pair kk: lanes 0-31 read pixel 2 kk, lanes 32-63 pixel 2 kk + 1,
      // each lane its fragment column lane % 32 -- the 32x32x2 MFMA's k = lane / 32
      constexpr int KK = BK / 2;
      float a[KK][FMW], b[KK][FNW];
      const char *la = Ls + (lane >> 5) * RSA + 4 * (wm * 32 * FMW + (lane & 31));
      const char *lb = Ls + BK * RSA + (lane >> 5) * RSB + 4 * (wn * 32 * FNW + (lane & 31));
#pragma unroll
      for (int kk = 0; kk < KK; ++kk) {
#pragma unroll
        for (int f = 0; f < FMW; ++f) a[kk][f] = *(const float *)(la + 2 * kk * RSA + 128 * f);
#pragma unroll
        for (int f = 0; f < FNW; ++f) b[kk][f] = *(const float *)(lb + 2 * kk * RSB + 128 * f);
      }
#pragma unroll
      for (int kk = 0; kk < KK; ++kk)
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int j = 0; j < FNW; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kk][i], b[kk][j], acc[i][j], 0, 0, 0);
      // after the last step this fills a stage nobody reads again
      store(rn, cur ^ 1);
      load(rn);
      __builtin_amdgcn_sched_group_barrier(0x100, KK * (FMW + FNW), 0);
      __builtin_amdgcn_sched_group_barrier(0x008, KK * FNW * FMW, 0);
      __syncthreads();
      return;
    }
    // every 16-pixel quarter's fragments in their own registers, so all the
    // transposed reads can be in flight before the first MFMA
    constexpr int NQ = BK / 16;
    v4s16 a[NQ][FMW][2], b[NQ][FNW][2];  // [quarter][fragment][half]
#pragma unroll
    for (int s = 0; s < NQ; ++s)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
#pragma unroll
        for (int f = 0; f < FMW; ++f)
          a[s][f][hf] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
              (lds_v4s16 *)(Ls + tr_a + (16 * s + 4 * hf) * RSA + 2 * (wm * 32 * FMW + f * 32)));
#pragma unroll
        for (int f = 0; f < FNW; ++f)
          b[s][f][hf] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
              (lds_v4s16 *)(Ls + tr_b + (16 * s + 4 * hf) * RSB + 2 * (wn * 32 * FNW + f * 32)));
      }
#pragma unroll
    for (int s = 0; s < NQ; ++s)
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int j = 0; j < FNW; ++j)
          acc[i][j] = mfma16<TI>(a[s][i][0], a[s][i][1], b[s][j][0], b[s][j][1], acc[i][j]);
    // after the last step this fills a stage nobody reads again
    store(rn, cur ^ 1);
    load(rn);
    // scheduling: the LDS reads as one group ahead of the MFMAs (an
    // MFMA-by-MFMA interleave with the loads and writes ran slower, and so
    // did the second quarter's reads spread between the first quarter's MFMAs)
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * NQ * (FMW + FNW), 0);
    __builtin_amdgcn_sched_group_barrier(0x008, NQ * FMW * FNW, 0);
    __syncthreads();
  };

  if (nk > 0) {
    load(rs0);  // step 0
    load(rs1);  // step 1
    store(rs0, 0);
    load(rs0);  // step 2
    __syncthreads();
    // Unconditional pairs of steps, then an odd last step: every path
    // through the loop issues the same loads, so the compiler's waits keep
    // two steps of loads in flight.  A conditional second step (if ks + 1 <
    // nk) made it wait for every load in flight at each step (BF16
    // 64x56x56x64 -> 128 layer 37.4 -> 35.9 us, 128x28x28x256 -> 256 127.5
    // -> 121.4 us with the pairs).
    int ks = 0;
    for (; ks + 1 < nk; ks += 2) {
      kstep(ks, rs1);
      kstep(ks + 1, rs0);
    }
    if (ks < nk) kstep(ks, rs1);
  }

  // ---- KG 2: add the halves.  With two fragment columns group 0 keeps the
  // j = 0 fragments and group 1 the j = 1 ones, each handing the other half
  // over through LDS; with one, group 1 hands over everything and group 0
  // stores the tile.  Every sum is (first half) + (second half).
  if constexpr (KG == 2) {
    __syncthreads();
    // regions of FMW * 16 * 256 floats: one per group (two fragment
    // columns), or one that group 1 writes and group 0 reads (one column)
    float *X = (float *)&lds_all[0][0][0];
    constexpr int REG = FMW * 16 * 256;
    static_assert(FNW * REG * 4 <= KG * 2 * STAGE, "exchange fits the stages");
    float *mine = X + (FNW == 2 ? grp : 0) * REG;
    const float *theirs = X + (FNW == 2 ? 1 - grp : 0) * REG;
    if constexpr (FNW == 2) {
      if (grp == 0) {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) mine[(i * 16 + r) * 256 + tid] = acc[i][1][r];
      } else {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) mine[(i * 16 + r) * 256 + tid] = acc[i][0][r];
      }
      __syncthreads();
      if (grp == 0) {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[i][0][r] = acc[i][0][r] + theirs[(i * 16 + r) * 256 + tid];
      } else {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[i][1][r] = theirs[(i * 16 + r) * 256 + tid] + acc[i][1][r];
      }
    } else {
      if (grp == 1) {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) mine[(i * 16 + r) * 256 + tid] = acc[i][0][r];
      }
      __syncthreads();
      if (grp == 0) {
#pragma unroll
        for (int i = 0; i < FMW; ++i)
#pragma unroll
          for (int r = 0; r < 16; ++r) acc[i][0][r] = acc[i][0][r] + theirs[(i * 16 + r) * 256 + tid];
      }
    }
  }

  // ---- epilogue (same layout as the other MFMA kernels) ----
  const int frow = lane & 31, fkg = lane >> 5;
  const bool partial = p.split > 1;
  const uint32_t c_bytes = partial ? (uint32_t)((int64_t)M * N * 4) : (uint32_t)((int64_t)M * N * sizeof(TO));
  char *cbase = (char *)p.c + (partial ? (int64_t)split_idx * M * N * 4 : 0);
  const __amdgpu_buffer_rsrc_t cr = make_rsrc(cbase, c_bytes);
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < FNW; ++j) {
      if (KG == 2 && (FNW == 2 ? j != grp : grp != 0)) continue;  // the other group stores it
      const int n = n0 + wn * 32 * FNW + j * 32 + frow;
      const int mrow = m0 + wm * 32 * FMW + i * 32 + 4 * fkg;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = mrow + 8 * (r >> 2) + (r & 3);
        const bool ok = m < M && n < N;
        // take the element as a scalar first: bit-casting the vector element
        // directly was compiled into a dynamically indexed read that stored
        // element 0 for every r on this path
        const float v = acc[i][j][r];
        if constexpr (sizeof(TO) == 2) {
          if (!partial) {
            const uint32_t vo = ok ? (uint32_t)(m * N + n) * 2u : OOB;
            const TO h = (TO)v;
            __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), cr, (int)vo, 0, ST_NT);
            continue;
          }
        }
        const uint32_t vo = ok ? (uint32_t)(m * N + n) * 4u : OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), cr, (int)vo, 0, ST_NT);
      }
    }
}

// ---------------------------------------------------------------------------
// Weight gradient with the input staged as a halo (16-bit storage, 2-D,
// kw = 3, unit stride and rate along W, Ci % 64 == 0 and Ci >= 128,
// Wo >= 32, B Ho Wo a multiple of 32).
//
// A tile is one kernel row a, its three taps c and a block of 64 input
// channels (192 rows m = (a kw + c) Ci + 64 cb + q) by 128 output channels.
// The 32 output pixels of a K-step lie in at most two output rows
// (Wo >= 32); for each such run of n pixels the x pixels it reads through
// the three taps, one row of n + 2, are staged in LDS once, and the tap c
// fragment of pixel k is read at halo slot slot(k) + c, where slot(k) = k in
// the first run and k + 2 in the second.  The im2col form (conv_wgrad_kernel)
// stages and loads each x pixel three times, once per tap: here the A tile
// is 36 pixels x 64 channels instead of 32 x 192.  Measured against it (one
// box, alternating processes): 64x56x56x128 -> 128 layer 88.1 -> 73.7 us.
// Not taken: 64 input channels (64x56x56x64 -> 128: 40.9 against 38.7 us),
// and rows under 32 pixels (a form with three runs per step, for Wo >= 16,
// ran a 128x28x28x256 -> 256 layer in 168 against 132 us).  The MFMA loop,
// the two wave groups per workgroup and the FP32 split-K partials are those
// of conv_wgrad_kernel (same split plan and tile count: kh Ci / 64 row tiles
// of 192).  Every K-step is either inside the split or wholly past its end
// (B Ho Wo and the split lengths are multiples of 32), and a step past the
// end loads zeros for both operands, so the products are exactly those of
// the pixels the split owns.
// ---------------------------------------------------------------------------
template <typename TI, typename TO, int KG>
__global__ __launch_bounds__(256 * KG, KG == 1 ? 2 : 1) void conv_wgrad_halo_kernel(ConvParams p) {
  constexpr int FMW = 3, KW = 3, CB = 64;
  constexpr int BM = 64 * FMW, BN = 128, BK = 32;
  constexpr int HS = BK + 2 * (KW - 1);  // halo slots: two runs of up to BK pixels, each + KW - 1
  constexpr int RSA = CB * 2 + 64;       // halo slot row: 64 channels + 64 bytes pad (192 B: conflict-free tr reads)
  constexpr int RSB = BN * 2 + 64;
  constexpr int SA = 64 * RSA, STAGE = SA + BK * RSB;  // 64 slots: HS read, the rest take dummy stores
  constexpr int REG = FMW * 16 * 256;    // floats one group hands the other at the end (KG 2)
  constexpr int LDS_BYTES = (KG * 2 * STAGE > 2 * REG * 4) ? KG * 2 * STAGE : 2 * REG * 4;
  __shared__ __attribute__((aligned(16))) char lds_raw[LDS_BYTES];
  // the group is uniform per wave: as a scalar, the run rows and bases
  // derived from it below stay in scalar registers
  const int tid = threadIdx.x & 255, grp = __builtin_amdgcn_readfirstlane(threadIdx.x >> 8), lane = tid & 63,
            wave = tid >> 6;
  char *lds_grp = lds_raw + grp * 2 * STAGE;
  const int wm = wave >> 1, wn = wave & 1;
  const int M = (int)p.M, N = p.N;
  const int n_tiles = (N + BN - 1) / BN;
  const int cblocks = p.Ci / CB;
  const int mn_tiles = p.kh * cblocks * n_tiles;
  // the tiles of one split share their pixels: one XCD's L2 for them
  const int lid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int tile = lid % mn_tiles, split_idx = lid / mn_tiles;
  const int mt = tile / n_tiles, n0 = (tile % n_tiles) * BN;
  const int ta = mt / cblocks, cb = mt - (mt / cblocks) * cblocks;  // kernel row a, channel block
  const int64_t sbeg = (int64_t)split_idx * p.k_per_split;
  const int64_t send = sbeg + p.k_per_split < p.K ? sbeg + p.k_per_split : p.K;
  const int64_t half = p.k_per_split / KG;
  const int64_t kbeg = sbeg + grp * half;
  int64_t kend = kbeg + half < send ? kbeg + half : send;
  if (kend < kbeg) kend = kbeg;
  const int64_t len0 = half < send - sbeg ? half : send - sbeg;
  const int nk = (int)((len0 + BK - 1) / BK);

  const uint32_t x_bytes = (uint32_t)((int64_t)p.B * p.H * p.W * p.Ci * 2);
  const uint32_t g_bytes = (uint32_t)((int64_t)p.K * p.Co * 2);
  const __amdgpu_buffer_rsrc_t xr = make_rsrc(p.a, x_bytes);
  const __amdgpu_buffer_rsrc_t gr = make_rsrc(p.b, g_bytes);

  // ---- loader assignment.  A: HS x 8 chunks of 16 bytes (288): thread t
  // loads chunk t and, for t < 32, chunk 256 + t.  B: as conv_wgrad_kernel,
  // pixel kp = t / 8 of the step, chunks c8 and c8 + 8 of its 128 columns.
  const int c8 = tid & 7, kp = tid >> 3;
  const int sl0 = tid >> 3, sl1 = (tid + 256) >> 3;  // halo slots of the two A chunks
  const bool has1 = tid < HS * 8 - 256;
  const int coff = cb * CB + 8 * c8;                  // channel of this thread's A chunks
  const int bn = n0 + 8 * c8;
  const bool bok0 = bn < N, bok1 = bn + 64 < N;
  const int ah = ta * p.rh - p.ph;                    // x row offset of kernel row a

  // loader position: first pixel of the step being loaded, as (image, row,
  // column) digits advanced by BK without division (Wo >= BK: one carry)
  int lb, li, lj;
  {
    int t = (int)kbeg;
    lj = t % p.Wo;
    t /= p.Wo;
    li = t % p.Ho;
    lb = t / p.Ho;
  }
  int lk = (int)kbeg;          // absolute pixel index of the loaded step
  int gbase = (int)kbeg * p.Co + bn + kp * p.Co;  // dy element offset of the thread's B chunk 0
  const int kend32 = (int)kend;

  typedef uint4 Stage[4];  // A chunk 0, A chunk 1, B chunks 0 and 1
  Stage rs0, rs1;
  // The run rows, bases and bounds are uniform over the group (scalar
  // work); per lane a chunk needs a select, an add and its column check.
  // Threads without a second A chunk load zeros and store them past the
  // last halo slot (slots HS .. 63 are never read), so no branch splits the
  // K loop.
  const int sc0 = sl0 * p.Ci, sc1 = sl1 * p.Ci;
  auto a_off = [&](int slot, int sc, bool ok, int lim, int base1, int base2, bool r1, bool r2)
      __attribute__((always_inline)) {
    const bool s1 = slot < lim;
    const int col = s1 ? lj - p.pw + slot : slot - lim - p.pw;
    const bool in = ok & (s1 ? r1 : r2) & ((unsigned)col < (unsigned)p.W);
    return in ? (uint32_t)((s1 ? base1 : base2) + sc + coff) * 2u : OOB;
  };
  auto load = [&](Stage &rg) __attribute__((always_inline)) {
    const bool ok = lk < kend32;  // the whole step is inside the split or past its end
    const int n1 = p.Wo - lj < BK ? p.Wo - lj : BK;
    const int lim = n1 + KW - 1;
    const bool wrap = li + 1 == p.Ho;
    const int b1 = wrap ? lb + 1 : lb, i1 = wrap ? 0 : li + 1;
    const int h1 = li * p.sh + ah, h2 = i1 * p.sh + ah;
    const bool r1 = (lb < p.B) & ((unsigned)h1 < (unsigned)p.H);
    const bool r2 = (b1 < p.B) & ((unsigned)h2 < (unsigned)p.H);
    const int base1 = ((lb * p.H + h1) * p.W + lj - p.pw) * p.Ci;
    const int base2 = ((b1 * p.H + h2) * p.W - p.pw - lim) * p.Ci;
    rg[0] = bload(xr, a_off(sl0, sc0, ok, lim, base1, base2, r1, r2));
    rg[1] = bload(xr, a_off(sl1, sc1, ok & has1, lim, base1, base2, r1, r2));
    const uint32_t goff = (uint32_t)gbase * 2u;
    rg[2] = bload(gr, (ok & bok0) ? goff : OOB);
    rg[3] = bload(gr, (ok & bok1) ? goff + 128u : OOB);
    // advance by BK pixels
    lk += BK;
    gbase += BK * p.Co;
    lj += BK;
    if (lj >= p.Wo) {
      lj -= p.Wo;
      if (++li == p.Ho) {
        li = 0;
        ++lb;
      }
    }
  };
  auto store = [&](const Stage &rg, int st) __attribute__((always_inline)) {
    char *L = lds_grp + st * STAGE;
    *(uint4 *)(&L[sl0 * RSA + c8 * 16]) = rg[0];
    *(uint4 *)(&L[sl1 * RSA + c8 * 16]) = rg[1];
#pragma unroll
    for (int q = 0; q < 2; ++q) *(uint4 *)(&L[SA + kp * RSB + (c8 + 8 * q) * 16]) = rg[2 + q];
  };

  f32x16 acc[FMW][2];
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

  // transposed-read sources: lane j of a 16-lane group reads 4 channels of
  // pixel (j % 16) / 4 (+ 8 for the upper k-group, + 4 per half, + 16 per
  // quarter) of the step
  const int j16 = lane & 15;
  const int tr_row = (j16 >> 2) + 8 * (lane >> 5), tr_col = 2 * (16 * ((lane >> 4) & 1) + 4 * (j16 & 3));
  const int tr_b = SA + tr_row * RSB + tr_col;
  // compute-side column of the step's first pixel (for the run split)
  int cj = (int)(kbeg % p.Wo);

  auto kstep = [&](int ks, Stage &rn) __attribute__((always_inline)) {
    const char *Ls = lds_grp + (ks & 1) * STAGE;
    const int n1 = p.Wo - cj < BK ? p.Wo - cj : BK;
    constexpr int NQ = BK / 16;
    v4s16 a[NQ][FMW][2], b[NQ][2][2];
#pragma unroll
    for (int s = 0; s < NQ; ++s)
#pragma unroll
      for (int hf = 0; hf < 2; ++hf) {
        const int k = tr_row + 16 * s + 4 * hf;
        const int slot = k < n1 ? k : k + KW - 1;
        const char *pa = Ls + slot * RSA + tr_col;
#pragma unroll
        for (int f = 0; f < FMW; ++f) {
          const int F = wm * FMW + f;  // tap F / 2, channels 32 (F % 2) ..
          a[s][f][hf] = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s16 *)(pa + (F >> 1) * RSA + (F & 1) * 64));
        }
#pragma unroll
        for (int f = 0; f < 2; ++f)
          b[s][f][hf] = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
              (lds_v4s16 *)(Ls + tr_b + (16 * s + 4 * hf) * RSB + 2 * (wn * 64 + f * 32)));
      }
#pragma unroll
    for (int s = 0; s < NQ; ++s)
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = mfma16<TI>(a[s][i][0], a[s][i][1], b[s][j][0], b[s][j][1], acc[i][j]);
    store(rn, (ks & 1) ^ 1);
    load(rn);
    cj += BK;
    if (cj >= p.Wo) cj -= p.Wo;
    __builtin_amdgcn_sched_group_barrier(0x100, 2 * NQ * (FMW + 2), 0);
    __builtin_amdgcn_sched_group_barrier(0x008, 2 * NQ * FMW, 0);
    __syncthreads();
  };

  if (nk > 0) {
    load(rs0);  // step 0
    load(rs1);  // step 1
    store(rs0, 0);
    load(rs0);  // step 2
    __syncthreads();
    int ks = 0;
    for (; ks + 1 < nk; ks += 2) {
      kstep(ks, rs1);
      kstep(ks + 1, rs0);
    }
    if (ks < nk) kstep(ks, rs1);
  }

  if constexpr (KG == 2) {
    __syncthreads();
    float *X = (float *)lds_raw;
    float *mine = X + grp * REG;
    const float *theirs = X + (1 - grp) * REG;
    if (grp == 0) {
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) mine[(i * 16 + r) * 256 + tid] = acc[i][1][r];
    } else {
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) mine[(i * 16 + r) * 256 + tid] = acc[i][0][r];
    }
    __syncthreads();
    if (grp == 0) {
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][0][r] = acc[i][0][r] + theirs[(i * 16 + r) * 256 + tid];
    } else {
#pragma unroll
      for (int i = 0; i < FMW; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][1][r] = theirs[(i * 16 + r) * 256 + tid] + acc[i][1][r];
    }
  }

  // ---- epilogue: fragment F = wm FMW + i holds tap c = F / 2, channels
  // 32 (F % 2) .. of the block; row m = (a kw + c) Ci + 64 cb + q
  const int frow = lane & 31, fkg = lane >> 5;
  const bool partial = p.split > 1;
  const uint32_t c_bytes = partial ? (uint32_t)((int64_t)M * N * 4) : (uint32_t)((int64_t)M * N * sizeof(TO));
  char *cbase = (char *)p.c + (partial ? (int64_t)split_idx * M * N * 4 : 0);
  const __amdgpu_buffer_rsrc_t cr = make_rsrc(cbase, c_bytes);
#pragma unroll
  for (int i = 0; i < FMW; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      if (KG == 2 && j != grp) continue;
      const int n = n0 + wn * 64 + j * 32 + frow;
      const int F = wm * FMW + i;
      const int mrow = (ta * KW + (F >> 1)) * p.Ci + cb * CB + (F & 1) * 32 + 4 * fkg;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int m = mrow + 8 * (r >> 2) + (r & 3);
        const bool ok = n < N;
        const float v = acc[i][j][r];
        if constexpr (sizeof(TO) == 2) {
          if (!partial) {
            const uint32_t vo = ok ? (uint32_t)(m * N + n) * 2u : OOB;
            const TO h = (TO)v;
            __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, h), cr, (int)vo, 0, ST_NT);
            continue;
          }
        }
        const uint32_t vo = ok ? (uint32_t)(m * N + n) * 4u : OOB;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), cr, (int)vo, 0, ST_NT);
      }
    }
}

// The halo form of the weight gradient takes the problem (see
// conv_wgrad_halo_kernel for the conditions); OWL_AMD_WGRAD_HALO=0 turns it
// off (cross-checks against the im2col form).
bool wgrad_halo_fits(const ConvParams &p) {
  static const bool on = [] {
    const char *e = getenv("OWL_AMD_WGRAD_HALO");
    return !(e && atoi(e) == 0);
  }();
  return on && p.D == 1 && p.Do == 1 && p.kd == 1 && p.pd == 0 && p.kw == 3 && p.sw == 1 && p.rw == 1 &&
         p.Ci % 64 == 0 && p.Ci >= 128 && p.Wo >= 32 && p.K % 32 == 0 && wgrad_tile_m(p.M, 2) == 192 &&
         p.Co % 8 == 0 &&
         (p.split <= 1 || p.k_per_split % 64 == 0);
}

// Rows (tap, ci) per weight-gradient tile: 192 (fewer LDS reads per MFMA)
// when that pads M no more than 128 does, for 16-bit storage up to 10% more
// (conv3d 32 -> 64 channels, M = 864: 960 rows at 54.6 us against 896 at
// 60.3 us), else 128.  OWL_AMD_WGRAD_BM = 128 or 192 forces one (tuning).
int wgrad_tile_m(int64_t M, int es) {
  static const int forced = [] {
    const char *e = getenv("OWL_AMD_WGRAD_BM");
    return e ? atoi(e) : 0;
  }();
  if (forced == 128 || forced == 192) return forced;
  const int64_t p192 = (M + 191) / 192 * 192, p128 = (M + 127) / 128 * 128;
  return (es == 2 ? p192 * 10 <= p128 * 11 : p192 <= p128) ? 192 : 128;
}

// Groups of four waves per weight-gradient workgroup.  2: one 512-thread
// workgroup per CU whose groups sum the two halves of a split (half the
// partials).  1 for 16-bit tiles of 128 rows x 64 columns (N <= 64, about 120
// VGPRs): three 256-thread workgroups per CU then hide more latency than the
// halved partials save (conv3d 16 -> 32 channels: 30.5 + 5.4 us for the
// reduce against 35.3 + 4.1).  OWL_AMD_WGRAD_KG = 1 or 2 forces one.  Read
// on every call; the split-K plan, and so the workspace size, depends on it.
int wgrad_kgroups(const ConvParams &p, int es) {
  const char *e = getenv("OWL_AMD_WGRAD_KG");
  if (e && (atoi(e) == 1 || atoi(e) == 2)) return atoi(e);
  return es == 2 && p.N <= 64 && wgrad_tile_m(p.M, es) == 128 ? 1 : 2;
}

template <typename TI, typename TO>
void run_wgrad(const ConvParams &p, hipStream_t st) {
  const int bm = wgrad_tile_m(p.M, (int)sizeof(TI));
  const int kg = p.split > 1 ? wgrad_kgroups(p, (int)sizeof(TI)) : 1;
  dim3 grid((unsigned)(((p.M + bm - 1) / bm) * ((p.N + 127) / 128) * p.split)), blk(256 * kg);
  if constexpr (sizeof(TI) == 2) {
    if (wgrad_halo_fits(p)) {
      if (kg == 2) launch_k((conv_wgrad_halo_kernel<TI, TO, 2>), grid, blk, st, p);
      else launch_k((conv_wgrad_halo_kernel<TI, TO, 1>), grid, blk, st, p);
      return;
    }
  }
  const bool d1 = p.D == 1 && p.Do == 1 && p.kd == 1 && p.pd == 0;
  if (p.N <= 64) {
    // one 64-wide tile covers the output channels (the grid above counts
    // ceil(N / 128) = 1 column of tiles either way)
    if (kg == 2) {
      if (bm == 192) {
        if (d1) launch_k((conv_wgrad_kernel<TI, TO, 3, true, 2, 1>), grid, blk, st, p);
        else launch_k((conv_wgrad_kernel<TI, TO, 3, false, 2, 1>), grid, blk, st, p);
      } else {
        if (d1) launch_k((conv_wgrad_kernel<TI, TO, 2, true, 2, 1>), grid, blk, st, p);
        else launch_k((conv_wgrad_kernel<TI, TO, 2, false, 2, 1>), grid, blk, st, p);
      }
    } else if (bm == 192) {
      if (d1) launch_k((conv_wgrad_kernel<TI, TO, 3, true, 1, 1>), grid, blk, st, p);
      else launch_k((conv_wgrad_kernel<TI, TO, 3, false, 1, 1>), grid, blk, st, p);
    } else {
      if (d1) launch_k((conv_wgrad_kernel<TI, TO, 2, true, 1, 1>), grid, blk, st, p);
      else launch_k((conv_wgrad_kernel<TI, TO, 2, false, 1, 1>), grid, blk, st, p);
    }
    return;
  }
  if (kg == 2) {
    if (bm == 192) {
      if (d1) launch_k((conv_wgrad_kernel<TI, TO, 3, true, 2>), grid, blk, st, p);
      else launch_k((conv_wgrad_kernel<TI, TO, 3, false, 2>), grid, blk, st, p);
    } else {
      if (d1) launch_k((conv_wgrad_kernel<TI, TO, 2, true, 2>), grid, blk, st, p);
      else launch_k((conv_wgrad_kernel<TI, TO, 2, false, 2>), grid, blk, st, p);
    }
  } else if (bm == 192) {
    if (d1) launch_k((conv_wgrad_kernel<TI, TO, 3, true, 1>), grid, blk, st, p);
    else launch_k((conv_wgrad_kernel<TI, TO, 3, false, 1>), grid, blk, st, p);
  } else {
    if (d1) launch_k((conv_wgrad_kernel<TI, TO, 2, true, 1>), grid, blk, st, p);
    else launch_k((conv_wgrad_kernel<TI, TO, 2, false, 1>), grid, blk, st, p);
  }
}
template void run_wgrad<bf16_t, bf16_t>(const ConvParams &, hipStream_t);
template void run_wgrad<bf16_t, float>(const ConvParams &, hipStream_t);
template void run_wgrad<f16_t, f16_t>(const ConvParams &, hipStream_t);
template void run_wgrad<f16_t, float>(const ConvParams &, hipStream_t);
template void run_wgrad<float, float>(const ConvParams &, hipStream_t);

// ---------------------------------------------------------------------------
// Stride-1 2-D convolution with an LDS-staged input halo.
//
// A workgroup owns TR full output rows of one image and a BN-wide slice of
// output channels: 4 waves of FM x FN 32x32 fragments, so TR * Wo <= 512
// pixels for 64-wide slices (FM 4) and <= 256 for 128-wide ones (FM 2).
// For each block
// of BK input channels it stages the input rows those output rows read --
// (TR - 1) + (kh - 1) rh + 1 rows of Wo + (kw - 1) rw columns, zero-filled
// outside the image -- into LDS once, and every tap then reads its A
// fragments from that halo at a fixed offset.  Each activation is fetched
// from memory once per channel block instead of once per tap.
//
// PASS 0 is the forward convolution.  PASS 1 is the backward input pass
// written as a forward pass over dy: with stride 1,
//   dx[h] = sum_a dy[h - p' + a r] w[kh-1-a]^T,  p' = (kh-1) r - p,
// so the same kernel runs with dy as input, the taps flipped and the kernel
// read as w[tap][ci][co] with co contiguous.
// ---------------------------------------------------------------------------
struct HaloGeom {
  int Hi, Wi, C;       // input image (x, or dy for PASS 1) and contracted channels
  int Ho, Wo, N;       // output image and output channels
  int kh, kw, rh, rw;  // taps and dilation
  int ph, pw;          // leading pads (of this forward view)
  int B, TR;           // batch, output rows per tile
  int ms;              // MFMA fragment edge: 32 or 16
  int bn;              // output channels per tile (channel slice): 32, 64 or 128
  int wCi, wCo;        // the kernel buffer's [kh][kw][wCi][wCo] sizes
  // 3-D layers (stride 1): each "image" above is one (w, d) plane of an
  // NDHWC tensor -- for fixed (b, h) the W x D pixels are contiguous -- so B
  // counts B x HOA output planes and the kernel's h taps are KA outer
  // blocks around the channel blocks: block a restages the halo from input
  // plane i + a ra - pa of the same image (zeros outside [0, HA)) and uses
  // kernel taps a kh kw + (c, e).  2-D: KA = HA = HOA = ra = 1, pa = 0.
  int KA, ra, pa, HA, HOA;
  int IB;              // input images (planes) in the input buffer
  // forward only: fused bias / activation and the pre-activation output
  const float *bias;
  void *pre;
  int act;
  // forward, 16-bit output only: 2x2 stride-2 max-pool of y written here
  // from the staged tile (TR even, or one row tile), else null
  void *pool;
  // FP8 forward: per-tensor scales of x and w (device floats), else null
  const float *sa, *sb;
};

// MFMA of the halo kernel by fragment edge MS: 32 -> v_mfma_f32_32x32x16
// (two per 32-channel K-step, 16 accumulators per lane), 16 ->
// v_mfma_f32_16x16x32 (one per K-step, 4 accumulators per lane).  Both run
// at the same rate; 16-pixel fragments tile rows whose pixel count is not a
// multiple of 32 with less padding.
// compile-time flag for the staged epilogue's variants
template <bool V> struct HaloFlag {
  static constexpr bool value = V;
};

template <typename TI, int MS> struct HaloMfma;
template <> struct HaloMfma<bf16_t, 32> {
  typedef f32x16 acc_t;
  __device__ static __forceinline__ acc_t mma(const uint4 &a, const uint4 &b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
  }
};
template <> struct HaloMfma<f16_t, 32> {
  typedef f32x16 acc_t;
  __device__ static __forceinline__ acc_t mma(const uint4 &a, const uint4 &b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                  0);
  }
};
template <> struct HaloMfma<bf16_t, 16> {
  typedef f32x4 acc_t;
  __device__ static __forceinline__ acc_t mma(const uint4 &a, const uint4 &b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                   0, 0);
  }
};
template <> struct HaloMfma<f16_t, 16> {
  typedef f32x4 acc_t;
  __device__ static __forceinline__ acc_t mma(const uint4 &a, const uint4 &b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a), __builtin_bit_cast(f16x8, b), c, 0, 0,
                                                  0);
  }
};

// MS: fragment edge (32 or 16).  FM x FN fragments per wave; the four waves
// form (4 / WGN) x WGN along (pixels, channels), so a tile holds
// CAP = (4 / WGN) FM MS pixels and BN = WGN FN MS channels.
// KROWS: the kernel operand is laid out [tap][n][k] (k contiguous), the
// backward-input view of w; its tile is copied to LDS as [n][k] rows and
// read with 16-byte loads.  Otherwise (forward weights in place,
// [tap][k][n]) the tile is copied in that natural layout and the fragments
// are read with the transposing LDS load (ds_read_b64_tr_b16, as in the
// weight-gradient kernel), so no pass rearranges the weights first.
// F8: FP8 e4m3 forward (PASS 0, MS 32, KROWS).  The operands are bytes, and
// every offset is counted in pairs of them, so a K-step of 32 "channels" is
// 64 FP8 channels (64 bytes per halo pixel, as for 32 16-bit channels) and
// each lane's two 16-byte A and B chunks are exactly the 32 bytes one
// v_mfma_scale_f32_32x32x64_f8f6f4 takes: the two 32x32x16 MFMAs of a step
// become one.  A and B give each lane group the same bytes in the same
// order, so the products pair up whatever order the instruction sums them
// in.  The per-tensor scales multiply the sums in the epilogue.  Waves hold
// 2 x 2 fragments (tiles of 256 pixels x 64 channels): the scaled FP8 MFMA
// keeps its result apart from its accumulator input, and 2 x 4 fragments
// then spilled about 400 bytes per lane.
// GEN: the general instance -- 3-D layers as (w, d) planes with outer h
// blocks (see HaloGeom), and channel counts that are multiples of 8 but not
// of 32, whose last channel block loads the chunks and kernel rows past the
// last channel as zeros.  The other instances compile that logic out.
template <int PASS, typename TI, typename TO, int MS, int FM, int FN, int WGN, bool KROWS, bool F8 = false,
          bool GEN = false>
__global__ __launch_bounds__(256, 2) void conv_halo_kernel(HaloGeom g, const TI *__restrict__ in,
                                                           const TI *__restrict__ wt, TO *__restrict__ out) {
  static_assert(!F8 || (PASS == 0 && MS == 32 && KROWS && FM == 2 && FN == 2), "FP8: forward, 32-edge, k-rows kernel tile");
  typedef HaloMfma<TI, MS> MF;
  typedef typename MF::acc_t acc_t;
  constexpr int NR = MS == 32 ? 16 : 4;  // accumulators per lane and fragment
  constexpr int WGM = 4 / WGN;
  constexpr int CAP = WGM * FM * MS;     // pixels per tile
  // Halo pixels and [n][k] kernel rows: 32-edge fragments use 80-byte rows
  // (32 channels + 16 B pad), which their reads (32 consecutive pixels per
  // half-wave) hit without bank conflicts.  16-edge fragments read 16 pixels
  // x 4 chunks per instruction, which conflicts on padded rows (two LDS
  // cycles per read, three across a row end, measured on gfx950).  With
  // 64-byte rows and chunk c of pixel (or row) p at 16-byte slot
  // c ^ ((p >> 1) & 3) those reads are conflict-free at any tap shift, and
  // so are the 4-chunks-per-pixel writes, but each A read then needs its
  // address rebuilt per tap (five VALU operations).  That paid off only for
  // the backward input (38.8-39.4 against 39.1-40.0 us on a 64x56x56x64 ->
  // 128 layer; its B rows need no per-tap work), while the forward ran
  // slower (44.0 against 42.0 us), so only the KROWS instance swizzles.
  constexpr bool SWZ = MS == 16 && KROWS;
  constexpr int BK = 32, BN = WGN * FN * MS, PR = SWZ ? BK : BK + 8;
  // kernel-tile columns staged per step: at least 64, so every thread moves
  // whole 16-byte chunks; a 32-wide tile stages columns past N as zeros
  // (layers with at most 32 output channels) and its MFMAs read the first 32
  constexpr int BNL = BN < 64 ? 64 : BN;
  constexpr int LDR = PR;
  // The halo is one buffer of up to 2 HBUF pixels, restaged between channel
  // blocks behind a barrier with HQ 16-byte chunks per thread in flight at
  // once.  Loading the next block's halo into registers during a block (and
  // double-buffering it in LDS) measured slower everywhere once the
  // restaged form's chunk writes were made unconditional (see halo_commit):
  // the registers it holds across a block cost the fragment pipelining more
  // than the early loads save (64x56x56x64 -> 128 BF16 backward input 34.1
  // against 31.4 us, FP8 forward 28.0 against 26.0 us), and spreading the
  // next block's halo over the block's steps, one chunk per thread and
  // step, ran the forward at 39.9 against 35.2 us.
  constexpr int HQ = CAP > 256 ? 10 : 6, HBUF = 352;
  // one LDS block: the halo, then the two kernel tiles; after the K loop it
  // holds the output tile for the 16-byte store epilogue
  constexpr int HALO_BYTES = 2 * HBUF * PR * (int)sizeof(TI);
  // natural-layout kernel tile: BK rows of BN values, padded by 64 bytes so
  // the four rows one transposing read phase touches cover all 64 banks
  constexpr int RSB = BNL * (int)sizeof(TI) + 64;
  constexpr int B_BYTES = KROWS ? BNL * LDR * (int)sizeof(TI) : BK * RSB;
  // 16-bit outputs are staged for the 16-byte store epilogue in the same
  // block, which is sized for whichever use is larger
  constexpr int K_BYTES = HALO_BYTES + 2 * B_BYTES;
  constexpr int STG_BYTES = CAP * (BN * (int)sizeof(TO) + 16);
  constexpr int LDS_BYTES = sizeof(TO) == 2 && STG_BYTES > K_BYTES ? STG_BYTES : K_BYTES;
  // plus a 16-byte slot per lane past the end that takes the halo chunks
  // beyond the halo (q >= hq_total), so that every lane writes its chunks
  // unconditionally (one slot per lane: lanes writing one address conflict)
  __shared__ __attribute__((aligned(16))) char smem[LDS_BYTES + 64 * 16];
  TI *halo = (TI *)smem;
  char *const Bs0 = smem + HALO_BYTES;  // kernel tile buffer i at Bs0 + i * B_BYTES

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WGN, wn = wave % WGN;
  // lane -> fragment row / column (frow) and k group (fkg): 32-edge
  // fragments give each half-wave 16 of a K-step's 32 channels (8 per MFMA),
  // 16-edge fragments each quarter-wave 8
  const int frow = lane & (MS - 1), fkg = lane / MS;
  const int n_tiles = (g.N + BN - 1) / BN;
  const int row_tiles = (g.Ho + g.TR - 1) / g.TR;
  const int bid = xcd_swizzle(blockIdx.x, gridDim.x);
  const int nt = bid % n_tiles;
  const int rt = (bid / n_tiles) % row_tiles;
  const int b = bid / (n_tiles * row_tiles);  // output image (plane) of the tile
  // GEN: the plane's image and index (the other instances compile the
  // outer-block logic out: KA = 1, the halo always from image b)
  const int bimg = GEN ? b / g.HOA : b, bpl = GEN ? b - (b / g.HOA) * g.HOA : 0;
  const int KA = GEN ? g.KA : 1;
  const int n0 = nt * BN, i0 = rt * g.TR;
  const int TRv = min(g.TR, g.Ho - i0);     // output rows in this tile
  const int P = TRv * g.Wo;                 // output pixels in this tile
  const int HR = (TRv - 1) + (g.kh - 1) * g.rh + 1;
  const int HC = (g.Wo - 1) + (g.kw - 1) * g.rw + 1;
  const int hr0 = i0 - g.ph, hc0 = -g.pw;   // input coordinates of halo (0, 0)

  // per-lane fragment pixels -> halo offset of the tap (0,0) window origin
  int hoff[FM];
#pragma unroll
  for (int f = 0; f < FM; ++f) {
    int p = wm * FM * MS + f * MS + frow;
    if (p >= P) p = 0;
    const int ti = p / g.Wo, tj = p - (p / g.Wo) * g.Wo;
    // 16-edge: the halo pixel index (the swizzle needs the tap's pixel)
    hoff[f] = SWZ ? ti * HC + tj : (ti * HC + tj) * PR + fkg * (MS == 32 ? 16 : 8);
  }

  const uint32_t in_bytes = (uint32_t)((int64_t)g.IB * g.Hi * g.Wi * g.C * 2);
  const __amdgpu_buffer_rsrc_t inr = make_rsrc(in, in_bytes);
  const int T = g.kh * g.kw;  // taps of one outer block
  const __amdgpu_buffer_rsrc_t wr = make_rsrc(wt, (uint32_t)(KA * T * g.N * g.C * 2));
  const int cblocks = GEN ? (g.C + BK - 1) / BK : g.C / BK;
  // the input image of the halo for outer block a, and whether it exists
  int himg = b;
  bool hval = true;
  auto set_plane = [&](int a) __attribute__((always_inline)) {
    if constexpr (GEN) {
      const int ip = bpl + a * g.ra - g.pa;
      hval = (unsigned)ip < (unsigned)g.HA;
      himg = bimg * g.HA + (hval ? ip : 0);
    }
  };
  // 16-byte slot XOR of pixel / kernel row p (0 for 80-byte rows)
  auto swz = [](int p) __attribute__((always_inline)) -> int { return SWZ ? (p >> 1) & 3 : 0; };

  // halo staging: chunk q -> pixel q / 4, 16-byte channel chunk q % 4;
  // OOB marks padding, which the buffer loads fill with zeros.
  const int hq_total = HR * HC * 4;
  auto chunk_src = [&](int q) __attribute__((always_inline)) -> uint32_t {
    const int pix = q >> 2, ch = q & 3;
    const int r = pix / HC, c = pix - (pix / HC) * HC;
    const int hi = hr0 + r, wi = hc0 + c;
    if ((!GEN || hval) && q < hq_total && (unsigned)hi < (unsigned)g.Hi && (unsigned)wi < (unsigned)g.Wi)
      return (uint32_t)((((GEN ? himg : b) * g.Hi + hi) * g.Wi + wi) * g.C + ch * 8) * 2u;
    return OOB;
  };
  // This thread's chunks q = tid + 256 u sit 64 pixels apart: the pixel's
  // (row, column) in the halo steps by 64 in radix HC, one carry per step.
  const int hs_r = 64 / HC, hs_c = 64 - (64 / HC) * HC;
  const int hpix0 = tid >> 2, hch = tid & 3;
  const int hr_0 = hpix0 / HC, hc_0 = hpix0 - (hpix0 / HC) * HC;
  uint4 hreg[HQ];
  // the first HQ * 256 chunks go through registers (loads issued ahead of
  // their LDS writes); halos beyond that (rows wider than 64 HQ pixels
  // allow) copy their remaining chunks load by load
  auto halo_issue = [&](int c0) __attribute__((always_inline)) {
    int r = hr_0, c = hc_0;
#pragma unroll
    for (int u = 0; u < HQ; ++u) {
      const int hi = hr0 + r, wi = hc0 + c;
      const bool ok = (!GEN || (hval && c0 + hch * 8 < g.C)) && tid + u * 256 < hq_total &&
                      (unsigned)hi < (unsigned)g.Hi && (unsigned)wi < (unsigned)g.Wi;
      const uint32_t so = (uint32_t)((((GEN ? himg : b) * g.Hi + hi) * g.Wi + wi) * g.C + c0 + hch * 8) * 2u;
      hreg[u] = bload(inr, ok ? so : OOB);
      c += hs_c;
      const int cy = c >= HC;
      if (cy) c -= HC;
      r += hs_r + cy;
    }
  };
  // Every lane writes all HQ chunks, those past the halo into the dump slot:
  // a write guarded per lane is an exec-masked branch, after which the
  // compiler cannot tell that the chunk's load has been waited for, and it
  // then waited for every load in flight (s_waitcnt vmcnt(0)) before the
  // next step's fragment reads reused those registers, draining the
  // kernel-tile ring at every step.
  auto halo_commit = [&](TI *hb) __attribute__((always_inline)) {
#pragma unroll
    for (int u = 0; u < HQ; ++u) {
      const int q = tid + u * 256;
      char *dst = q < hq_total ? (char *)&hb[(q >> 2) * PR + ((q & 3) ^ swz(q >> 2)) * 8] : smem + LDS_BYTES + lane * 16;
      *(uint4 *)dst = hreg[u];
    }
  };
  auto stage_halo = [&](TI *hb, int c0) __attribute__((always_inline)) {
    halo_issue(c0);
    halo_commit(hb);
    for (int q = HQ * 256 + tid; q < hq_total; q += 256) {
      const uint32_t so = chunk_src(q);
      const bool cok = !GEN || c0 + (q & 3) * 8 < g.C;
      *(uint4 *)&hb[(q >> 2) * PR + ((q & 3) ^ swz(q >> 2)) * 8] =
          bload(inr, so == OOB || !cok ? OOB : so + (uint32_t)c0 * 2u);
    }
  };

  // B staging for (tap, c0): natural layout reads rows w[tap][c0 + k][n0 ..]
  // (BN / 8 chunks of 16 bytes per row); KROWS reads rows w[tap][n][c0 ..].
  // Both use unconditional buffer loads (rows or columns past N read as
  // zeros), which keeps the number of loads in flight static, so the waits
  // stay precise across the loop.
  constexpr int BELEMS = BK * BNL / 256;
  constexpr int BNV = BELEMS / 8;              // 16-byte chunks per thread
  constexpr int CPRB = BNL / 8, RPP = 256 / CPRB;  // natural layout: chunks per row, rows per pass
  const int nb_ch = tid % CPRB, nb_row = tid / CPRB;
  const int b_kv = (tid % 4) * 8, b_row = tid / 4;
  TI rb0[BELEMS], rb1[BELEMS];  // register ring: B tiles two steps ahead
  auto load_b = [&](TI(&rb)[BELEMS], int tap, int c0) __attribute__((always_inline)) {
    if (!KROWS) {
      const int n = n0 + nb_ch * 8;
#pragma unroll
      for (int i = 0; i < BNV; ++i) {
        const int k = c0 + nb_row + i * RPP;
        const uint32_t wo = n < g.N && (!GEN || k < g.C) ? (uint32_t)((tap * g.wCi + k) * g.wCo + n) * 2u : OOB;
        const uint4 v = bload(wr, wo);
        __builtin_memcpy(&rb[i * 8], &v, 16);
      }
    } else {
      // rows of k: [tap][n][k] with N rows of C values per tap
      const int ft = PASS == 1 ? KA * T - 1 - tap : tap;  // backward input flips the taps
#pragma unroll
      for (int i = 0; i < BNV; ++i) {
        const int n = n0 + b_row + i * 64;
        const uint32_t wo =
            n < g.N && (!GEN || c0 + b_kv < g.C) ? (uint32_t)((ft * g.N + n) * g.C + c0 + b_kv) * 2u : OOB;
        const uint4 v = bload(wr, wo);
        __builtin_memcpy(&rb[i * 8], &v, 16);
      }
    }
  };
  // MS 16 reads k rows r and r + 8 of the natural tile in one phase, which
  // share banks at this row stride: rows with bit 3 set keep their 16-byte
  // chunks swapped in pairs two apart (columns XOR 16).  RPP is a multiple
  // of 16, so a thread's rows all have the same bit 3.
  const int nb_swz = MS == 16 && !KROWS ? (nb_ch ^ (2 * ((nb_row >> 3) & 1))) : nb_ch;
  auto store_b = [&](const TI(&rb)[BELEMS], int buf) __attribute__((always_inline)) {
    char *dst = Bs0 + buf * B_BYTES;
#pragma unroll
    for (int i = 0; i < BNV; ++i) {
      // KROWS rows b_row + 64 i share bits 1-2, so one slot XOR per thread
      const int off = KROWS ? ((b_row + i * 64) * LDR + ((b_kv >> 3) ^ swz(b_row)) * 8) * (int)sizeof(TI)
                            : (nb_row + i * RPP) * RSB + nb_swz * 16;
      __builtin_memcpy(dst + off, &rb[i * 8], 16);
    }
  };
  // Transposing fragment reads (natural layout): in each group of 16 lanes,
  // ds_read_b64_tr_b16 gives lane i halfword (i % 4) of the 8 bytes addressed
  // by lane 4e + i / 4 as its element e.  Lane i of a group pointing at k row
  // k0 + i / 4 and columns c0 + 4 (i % 4) .. + 3 therefore receives column
  // c0 + i at the four rows k0 .. k0 + 3; a second read 4 rows further on
  // completes the 8 k values a lane's fragment holds.  MS 32: the group of
  // lanes 16 h .. 16 h + 15 covers columns 16 (h % 2) .., rows 16 (h / 2) +
  // 8 s2 .. for the K-step's MFMA s2.  MS 16: group h covers the fragment's
  // 16 columns at rows 8 h ..
  // MS 16 with the swizzle above: the odd groups (rows with bit 3 set) read
  // fragment j's columns from block j ^ 1.
  const int j16 = lane & 15, hodd = (lane >> 4) & 1;
  const int tr_off = MS == 32 ? (fkg * 16 + (j16 >> 2)) * RSB + 2 * (16 * ((lane >> 4) & 1) + 4 * (j16 & 3))
                              : (8 * (lane >> 4) + (j16 >> 2)) * RSB + 2 * (4 * (j16 & 3));
  const int wn0 = wn * FN * MS;  // this wave's first column in the tile (a multiple of 32)
  auto read_b = [&](const char *bbuf, int j, int s2) __attribute__((always_inline)) -> uint4 {
    if constexpr (!KROWS) {
      const int jb = MS == 16 ? (j ^ hodd) : j;
      const char *base = bbuf + tr_off + (8 * s2) * RSB + 2 * (wn0 + MS * jb);
      const v4s16 lo = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s16 *)(base));
      const v4s16 hi = __builtin_amdgcn_ds_read_tr16_b64_v4i16((lds_v4s16 *)(base + 4 * RSB));
      return __builtin_bit_cast(uint4, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
    } else {
      // 16-edge rows wn0 + 16 j + frow: bits 1-2 come from frow
      const int kofs = MS == 32 ? fkg * 16 + 8 * s2 : 8 * (fkg ^ swz(frow));
      return *(const uint4 *)(bbuf + ((wn0 + j * MS + frow) * LDR + kofs) * (int)sizeof(TI));
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int r = 0; r < NR; ++r) acc[i][j][r] = 0.0f;

  const int nk = KA * cblocks * T;
  // load cursor: (outer block, channel block, tap) of the next B tile to
  // fetch.  Past the last step it fetches tile (0, 0, 0) again, so every
  // step issues the same loads and the waits on them stay exact.
  int ltap = 0, lcb = 0, la = 0;
  auto next_load = [&](TI(&rb)[BELEMS]) __attribute__((always_inline)) {
    if constexpr (GEN) {
      const bool in = la < KA;
      load_b(rb, in ? la * T + ltap : 0, in ? lcb * BK : 0);
      if (++ltap == T) {
        ltap = 0;
        if (++lcb == cblocks) {
          lcb = 0;
          ++la;
        }
      }
    } else {
      const bool in = lcb < cblocks;
      load_b(rb, in ? ltap : 0, in ? lcb * BK : 0);
      if (++ltap == T) {
        ltap = 0;
        ++lcb;
      }
    }
  };
  next_load(rb0);  // step 0
  next_load(rb1);  // step 1
  set_plane(0);
  stage_halo(halo, 0);
  store_b(rb0, 0);
  next_load(rb0);  // step 2
  __syncthreads();
  // current step: block cb (outer block sa, channel block sc), tap = (ta, tc)
  int cb = 0, tap = 0, ta = 0, tc = 0, sa = 0, sc = 0;
  // one K-step; rnext holds B(ks + 1) and is refilled with B(ks + 3)
  auto step = [&](int ks, TI(&rnext)[BELEMS]) __attribute__((always_inline)) {
    const int bcur = ks & 1;
    // next step's (tap, channel block)
    int ntap = tap + 1, ncb = cb, nta = ta, ntc = tc + 1;
    if (ntc == g.kw) {
      ntc = 0;
      ++nta;
    }
    if (ntap == T) {
      ntap = 0;
      nta = 0;
      ++ncb;
    }
    const bool more = ks + 1 < nk;
    // A from the halo at this tap's offset, B from LDS
    const TI *hcur = halo;
    const int tsh = ta * g.rh * HC + tc * g.rw;  // the tap's pixel shift in the halo
    const int toff = tsh * PR;
    const char *bs = Bs0 + bcur * B_BYTES;
    if constexpr (MS == 16) {
      // one MFMA per fragment pair covers the K-step's 32 channels
      uint4 fa[FM], fb[FN];
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        if constexpr (SWZ) {
          const int pix = hoff[f] + tsh;
          fa[f] = *(const uint4 *)&hcur[pix * PR + 8 * (fkg ^ swz(pix))];
        } else {
          fa[f] = *(const uint4 *)&hcur[hoff[f] + toff];
        }
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) fb[j] = read_b(bs, j, 0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(fb[j], fa[i], acc[i][j]);
    } else if constexpr (FM == 4) {
      // 128 accumulators leave no room for both K halves' fragments: read
      // and consume one half at a time
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2) {
        uint4 fa[FM], fb[FN];
#pragma unroll
        for (int f = 0; f < FM; ++f) fa[f] = ((const uint4 *)&hcur[hoff[f] + toff])[s2];
#pragma unroll
        for (int j = 0; j < FN; ++j) fb[j] = read_b(bs, j, s2);
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(fb[j], fa[i], acc[i][j]);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else if constexpr (F8) {
      // one 32-byte operand per fragment: the two 16-byte reads land in the
      // halves of one 8-register vector
      i32x8 fa[FM], fb[FN];
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        const uint4 *src = (const uint4 *)&hcur[hoff[f] + toff];
        fa[f] = __builtin_shufflevector(__builtin_bit_cast(i32x4, src[0]), __builtin_bit_cast(i32x4, src[1]), 0, 1, 2,
                                        3, 4, 5, 6, 7);
      }
#pragma unroll
      for (int j = 0; j < FN; ++j)
        fb[j] = __builtin_shufflevector(__builtin_bit_cast(i32x4, read_b(bs, j, 0)),
                                        __builtin_bit_cast(i32x4, read_b(bs, j, 1)), 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
          // cbsz = blgp = 0: both operands OCP e4m3; block scales 2^0
          acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(fb[j], fa[i], acc[i][j], 0, 0, 0, 127, 0, 127);
    } else {
      uint4 fa[FM][2], fb[FN][2];
#pragma unroll
      for (int f = 0; f < FM; ++f) {
        const uint4 *src = (const uint4 *)&hcur[hoff[f] + toff];
        fa[f][0] = src[0];
        fa[f][1] = src[1];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        fb[j][0] = read_b(bs, j, 0);
        fb[j][1] = read_b(bs, j, 1);
      }
#pragma unroll
      for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = MF::mma(fb[j][s2], fa[i][s2], acc[i][j]);
    }
    // unconditional, so every step issues and waits for the same loads;
    // after the last step this fills a buffer nobody reads again
    store_b(rnext, bcur ^ 1);
    next_load(rnext);
    if (more && ncb != cb) {
      if constexpr (GEN) {
        if (++sc == cblocks) {
          sc = 0;
          set_plane(++sa);
        }
      }
      __syncthreads();  // every wave is done with this channel block's halo
      stage_halo(halo, (GEN ? sc : ncb) * BK);
    }
    __syncthreads();
    tap = ntap;
    ta = nta;
    tc = ntc;
    cb = ncb;
  };
  // pairs of steps, both unconditional (the register ring alternates), then
  // an odd last step: no path through the loop skips a step's loads, which
  // keeps the compiler's load waits at the two steps of lookahead
  int ks = 0;
  for (; ks + 1 < nk; ks += 2) {
    step(ks, rb1);
    step(ks + 1, rb0);
  }
  if (ks < nk) step(ks, rb1);

  // ---- epilogue: output pixel (b, i0 + p / Wo, p % Wo), channel n ----
  // The MFMAs take the kernel tile as their first operand, so each result
  // block is channels x pixels: a 32x32 result gives lane l column (pixel)
  // l % 32 of rows (channels) 8 (r / 4) + 4 (l / 32) + r % 4, a 16x16 result
  // column l % 16 of rows 4 (l / 16) + r.  Accumulator r of fragment (i, j)
  // thus holds tile pixel pix(i) = wm FM MS + i MS + frow and tile channel
  // chn(j) + rofs(r), chn(j) = wn0 + j MS + 4 fkg: each run of four
  // accumulators is four consecutive channels of one pixel, written as one
  // 8-byte (16-bit) or 16-byte (FP32) piece of its NHWC row.
  const uint32_t o_bytes = (uint32_t)((int64_t)g.B * g.Ho * g.Wo * g.N * sizeof(TO));
  const int obase = (b * g.Ho + i0) * g.Wo;  // first output pixel of the tile
  auto rofs = [](int r) __attribute__((always_inline)) -> int { return MS == 32 ? 8 * (r >> 2) + (r & 3) : r; };
  auto pix = [&](int i) __attribute__((always_inline)) -> int { return wm * FM * MS + i * MS + frow; };
  auto chn = [&](int j) __attribute__((always_inline)) -> int { return wn0 + j * MS + 4 * fkg; };
  // fused bias and activation (forward only); relu is fmax(v, 0): NaN -> 0
  const bool epi = PASS == 0 && (g.bias || g.act);
  const float fsc = F8 ? g.sa[0] * g.sb[0] : 1.0f;
  // the bias of run q (accumulators 4q .. 4q + 3) of fragment column j
  auto bias4 = [&](int j, int q, float (&bv)[4]) __attribute__((always_inline)) {
    const int n = n0 + chn(j) + rofs(4 * q);
#pragma unroll
    for (int e = 0; e < 4; ++e) bv[e] = (epi && g.bias && n < g.N) ? g.bias[n + e] : 0.0f;
  };
  // run q of fragment (i, j) as four outputs of type TO
  auto run4 = [&](int i, int j, int q, const float (&bv)[4], bool add_bias, bool relu,
                  TO (&o)[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      float v = acc[i][j][4 * q + e];
      if (F8) v *= fsc;
      if (add_bias) v += bv[e];
      if (relu) v = fmaxf(v, 0.0f);
      o[e] = (TO)v;
    }
  };
  // staged output row: BN values, plus 16 B of padding when that still fits
  constexpr int ORS_PAD = BN * (int)sizeof(TO) + 16;
  constexpr int ORS = CAP * ORS_PAD <= LDS_BYTES ? ORS_PAD : BN * (int)sizeof(TO);
  // dst may be null when the tile only feeds the fused pooling.  dst_pre
  // (staged 16-bit outputs): the tile is staged once, before the activation,
  // and stored to dst_pre; y is formed from it per 16-bit value on the way
  // out.  relu commutes with rounding to 16 bits, and the integer form
  // below maps negatives, -0 and NaN to +0 as fmaxf(v, 0) does before
  // rounding (v is never -0 there: the sums start from +0), so y has the
  // bits of the two-pass epilogue.
  auto emit = [&](TO *dst, bool act, TO *dst_pre) __attribute__((always_inline)) {
    const __amdgpu_buffer_rsrc_t orr = make_rsrc(dst, o_bytes);
    if constexpr (CAP * ORS <= LDS_BYTES) {
      // Stage the tile in LDS as [pixel][channel] rows (the LDS is free after
      // the K loop's last barrier), then store whole 16-byte chunks.  The
      // bias and relu choices are made once, outside the unrolled loops:
      // tested per value they cost 3.5 us on a 64x56x56x64 -> 128 layer.
      auto stage = [&](auto HB, auto HA) __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int q = 0; q < NR / 4; ++q) {
            float bv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            if constexpr (decltype(HB)::value) bias4(j, q, bv);
#pragma unroll
            for (int i = 0; i < FM; ++i) {
              TO o[4];
              run4(i, j, q, bv, decltype(HB)::value, decltype(HA)::value, o);
              __builtin_memcpy(smem + pix(i) * ORS + (chn(j) + rofs(4 * q)) * (int)sizeof(TO), o, sizeof(o));
            }
          }
      };
      const bool hb = epi && g.bias, ha = epi && act && g.act == OWL_AMD_ACT_RELU;
      if (hb && ha) stage(HaloFlag<true>(), HaloFlag<true>());
      else if (hb) stage(HaloFlag<true>(), HaloFlag<false>());
      else if (ha) stage(HaloFlag<false>(), HaloFlag<true>());
      else stage(HaloFlag<false>(), HaloFlag<false>());
      __syncthreads();
      constexpr int CPR = BN * (int)sizeof(TO) / 16;  // 16-byte chunks per row
      constexpr int EPC = 16 / (int)sizeof(TO);       // values per chunk
      const bool relu_y = dst_pre && g.act == OWL_AMD_ACT_RELU;
      // relu of eight 16-bit values: sign bit or NaN -> +0
      auto relu8 = [](uint4 v) __attribute__((always_inline)) -> uint4 {
        constexpr unsigned INF16 = __is_same(TO, f16_t) ? 0x7c00u : 0x7f80u;
        unsigned short h[8];
        __builtin_memcpy(h, &v, 16);
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if ((h[e] & 0x8000u) || (h[e] & 0x7fffu) > INF16) h[e] = 0;
        __builtin_memcpy(&v, h, 16);
        return v;
      };
      if (dst || dst_pre) {
        const __amdgpu_buffer_rsrc_t qrr = make_rsrc(dst_pre, o_bytes);
        for (int c = tid; c < P * CPR; c += 256) {
          const int pix = c / CPR, ch = c - (c / CPR) * CPR;
          const int n = n0 + ch * EPC;
          uint4 v = *(const uint4 *)(smem + pix * ORS + ch * 16);
          const uint32_t vo = n < g.N ? (uint32_t)((obase + pix) * g.N + n) * (uint32_t)sizeof(TO) : OOB;
          if (dst_pre) __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v), qrr, (int)vo, 0, ST_NT);
          if (dst) {
            if (relu_y) v = relu8(v);
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, v), orr, (int)vo, 0, ST_NT);
          }
        }
      }
      if constexpr (PASS == 0 && sizeof(TO) == 2) {
        if ((act || dst_pre) && g.pool) {
          // Fused 2x2 stride-2 max-pool of y (the VALID windows, Hp = Ho / 2,
          // Wp = Wo / 2).  The tile starts at an even row, so its windows
          // are whole; each thread reads a window's four staged chunks and
          // keeps Owl's max (owl_ndarray_pool_impl.h:9-92 with the stub's
          // max macros): acc = -inf, then "if (acc < t) acc = t" over
          // (h, w), (h, w + 1), (h + 1, w), (h + 1, w + 1), compared in FP32
          // -- the first strict maximum, NaN never taken.
          const int Wp = g.Wo >> 1, Hp = g.Ho >> 1;
          const int PQ = (TRv >> 1) * Wp;
          const uint32_t p_bytes = (uint32_t)((int64_t)g.B * Hp * Wp * g.N * sizeof(TO));
          const __amdgpu_buffer_rsrc_t prr = make_rsrc(g.pool, p_bytes);
          const int pbase = (b * Hp + (i0 >> 1)) * Wp;
          for (int c = tid; c < PQ * CPR; c += 256) {
            const int pq = c / CPR, ch = c - (c / CPR) * CPR;
            const int pr = pq / Wp, pc = pq - (pq / Wp) * Wp;
            const char *s0 = smem + (2 * pr * g.Wo + 2 * pc) * ORS + ch * 16;
            const int wofs[4] = {0, ORS, g.Wo * ORS, (g.Wo + 1) * ORS};
            float m[EPC];
#pragma unroll
            for (int e = 0; e < EPC; ++e) m[e] = -__builtin_inff();
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              uint4 q = *(const uint4 *)(s0 + wofs[k]);
              if (relu_y) q = relu8(q);
              TO tv[EPC];
              __builtin_memcpy(tv, &q, 16);
#pragma unroll
              for (int e = 0; e < EPC; ++e) {
                const float v = (float)tv[e];
                if (m[e] < v) m[e] = v;
              }
            }
            TO o[EPC];
#pragma unroll
            for (int e = 0; e < EPC; ++e) o[e] = (TO)m[e];
            uint4 ov;
            __builtin_memcpy(&ov, o, 16);
            const int n = n0 + ch * EPC;
            const uint32_t vo = n < g.N ? (uint32_t)((pbase + pq) * g.N + n) * (uint32_t)sizeof(TO) : OOB;
            __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(v4u32, ov), prr, (int)vo, 0, ST_NT);
          }
        }
      }
    } else if (dst) {
      const bool relu = epi && act && g.act == OWL_AMD_ACT_RELU;
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int q = 0; q < NR / 4; ++q) {
          float bv[4];
          bias4(j, q, bv);
          const int n = n0 + chn(j) + rofs(4 * q);
#pragma unroll
          for (int i = 0; i < FM; ++i) {
            const int pp = pix(i);
            // N % 8 == 0: a run of four channels is wholly inside or outside
            const bool ok = pp < P && n < g.N;
            const uint32_t vo = ok ? (uint32_t)((obase + pp) * g.N + n) * (uint32_t)sizeof(TO) : OOB;
            TO o[4];
            run4(i, j, q, bv, epi && g.bias, relu, o);
            if constexpr (sizeof(TO) == 2) {
              v2u32 w2;
              __builtin_memcpy(&w2, o, 8);
              __builtin_amdgcn_raw_buffer_store_b64(w2, orr, (int)vo, 0, ST_NT);
            } else {
              v4u32 w4;
              __builtin_memcpy(&w4, o, 16);
              __builtin_amdgcn_raw_buffer_store_b128(w4, orr, (int)vo, 0, ST_NT);
            }
          }
        }
    }
  };
  if constexpr (sizeof(TO) == 2 && CAP * ORS <= LDS_BYTES) {
    // one staging pass also when the pre-activation is kept
    TO *pre = PASS == 0 ? (TO *)g.pre : nullptr;
    emit(out, pre == nullptr, pre);
  } else {
    if (PASS == 0 && g.pre) {
      emit((TO *)g.pre, false, nullptr);
      __syncthreads();  // the staging area is reused for y
    }
    emit(out, true, nullptr);
  }
}

// Pixel capacity of a halo tile of channel slice bn: 128-wide slices run 4
// waves of 2 x 4 32-edge fragments (256 pixels) or 2 x 2 waves of 7 x 4
// 16-edge ones (224 pixels); 64- and 32-wide slices 4 waves along the
// pixels, of 4 x 2 (4 x 1) 32-edge fragments (512 pixels) or of 7 x 4
// (7 x 2) 16-edge ones (448 pixels).  At most 112 or 128 accumulators per
// lane.
static int halo_cap(int bn, int ms) { return bn <= 64 ? (ms == 32 ? 512 : 448) : (ms == 32 ? 256 : 224); }
// Default channel slice for N output channels: 32 for at most 32 (instead
// of a 64-wide tile half on padding), 64 for at most 64, else 128.
static int halo_bn(int N) { return N <= 32 ? 32 : N <= 64 ? 64 : 128; }

// Whole output rows per tile for a capacity (0 when one row does not fit),
// and whether their input halo fits the LDS (704 pixels).
static int halo_rows(const HaloGeom &g, int cap) {
  const int tr = cap / g.Wo;
  return tr > g.Ho ? g.Ho : tr;
}
static bool halo_rows_fit(const HaloGeom &g, int tr) {
  if (tr < 1) return false;
  const int HR = (tr - 1) + (g.kh - 1) * g.rh + 1, HC = (g.Wo - 1) + (g.kw - 1) * g.rw + 1;
  return HR * HC <= 2 * 352;
}

// Geometry of the halo kernel for a problem, or false if it does not fit.
// The fragment edge is the one whose tiles pad the output least (ties go to
// 32): a 56-pixel row fills 16-edge tiles exactly (224 = 4 rows, 448 = 8),
// while 32-edge tiles leave 12.5% of their MFMA work on padding.
// OWL_AMD_HALO_MS = 16 or 32 forces an edge where it fits (tests, tuning).
template <int PASS>
static bool halo_geom(const ConvParams &p, int in_bytes, int out_bytes, HaloGeom &g, int force_ms = 0) {
  if (in_bytes != 2) return false;
  if (p.sh != 1 || p.sw != 1) return false;
  const bool d3 = p.D != 1 || p.kd != 1 || p.Do != 1;
  if (!d3) {
    if (PASS == 0) {
      g.Hi = p.H; g.Wi = p.W; g.C = p.Ci; g.Ho = p.Ho; g.Wo = p.Wo; g.N = p.Co;
      g.ph = p.ph; g.pw = p.pw;
    } else {
      g.Hi = p.Ho; g.Wi = p.Wo; g.C = p.Co; g.Ho = p.H; g.Wo = p.W; g.N = p.Ci;
      g.ph = (p.kh - 1) * p.rh - p.ph;
      g.pw = (p.kw - 1) * p.rw - p.pw;
      if (g.ph < 0 || g.pw < 0) return false;
    }
    g.kh = p.kh; g.kw = p.kw; g.rh = p.rh; g.rw = p.rw;
    g.KA = 1; g.ra = 1; g.pa = 0; g.HA = 1; g.HOA = 1;
    g.B = p.B; g.IB = p.B;
  } else {
    // 3-D, unit strides: (w, d) planes, the h taps as outer blocks
    if (p.sd != 1 || p.pool) return false;
    if (PASS == 0) {
      g.Hi = p.W; g.Wi = p.D; g.C = p.Ci; g.Ho = p.Wo; g.Wo = p.Do; g.N = p.Co;
      g.ph = p.pw; g.pw = p.pd; g.pa = p.ph;
      g.HA = p.H; g.HOA = p.Ho;
    } else {
      g.Hi = p.Wo; g.Wi = p.Do; g.C = p.Co; g.Ho = p.W; g.Wo = p.D; g.N = p.Ci;
      g.ph = (p.kw - 1) * p.rw - p.pw;
      g.pw = (p.kd - 1) * p.rd - p.pd;
      g.pa = (p.kh - 1) * p.rh - p.ph;
      if (g.ph < 0 || g.pw < 0 || g.pa < 0) return false;
      g.HA = p.Ho; g.HOA = p.H;
    }
    g.kh = p.kw; g.kw = p.kd; g.rh = p.rw; g.rw = p.rd;
    g.KA = p.kh; g.ra = p.rh;
    g.B = p.B * g.HOA; g.IB = p.B * g.HA;
    if ((int64_t)p.B * g.HOA >= (1LL << 31) || (int64_t)p.B * g.HA >= (1LL << 31)) return false;
  }
  g.wCi = p.Ci; g.wCo = p.Co;
  g.bias = PASS == 0 ? (const float *)p.bias : nullptr;
  g.pre = PASS == 0 ? p.pre : nullptr;
  g.act = PASS == 0 ? p.act : OWL_AMD_ACT_NONE;
  g.pool = PASS == 0 ? p.pool : nullptr;
  g.sa = g.sb = nullptr;
  // the fused pooling reads the staged 16-bit tile
  if (g.pool && out_bytes != 2) return false;
  // channel counts that are not multiples of 32 run the general instance
  if (g.C % 8 != 0 || g.N % 8 != 0 || g.Wo > 256 || g.Ho < 1) return false;
  if ((int64_t)g.IB * g.Hi * g.Wi * g.C * 2 >= (1LL << 31) || (int64_t)g.B * g.Ho * g.Wo * g.N * out_bytes >= (1LL << 31) ||
      (int64_t)g.KA * g.kh * g.kw * g.C * g.N * 2 >= (1LL << 31))
    return false;  // 32-bit buffer offsets
  const char *fe = getenv("OWL_AMD_HALO_MS");
  const int forced = force_ms ? force_ms : fe ? atoi(fe) : 0;
  g.ms = 0;
  double best = -1.0;
  // Candidates (slice, edge), scored by the share of the tiles' MFMA work
  // on real pixels and channels; earlier candidates win ties.  The
  // backward input also tries 64-wide slices for N > 64, first: where their
  // 448-pixel tiles fill the rows as well as the 224-pixel ones they ran 2-3%
  // faster (BF16 backward input, 28- and 56-pixel rows, N 128 and 256; same
  // bits); on the forward they ran 0-7% slower.
  const int bn0 = halo_bn(g.N);
  const int nbn = PASS == 1 && g.N > 64 ? 2 : 1;
  const int bns[2] = {nbn == 2 ? 64 : bn0, bn0};
  for (int bi = 0; bi < nbn; ++bi) {
    const int bn = bns[bi];
    const double ch_use = (double)g.N / ((double)((g.N + bn - 1) / bn) * bn);
    for (int ms : {32, 16}) {
      const int cap = halo_cap(bn, ms);
      int tr = halo_rows(g, cap);
      // fused pooling: tiles start at even rows (whole 2x2 windows)
      if (g.pool && tr < g.Ho) tr &= ~1;
      if (!halo_rows_fit(g, tr)) continue;
      const double use = ch_use * (double)g.Ho * g.Wo / ((double)((g.Ho + tr - 1) / tr) * cap);
      // a forced edge (tests, tuning) ranks first, then the default slice;
      // an edge that does not fit falls back to one that does
      const double score = forced == 16 || forced == 32 ? 2.0 * (ms == forced) + (bn == bn0) : use;
      if (score > best + 1e-9) {
        g.ms = ms;
        g.bn = bn;
        g.TR = tr;
        best = score;
      }
    }
  }
  return g.ms != 0;
}

bool halo_fits(const ConvParams &p, int pass, int in_bytes, int out_bytes) {
  HaloGeom g;
  return pass == 0 ? halo_geom<0>(p, in_bytes, out_bytes, g) : halo_geom<1>(p, in_bytes, out_bytes, g);
}

// Share of a halo launch's MFMA work on real output pixels and channels (0
// when the problem does not fit the halo kernel): tiles hold whole rows of one
// image, so small images leave most of a tile empty (7x7: 49 of 224 pixels).
double halo_fill(const ConvParams &p, int pass, int in_bytes, int out_bytes) {
  HaloGeom g;
  const bool ok = pass == 0 ? halo_geom<0>(p, in_bytes, out_bytes, g) : halo_geom<1>(p, in_bytes, out_bytes, g);
  if (!ok) return 0.0;
  const double ch = (double)g.N / ((double)((g.N + g.bn - 1) / g.bn) * g.bn);
  return ch * (double)g.Ho * g.Wo / ((double)((g.Ho + g.TR - 1) / g.TR) * halo_cap(g.bn, g.ms));
}

// Runs the halo kernel if the problem fits it; returns false otherwise.
// The forward reads the weights in place ([tap][ci][co], transposing LDS
// reads); the backward input reads them as [tap][n = ci][k = co] rows.
template <int PASS, typename TI, typename TO>
bool run_halo(const ConvParams &p, hipStream_t st) {
  HaloGeom g;
  if (!halo_geom<PASS>(p, (int)sizeof(TI), (int)sizeof(TO), g)) return false;
  const int BN = g.bn;
  const int64_t blocks = (int64_t)g.B * ((g.Ho + g.TR - 1) / g.TR) * ((g.N + BN - 1) / BN);
  const TI *in = (const TI *)p.a;
  const TI *w = (const TI *)p.b;
  TO *out = (TO *)p.c;
  constexpr bool KROWS = PASS == 1;
  dim3 grid((unsigned)blocks), blk(256);
  KTimer kt(PASS, "conv_halo_kernel", st);
  const bool gen = p.D != 1 || p.kd != 1 || p.Do != 1 || g.C % 32 != 0;
  if (gen) {
    if (g.ms == 16) {
      if (BN == 32) launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 2, 1, KROWS, false, true>), grid, blk, st, g, in, w, out);
      else if (BN == 64) launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 4, 1, KROWS, false, true>), grid, blk, st, g, in, w, out);
      else launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 4, 2, KROWS, false, true>), grid, blk, st, g, in, w, out);
    } else {
      if (BN == 32) launch_k((conv_halo_kernel<PASS, TI, TO, 32, 4, 1, 1, KROWS, false, true>), grid, blk, st, g, in, w, out);
      else if (BN == 64) launch_k((conv_halo_kernel<PASS, TI, TO, 32, 4, 2, 1, KROWS, false, true>), grid, blk, st, g, in, w, out);
      else launch_k((conv_halo_kernel<PASS, TI, TO, 32, 2, 4, 1, KROWS, false, true>), grid, blk, st, g, in, w, out);
    }
    return true;
  }
  if (g.ms == 16) {
    if (BN == 32) launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 2, 1, KROWS>), grid, blk, st, g, in, w, out);
    else if (BN == 64) launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 4, 1, KROWS>), grid, blk, st, g, in, w, out);
    else launch_k((conv_halo_kernel<PASS, TI, TO, 16, 7, 4, 2, KROWS>), grid, blk, st, g, in, w, out);
  } else {
    if (BN == 32) launch_k((conv_halo_kernel<PASS, TI, TO, 32, 4, 1, 1, KROWS>), grid, blk, st, g, in, w, out);
    else if (BN == 64) launch_k((conv_halo_kernel<PASS, TI, TO, 32, 4, 2, 1, KROWS>), grid, blk, st, g, in, w, out);
    else launch_k((conv_halo_kernel<PASS, TI, TO, 32, 2, 4, 1, KROWS>), grid, blk, st, g, in, w, out);
  }
  return true;
}
// FP8: the layer seen as one over pairs of input channels (Ci / 2 of them,
// 16-bit units), 32-edge tiles of 256 pixels x 64 channels.
static bool halo_geom_fp8(const ConvParams &p, int out_bytes, HaloGeom &g) {
  if (p.Ci % 64 != 0 || p.Co % 8 != 0) return false;
  if (p.D != 1 || p.kd != 1 || p.Do != 1) return false;  // 2-D layers only
  ConvParams q = p;
  q.Ci = p.Ci / 2;
  q.bias = nullptr;
  q.pre = nullptr;
  q.act = OWL_AMD_ACT_NONE;
  q.pool = nullptr;
  if (!halo_geom<0>(q, 2, out_bytes, g, 32) || g.ms != 32) return false;
  g.TR = 256 / g.Wo;
  if (g.TR > g.Ho) g.TR = g.Ho;
  if (!halo_rows_fit(g, g.TR)) return false;
  g.sa = p.scale_a;
  g.sb = p.scale_b;
  return true;
}

bool halo_fits_fp8(const ConvParams &p, int out_bytes) {
  HaloGeom g;
  return halo_geom_fp8(p, out_bytes, g);
}

template <typename TO>
bool run_halo_fp8(const ConvParams &p, const void *wp, hipStream_t st) {
  HaloGeom g;
  if (!halo_geom_fp8(p, (int)sizeof(TO), g)) return false;
  const int64_t blocks = (int64_t)g.B * ((g.Ho + g.TR - 1) / g.TR) * ((g.N + 63) / 64);
  KTimer kt(0, "conv_halo_kernel", st);
  launch_k((conv_halo_kernel<0, bf16_t, TO, 32, 2, 2, 1, true, true>), dim3((unsigned)blocks), dim3(256), st, g,
           (const bf16_t *)p.a, (const bf16_t *)wp, (TO *)p.c);
  return true;
}
template bool run_halo_fp8<float>(const ConvParams &, const void *, hipStream_t);
template bool run_halo_fp8<bf16_t>(const ConvParams &, const void *, hipStream_t);

#define OWL_HALO(PASS, TI, TO) template bool run_halo<PASS, TI, TO>(const ConvParams &, hipStream_t);
OWL_HALO(0, bf16_t, bf16_t)
OWL_HALO(0, bf16_t, float)
OWL_HALO(1, bf16_t, bf16_t)
OWL_HALO(1, bf16_t, float)
OWL_HALO(0, f16_t, f16_t)
OWL_HALO(0, f16_t, float)
OWL_HALO(1, f16_t, f16_t)
OWL_HALO(1, f16_t, float)
#undef OWL_HALO

template <int PASS, typename TI, typename TO, int FM, int FN, int WGM, int WGN>
static void launch_fast(const ConvParams &p, int pd, hipStream_t st) {
  constexpr int BM = WGM * FM * 32, BN = WGN * FN * 32;
  int64_t tiles = ((p.M + BM - 1) / BM) * (int64_t)((p.N + BN - 1) / BN);
  dim3 g((unsigned)tiles), b(256);
  if (pd <= 1) launch_k((conv_fast_kernel<PASS, TI, TO, FM, FN, WGM, WGN, 1>), g, b, st, p);
  else if (pd == 2) launch_k((conv_fast_kernel<PASS, TI, TO, FM, FN, WGM, WGN, 2>), g, b, st, p);
  else launch_k((conv_fast_kernel<PASS, TI, TO, FM, FN, WGM, WGN, 3>), g, b, st, p);
}

// Tile shapes: 0 -> 256 x 64, 1 -> 256 x 128, 2 -> 128 x 256, 3 -> 128 x 128.
template <int PASS, typename TI, typename TO>
void run_fast(const ConvParams &p, int shape, int pd, hipStream_t st) {
  switch (shape) {
    case 0: launch_fast<PASS, TI, TO, 2, 2, 4, 1>(p, pd, st); break;
    case 1: launch_fast<PASS, TI, TO, 2, 4, 4, 1>(p, pd, st); break;
    case 2: launch_fast<PASS, TI, TO, 2, 4, 2, 2>(p, pd, st); break;
    default: launch_fast<PASS, TI, TO, 2, 2, 2, 2>(p, pd, st); break;
  }
}

int fast_bk(int elem_bytes) { return elem_bytes == 4 ? 16 : 32; }

#define OWL_FAST(PASS, TI, TO) template void run_fast<PASS, TI, TO>(const ConvParams &, int, int, hipStream_t);
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

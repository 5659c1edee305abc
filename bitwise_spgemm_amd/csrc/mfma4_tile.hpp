// mfma4_tile.hpp — the direct-to-LDS 16-bit MFMA tile of the 0/1 bitmap SpMM
// (shared by the translation units that instantiate it with different
// register budgets) and the VALU construction of the 0/1 A fragment.
#pragma once

#include "mfma_common.hpp"

namespace bws {
namespace {

// 0/1 A fragment from a bitmap byte (see spmm_mfma.hip): with the word
// bit-reversed at staging, K slot 8q + e of the 32-wide step is bit e of byte
// q; for z = x | (x << 15), (z & (0x10001 << 2p)) * (ONE >> 2p) puts ONE in
// the low half iff bit 2p is set and in the high half iff bit 2p + 1 is set
// (ONE, the 16-bit 1.0, has at least 7 trailing zero bits, so ONE >> 2p is
// exact for p < 4).
template <int PREC>
__device__ __forceinline__ uint4 a_frag_bits(uint32_t rw, int q) {
  constexpr uint32_t ONE = PREC == 2 ? 0x3C00u : 0x3F80u;
  const uint32_t x = __builtin_amdgcn_ubfe(rw, 8 * q, 8);
  const uint32_t z = x | (x << 15);
  return make_uint4((uint32_t)__umul24(z & 0x00010001u, ONE),
                    (uint32_t)__umul24(z & 0x00040004u, ONE >> 2),
                    (uint32_t)__umul24(z & 0x00100010u, ONE >> 4),
                    (uint32_t)__umul24(z & 0x00400040u, ONE >> 6));
}

// ---------------------------------------------------------------------------
// Direct-to-LDS variant for 16-bit B (N and ldb multiples of 8, 16-byte
// aligned): no register staging at all.
//  - B stages of 64 K rows are copied row-major into LDS by
//    global_load_lds_dwordx4 (each wave instruction writes 1 KiB), NBUF stages
//    in flight.  Within a row, 32-byte unit c is stored at slot c ^ f(k) with
//    f(k) = (k & 3) | ((k >> 1) & 4), so the transposing reads below hit 8
//    different slots (all 64 banks) per 32 lanes.
//  - B fragments come from ds_read_b64_tr_b16: in each group of 16 lanes, lane
//    i addresses row kb + i/4, columns n + 4(i%4) .. +3, and receives column
//    n + i of rows kb .. kb+3; two reads give the 8 K values of an MFMA operand
//    (K slots 8q .. 8q+7 for lane group q, the same slots the A operand uses).
//  - bitmap words are copied by global_load_lds_dword (4 bytes per lane);
//    masking of bits beyond K and the bit reversal happen when they are read.
// Rows beyond K load row K-1 and columns beyond N load columns N-8..N-1: that
// data meets zero A bits or is never stored.  B must be finite (0 x Inf = NaN
// in a dense contraction, as in any GEMM).
// ---------------------------------------------------------------------------

template <int WM, int WN, int NBUF, int KS_>
struct Cfg4 {
  static constexpr int BM = 2 * WM, BN = 2 * WN, KS = KS_, KW = KS_ / 32;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int ROWB = BN * 2;                    // bytes per B row in LDS
  static constexpr int SB = KS * ROWB;                   // bytes per B stage
  static constexpr int NB_W = SB / 1024 / 4;             // B DMA instructions per wave per stage
  static constexpr int NA_W = BM * KW / 64 / 4;          // word DMA instructions per wave
  static constexpr int PER_STAGE = NB_W + NA_W;
  static constexpr int LANES_PER_ROW = ROWB / 16;        // 32 (BN 256) or 16 (BN 128)
};

template <int PREC, int WM, int WN, int NBUF, int KS_, bool SKIP = true>
__device__ __forceinline__ void mfma4_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K,
                                           int64_t ldw, const uint16_t* __restrict__ B,
                                           int64_t ldb, int64_t N, float* __restrict__ C,
                                           int64_t ldc, float alpha, float beta, int64_t tile,
                                           int64_t row_blocks) {
  using G = Cfg4<WM, WN, NBUF, KS_>;
  constexpr int BM = G::BM, BN = G::BN, KS = G::KS, KW = G::KW, TM = G::TM, TN = G::TN;
  constexpr int ROWB = G::ROWB, SB = G::SB, NB_W = G::NB_W, NA_W = G::NA_W;
  constexpr int LPR = G::LANES_PER_ROW;
  static_assert(NB_W >= 1 && NA_W >= 1, "stage must give every wave DMA work");
  __shared__ __attribute__((aligned(1024))) uint8_t sB[NBUF][SB];
  __shared__ __attribute__((aligned(256))) uint32_t sA[NBUF][BM * KW];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t m0 = rb * BM, n0 = cb * BN;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;
  const int64_t nst = (K + KS - 1) / KS;
  const int64_t wpr = (K + 31) / 32;

  // per-lane DMA sources (the stage only moves the row index)
  const uint16_t* bsrc[NB_W];
  int brow[NB_W];
#pragma unroll
  for (int t = 0; t < NB_W; ++t) {
    const int g = wave * NB_W + t;                 // instruction index within the stage
    const int r = g * (1024 / ROWB) + lane / LPR;  // B row within the stage
    const int u = lane % LPR;                      // 16-byte piece of the LDS row
    const int slot = u >> 1, half = u & 1;
    const int c = slot ^ ((r & 3) | ((r >> 1) & 4));
    int64_t col = n0 + 16 * c + 8 * half;
    if (col + 8 > N) col = N - 8;
    bsrc[t] = B + col;
    brow[t] = r;
  }
  const uint32_t* asrc[NA_W];
  int aw[NA_W];
#pragma unroll
  for (int t = 0; t < NA_W; ++t) {
    const int idx = (wave * NA_W + t) * 64 + lane;
    int64_t row = m0 + idx / KW;
    if (row > M - 1) row = M - 1;
    asrc[t] = W + row * ldw;
    aw[t] = idx % KW;
  }
  const uint32_t sB0 = lds_u32(&sB[0][0]), sA0 = lds_u32(&sA[0][0]);
  // stage 0 sources; a full stage only adds the stage's K offset, the last
  // (partial) stage clamps rows and words
  const uint16_t* bp0[NB_W];
  const uint32_t* ap0[NA_W];
#pragma unroll
  for (int t = 0; t < NB_W; ++t) bp0[t] = bsrc[t] + (int64_t)brow[t] * ldb;
#pragma unroll
  for (int t = 0; t < NA_W; ++t) ap0[t] = asrc[t] + aw[t];
  auto buf_of = [](int64_t st) -> int {
    return (NBUF & (NBUF - 1)) == 0 ? (int)(st & (NBUF - 1)) : (int)(st % NBUF);
  };

  auto issue = [&](int64_t st) {
    const int buf = buf_of(st);
    if ((st + 1) * KS <= K) {
      const int64_t boff = st * KS * ldb;
#pragma unroll
      for (int t = 0; t < NB_W; ++t)
        dma16(bp0[t] + boff, sB0 + (uint32_t)(buf * SB + (wave * NB_W + t) * 1024));
#pragma unroll
      for (int t = 0; t < NA_W; ++t)
        dma4(ap0[t] + st * KW, sA0 + (uint32_t)(buf * BM * KW * 4 + (wave * NA_W + t) * 256));
      return;
    }
#pragma unroll
    for (int t = 0; t < NB_W; ++t) {
      int64_t k = st * KS + brow[t];
      if (k > K - 1) k = K - 1;
      dma16(bsrc[t] + k * ldb, sB0 + (uint32_t)(buf * SB + (wave * NB_W + t) * 1024));
    }
#pragma unroll
    for (int t = 0; t < NA_W; ++t) {
      int64_t w = st * KW + aw[t];
      if (w > wpr - 1) w = wpr - 1;
      dma4(asrc[t] + w, sA0 + (uint32_t)(buf * BM * KW * 4 + (wave * NA_W + t) * 256));
    }
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

#pragma unroll
  for (int p = 0; p < NBUF - 1; ++p)
    if (p < nst) issue(p);

  // lane constants of the transposing reads: row 8q + i/4 (+ 4h + 32 ks),
  // 8-byte piece i%4 of the slot
  const int i4 = j >> 2;
  const int fsw = i4 | ((q & 1) << 2);
  const uint32_t rd_lane = (uint32_t)((8 * q + i4) * ROWB + 8 * (j & 3));
  const int c0 = (WN * wn) / 16;

  for (int64_t st = 0; st < nst; ++st) {
    const int buf = buf_of(st);
    if (st + NBUF - 2 < nst) wait_vmcnt<(NBUF - 2) * G::PER_STAGE>();
    else wait_vmcnt<0>();
    __syncthreads();
    if (st + NBUF - 1 < nst) issue(st + NBUF - 1);

    uint32_t wr[TM][KW];
    uint32_t any = 0;
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const uint32_t* src = &sA[buf][(WM * wm + 16 * a + j) * KW];
      uint32_t v[KW];
      if constexpr (KW == 4) {
        const uint4 x = *reinterpret_cast<const uint4*>(src);
        v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
      } else {
        const uint2 x = *reinterpret_cast<const uint2*>(src);
        v[0] = x.x, v[1] = x.y;
      }
#pragma unroll
      for (int x = 0; x < KW; ++x) {
        wr[a][x] = __brev(v[x] & valid_bits(st * KW + x, K));
        any |= wr[a][x];
      }
    }
    bool run = true;
    if constexpr (SKIP) run = __ballot(any != 0u) != 0ull;
    if (run) {
      // the wave in its MFMA phase issues first: of the two waves on a SIMD,
      // the other is copying, at the barrier or building A fragments
      // (98.8 µs against 109.5 µs at 4096³ without it)
      __builtin_amdgcn_s_setprio(1);
      const uint32_t sbase = sB0 + (uint32_t)(buf * SB) + rd_lane;
      // all B fragments of the stage are requested first (two K steps at
      // KS = 64), then the MFMAs run as they arrive
      constexpr int KP = KW < 2 ? KW : 2;  // K steps whose fragments are held at once
#pragma unroll
      for (int k0 = 0; k0 < KW; k0 += KP) {
        uint4 bf[KP][TN];
#pragma unroll
        for (int d = 0; d < KP; ++d)
#pragma unroll
          for (int b = 0; b < TN; ++b) {
            const uint32_t a0 = sbase + (uint32_t)(32 * ((c0 + b) ^ fsw)) +
                                (uint32_t)(32 * (k0 + d) * ROWB);
            const uint2 lo = tr_read(a0);
            const uint2 hi = tr_read(a0 + 4 * ROWB);
            bf[d][b] = make_uint4(lo.x, lo.y, hi.x, hi.y);
          }
        if constexpr (TN == 8) {
          // 64 x 128 wave tile: the A fragment of the next row tile is built
          // while this row tile's 8 MFMAs run, one MFMA per two VALU ops
          uint4 afc = a_frag_bits<PREC>(wr[0][k0], q);
#pragma unroll
          for (int d = 0; d < KP; ++d)
#pragma unroll
            for (int a = 0; a < TM; ++a) {
              const int an = (a + 1 < TM) ? a + 1 : 0;
              const int dn = (a + 1 < TM) ? d : d + 1;
              uint4 afn = afc;
              if (dn < KP) afn = a_frag_bits<PREC>(wr[an][k0 + dn], q);
#pragma unroll
              for (int b = 0; b < TN; ++b) acc[a][b] = mfma<PREC>(afc, bf[d][b], acc[a][b]);
#pragma unroll
              for (int b = 0; b < TN; ++b) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
                __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);  // two VALU
              }
              afc = afn;
            }
        } else {
#pragma unroll
          for (int d = 0; d < KP; ++d)
#pragma unroll
            for (int a = 0; a < TM; ++a) {
              const uint4 af = a_frag_bits<PREC>(wr[a][k0 + d], q);
#pragma unroll
              for (int b = 0; b < TN; ++b) acc[a][b] = mfma<PREC>(af, bf[d][b], acc[a][b]);
            }
        }
      }
      __builtin_amdgcn_s_setprio(0);
    }
  }

#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) {
      const int64_t col = n0 + WN * wn + 16 * b + j;
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + WM * wm + 16 * a + 4 * q + e;
        if (row < M) {
          float* cp = C + row * ldc + col;
          const float v = alpha * acc[a][b][e];
          *cp = beta != 0.0f ? fmaf(beta, *cp, v) : v;
        }
      }
    }
}

}  // namespace
}  // namespace bws

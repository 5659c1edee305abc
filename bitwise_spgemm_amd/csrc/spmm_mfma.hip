// spmm_mfma.hip — 16-bit MFMA tile GEMM for a 0/1 bitmap A (gfx950,
// v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulate).  Three kernels share the
// operand layout below:
//  - k_spmm_mfma4 (default for 16-bit B): workgroup 128 x 128, 256 x 128 or
//    128 x 256 of C, B copied straight into LDS and read with the transposing
//    LDS read; see the comment above mfma4_tile.
//  - k_spmm_mfma3: the same tiles with B loaded into registers, transposed with
//    v_perm_b32 and written to LDS (FP32 B rounded on the fly, BF16X2, and B
//    layouts the direct copy cannot take).
//  - k_spmm_mfma (BWS_MFMA_LUT): 64 x 64 per wave, A fragments from an LDS
//    table, kept as a reference point for the other two.
// Operand layout used for both operands: lane l covers row/column (l % 16)
// and the K slots 8*(l/16) .. 8*(l/16)+7 of the 32-wide step; since A and B use
// the same K assignment the product does not depend on the hardware's K order,
// and all kernels accumulate the same K steps in the same order (their results
// are bit-identical).  The accumulator layout is D[4*(l/16) + e][l % 16].
// A stage whose bitmap block (a wave's rows x the stage's K columns) is empty
// skips its MFMAs.
//
// Below, the table-lookup kernel: workgroup = 4 waves (2 x 2), C tile
// 128 x 128, each wave 64 x 64 (4 x 4 MFMA tiles), K stage 128 rows.
//  - A operand: the byte of the bitmap word that a lane's 8 K-slots cover
//    indexes a 256-entry LDS table of eight 16-bit 0/1 values.
//  - B operand: staged in LDS transposed (N-major, K contiguous, 272-byte rows)
//    so a lane's 8 K-consecutive values are one 16-byte read.  Each thread
//    loads an 8(K) x 8(N) block with 16-byte global loads, transposes it in
//    registers with v_perm_b32 and writes eight 16-byte LDS rows.
//
// BF16X2: B = hi + lo with hi = bf16(b), lo = bf16(b - hi); both are staged
// and both products accumulate into the same FP32 accumulators.
#include "mfma4_tile.hpp"

namespace bws {
namespace {


constexpr int BM = 128, BN = 128, KS = 128;
constexpr int LDK = KS + 8;  // 16-bit elements per LDS row: 272 bytes, 16-byte aligned




// 16-bit pair (lo from a, hi from b) of element x (0..7) of two packed rows.
__device__ __forceinline__ uint32_t pick2(const uint4& ra, const uint4& rb, int x) {
  const uint32_t da = x < 2 ? ra.x : x < 4 ? ra.y : x < 6 ? ra.z : ra.w;
  const uint32_t db = x < 2 ? rb.x : x < 4 ? rb.y : x < 6 ? rb.z : rb.w;
  // bytes 0-1 from da, bytes 2-3 from db; low or high half by parity of x
  return __builtin_amdgcn_perm(db, da, (x & 1) ? 0x07060302u : 0x05040100u);
}

// Raw B block of one thread: 8 K rows x 8 N columns in the input type.
template <int TB>
struct BBlock;
template <>
struct BBlock<0> {
  float4 v[8][2];
};
template <>
struct BBlock<1> {
  uint4 v[8];
};
template <>
struct BBlock<2> {
  uint4 v[8];
};

template <int TB>
__device__ __forceinline__ void load_block(BBlock<TB>& blk, const void* __restrict__ Bv,
                                           int64_t k0, int64_t n0, int64_t ldb, int64_t K,
                                           int64_t N, bool vec) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int64_t k = k0 + r;
    if constexpr (TB == 0) {
      const float* p = static_cast<const float*>(Bv) + k * ldb + n0;
      if (k < K && vec && n0 + 8 <= N) {
        blk.v[r][0] = *reinterpret_cast<const float4*>(p);
        blk.v[r][1] = *reinterpret_cast<const float4*>(p + 4);
      } else {
        float e[8];
#pragma unroll
        for (int x = 0; x < 8; ++x) e[x] = (k < K && n0 + x < N) ? p[x] : 0.0f;
        blk.v[r][0] = make_float4(e[0], e[1], e[2], e[3]);
        blk.v[r][1] = make_float4(e[4], e[5], e[6], e[7]);
      }
    } else {
      const uint16_t* p = static_cast<const uint16_t*>(Bv) + k * ldb + n0;
      if (k < K && vec && n0 + 8 <= N) {
        blk.v[r] = *reinterpret_cast<const uint4*>(p);
      } else {
        uint32_t d[4];
#pragma unroll
        for (int x = 0; x < 4; ++x) {
          const uint32_t lo = (k < K && n0 + 2 * x < N) ? p[2 * x] : 0u;
          const uint32_t hi = (k < K && n0 + 2 * x + 1 < N) ? p[2 * x + 1] : 0u;
          d[x] = lo | (hi << 16);
        }
        blk.v[r] = make_uint4(d[0], d[1], d[2], d[3]);
      }
    }
  }
}

// Convert one raw block to packed 16-bit rows (hi part, and lo part for BF16X2).
template <int PREC, int TB>
__device__ __forceinline__ void pack_block(const BBlock<TB>& blk, uint4 (&hi)[8], uint4 (&lo)[8]) {
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    if constexpr (TB == 0) {
      const float e[8] = {blk.v[r][0].x, blk.v[r][0].y, blk.v[r][0].z, blk.v[r][0].w,
                          blk.v[r][1].x, blk.v[r][1].y, blk.v[r][1].z, blk.v[r][1].w};
      uint32_t h[8], l[8];
#pragma unroll
      for (int x = 0; x < 8; ++x) {
        if constexpr (PREC == 2) {
          h[x] = f32_to_f16_rne(e[x]);
          l[x] = 0;
        } else {
          h[x] = f32_to_bf16_rne(e[x]);
          l[x] = PREC == 3 ? f32_to_bf16_rne(e[x] - bf16_to_f32(h[x])) : 0u;
        }
      }
      hi[r] = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16),
                         h[6] | (h[7] << 16));
      lo[r] = make_uint4(l[0] | (l[1] << 16), l[2] | (l[3] << 16), l[4] | (l[5] << 16),
                         l[6] | (l[7] << 16));
    } else {
      hi[r] = blk.v[r];
      lo[r] = make_uint4(0, 0, 0, 0);
    }
  }
}

// Transpose packed rows[k][n] (8 x 8 of 16-bit) into columns and store each
// column n as 8 K-consecutive values: dst + n * LDK.
__device__ __forceinline__ void store_transposed(const uint4 (&rows)[8], uint16_t* dst) {
#pragma unroll
  for (int x = 0; x < 8; ++x) {
    const uint4 col = make_uint4(pick2(rows[0], rows[1], x), pick2(rows[2], rows[3], x),
                                 pick2(rows[4], rows[5], x), pick2(rows[6], rows[7], x));
    *reinterpret_cast<uint4*>(dst + x * LDK) = col;
  }
}

// One 128 x 128 output tile of one problem (tile index column-block-major).
template <int PREC, int TB>
__device__ __forceinline__ void mfma_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K,
                                          int64_t ldw, const void* __restrict__ Bv, int64_t ldb,
                                          int64_t N, float* __restrict__ C, int64_t ldc,
                                          float alpha, float beta, int64_t tile,
                                          int64_t row_blocks, int vec) {
  constexpr int NS = PREC == 3 ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t sBt[2][NS][BN][LDK];
  __shared__ __attribute__((aligned(16))) uint32_t sA[2][BM][4];
  __shared__ uint4 sLut[256];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t m0 = rb * BM, n0 = cb * BN;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;

  {  // byte -> eight 16-bit 0/1 values; element e <-> bit (7 - e) (MSB-first)
    const uint32_t one = PREC == 2 ? 0x3C00u : 0x3F80u;
    const uint32_t b = (uint32_t)tid;
    uint32_t d[4];
#pragma unroll
    for (int p = 0; p < 4; ++p)
      d[p] = (((b >> (7 - 2 * p)) & 1u) * one) | ((((b >> (6 - 2 * p)) & 1u) * one) << 16);
    sLut[tid] = make_uint4(d[0], d[1], d[2], d[3]);
  }

  const int64_t nst = (K + KS - 1) / KS;
  const int64_t wpr = (K + 31) / 32;
  // staging roles: B block (kb, nb) = K rows 8kb..8kb+7, N columns 8nb..8nb+7;
  // A words: row tid/2, words 2*(tid&1), 2*(tid&1)+1 of the stage
  const int kb = tid & 15, nb = tid >> 4;
  const int arow = tid >> 1, apair = 2 * (tid & 1);
  const bool arow_ok = m0 + arow < M;

  BBlock<TB> blk;
  uint32_t aw[2];
  auto load_stage = [&](int64_t st) {
    load_block<TB>(blk, Bv, st * KS + 8 * kb, n0 + 8 * nb, ldb, K, N, vec != 0);
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int64_t w = 4 * st + apair + x;
      aw[x] = (arow_ok && w < wpr) ? (W[(m0 + arow) * ldw + w] & valid_bits(w, K)) : 0u;
    }
  };
  auto store_stage = [&](int buf) {
    uint4 hi[8], lo[8];
    pack_block<PREC, TB>(blk, hi, lo);
    store_transposed(hi, &sBt[buf][0][8 * nb][8 * kb]);
    if constexpr (NS == 2) store_transposed(lo, &sBt[buf][NS - 1][8 * nb][8 * kb]);
    sA[buf][arow][apair] = aw[0];
    sA[buf][arow][apair + 1] = aw[1];
  };

  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  load_stage(0);
  store_stage(0);
  __syncthreads();

  for (int64_t st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nst;
    if (more) load_stage(st + 1);

    uint4 wrow[4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
      wrow[a] = *reinterpret_cast<const uint4*>(&sA[buf][64 * wm + 16 * a + j][0]);
    uint32_t any = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a) any |= wrow[a].x | wrow[a].y | wrow[a].z | wrow[a].w;
    if (__ballot(any != 0u) != 0ull) {  // skip a stage whose 64 x 128 A block is empty
      // Fragments of K-step s+1 are read while the 16 MFMAs of step s run.
      auto frags = [&](int s, uint4 (&af)[4], uint4 (&bh)[4], uint4 (&bl)[4]) {
#pragma unroll
        for (int a = 0; a < 4; ++a) {
          const uint32_t w =
              s == 0 ? wrow[a].x : s == 1 ? wrow[a].y : s == 2 ? wrow[a].z : wrow[a].w;
          af[a] = sLut[(w >> (24 - 8 * q)) & 0xFFu];
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          bh[b] = *reinterpret_cast<const uint4*>(
              &sBt[buf][0][64 * wn + 16 * b + j][32 * s + 8 * q]);
          if constexpr (NS == 2)
            bl[b] = *reinterpret_cast<const uint4*>(
                &sBt[buf][NS - 1][64 * wn + 16 * b + j][32 * s + 8 * q]);
        }
      };
      uint4 af0[4], bh0[4], bl0[4], af1[4], bh1[4], bl1[4];
      auto mma = [&](const uint4 (&af)[4], const uint4 (&bh)[4], const uint4 (&bl)[4]) {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = 0; b < 4; ++b) {
            acc[a][b] = mfma<PREC>(af[a], bh[b], acc[a][b]);
            if constexpr (NS == 2) acc[a][b] = mfma<PREC>(af[a], bl[b], acc[a][b]);
          }
      };
      frags(0, af0, bh0, bl0);
      frags(1, af1, bh1, bl1);
      mma(af0, bh0, bl0);
      frags(2, af0, bh0, bl0);
      mma(af1, bh1, bl1);
      frags(3, af1, bh1, bl1);
      mma(af0, bh0, bl0);
      mma(af1, bh1, bl1);
    }
    if (more) store_stage(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      const int64_t col = n0 + 64 * wn + 16 * b + j;
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + 64 * wm + 16 * a + 4 * q + e;
        if (row < M) {
          float* c = C + row * ldc + col;
          const float v = alpha * acc[a][b][e];
          *c = beta != 0.0f ? fmaf(beta, *c, v) : v;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// Register-staged variant: 2 x 2 waves, each wave WM x WN of C (64 x 64,
// 64 x 128 or 128 x 64: at most 128 accumulator VGPRs, so two workgroups fit
// on a CU), so each LDS byte read feeds up to 2x more MFMAs than a 64 x 64
// wave tile.  K stage KS = 16384 / BN, so the B stage is always 256 blocks of
// 8 (K) x 8 (N), one per thread.  The 0/1 A fragment is built in VALU from the
// bitmap byte (a_frag_bits, no table lookup).
// LDS layout: B transposed [n][KS + 8] (row stride 9 or 17 sixteen-byte units,
// odd, so the 16 lanes of a fragment read and the 16 lanes of a transposed
// write each touch 16 distinct bank groups); A words [BM][KW].
// ---------------------------------------------------------------------------
template <int PREC, int WM, int WN>
struct Cfg3 {
  static constexpr int BM = 2 * WM, BN = 2 * WN;
  static constexpr int KS = 16384 / BN;  // 64 (BN 256) or 128 (BN 128)
  static constexpr int KW = KS / 32;
  static constexpr int LDK = KS + 8;
  static constexpr int NS = PREC == 3 ? 2 : 1;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int AWPT = BM * KW / 256;  // A words staged per thread
  static constexpr int KBN = KS / 8;          // 8-row K blocks per stage
};


template <int PREC, int TB, int WM, int WN>
__device__ __forceinline__ void mfma3_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K,
                                           int64_t ldw, const void* __restrict__ Bv, int64_t ldb,
                                           int64_t N, float* __restrict__ C, int64_t ldc,
                                           float alpha, float beta, int64_t tile,
                                           int64_t row_blocks, int vec) {
  using G = Cfg3<PREC, WM, WN>;
  constexpr int BM = G::BM, BN = G::BN, KS = G::KS, KW = G::KW, LDK = G::LDK, NS = G::NS;
  constexpr int TM = G::TM, TN = G::TN, AWPT = G::AWPT, KBN = G::KBN;
  __shared__ __attribute__((aligned(16))) uint16_t sBt[2][NS][BN][LDK];
  __shared__ __attribute__((aligned(16))) uint32_t sA[2][BM][KW];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t m0 = rb * BM, n0 = cb * BN;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;

  const int64_t nst = (K + KS - 1) / KS;
  const int64_t wpr = (K + 31) / 32;
  // staging roles: B block (kb, nb) = K rows 8kb..8kb+7 x N columns 8nb..8nb+7
  // (kb fastest: conflict-free transposed writes); A words: AWPT consecutive
  // words of row arow starting at word aoff of the stage
  const int kb = tid % KBN, nb = tid / KBN;
  constexpr int TPR = KW / AWPT;  // threads per A row
  const int arow = tid / TPR, aoff = (tid % TPR) * AWPT;
  const bool arow_ok = m0 + arow < M;

  BBlock<TB> blk;
  uint32_t aw[AWPT];
  auto load_stage = [&](int64_t st) {
    load_block<TB>(blk, Bv, st * KS + 8 * kb, n0 + 8 * nb, ldb, K, N, vec != 0);
#pragma unroll
    for (int x = 0; x < AWPT; ++x) {
      const int64_t w = KW * st + aoff + x;
      aw[x] = (arow_ok && w < wpr) ? (W[(m0 + arow) * ldw + w] & valid_bits(w, K)) : 0u;
    }
  };
  auto store_stage = [&](int buf) {
    uint4 hi[8], lo[8];
    pack_block<PREC, TB>(blk, hi, lo);
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      const uint4 col = make_uint4(pick2(hi[0], hi[1], x), pick2(hi[2], hi[3], x),
                                   pick2(hi[4], hi[5], x), pick2(hi[6], hi[7], x));
      *reinterpret_cast<uint4*>(&sBt[buf][0][8 * nb + x][8 * kb]) = col;
      if constexpr (NS == 2) {
        const uint4 cl = make_uint4(pick2(lo[0], lo[1], x), pick2(lo[2], lo[3], x),
                                    pick2(lo[4], lo[5], x), pick2(lo[6], lo[7], x));
        *reinterpret_cast<uint4*>(&sBt[buf][1][8 * nb + x][8 * kb]) = cl;
      }
    }
#pragma unroll
    for (int x = 0; x < AWPT; ++x) sA[buf][arow][aoff + x] = __brev(aw[x]);
  };

  f32x4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  load_stage(0);
  store_stage(0);
  __syncthreads();

  for (int64_t st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nst;
    if (more) load_stage(st + 1);

    uint32_t wr[TM][KW];
    uint32_t any = 0;
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const uint32_t* src = &sA[buf][WM * wm + 16 * a + j][0];
      if constexpr (KW == 4) {
        const uint4 v = *reinterpret_cast<const uint4*>(src);
        wr[a][0] = v.x, wr[a][1] = v.y, wr[a][2] = v.z, wr[a][3] = v.w;
      } else {
        const uint2 v = *reinterpret_cast<const uint2*>(src);
        wr[a][0] = v.x, wr[a][1] = v.y;
      }
#pragma unroll
      for (int s = 0; s < KW; ++s) any |= wr[a][s];
    }
    if (__ballot(any != 0u) != 0ull) {  // skip a stage whose WM x KS A block is empty
      // B fragments of K-step s+1 are read while step s multiplies; the A
      // fragment of row tile a+1 is built while row tile a multiplies.
      uint4 bh[2][TN], bl[2][TN];
      auto frags = [&](int s, int d) {
#pragma unroll
        for (int b = 0; b < TN; ++b) {
          bh[d][b] = *reinterpret_cast<const uint4*>(
              &sBt[buf][0][WN * wn + 16 * b + j][32 * s + 8 * q]);
          if constexpr (NS == 2)
            bl[d][b] = *reinterpret_cast<const uint4*>(
                &sBt[buf][1][WN * wn + 16 * b + j][32 * s + 8 * q]);
        }
      };
      frags(0, 0);
#pragma unroll
      for (int s = 0; s < KW; ++s) {
        const int d = s & 1;
        if (s + 1 < KW) frags(s + 1, d ^ 1);
#pragma unroll
        for (int a = 0; a < TM; ++a) {
          const uint4 af = a_frag_bits<PREC>(wr[a][s], q);
#pragma unroll
          for (int b = 0; b < TN; ++b) {
            acc[a][b] = mfma<PREC>(af, bh[d][b], acc[a][b]);
            if constexpr (NS == 2) acc[a][b] = mfma<PREC>(af, bl[d][b], acc[a][b]);
          }
        }
      }
    }
    if (more) store_stage(buf ^ 1);
    __syncthreads();
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
          float* c = C + row * ldc + col;
          const float v = alpha * acc[a][b][e];
          *c = beta != 0.0f ? fmaf(beta, *c, v) : v;
        }
      }
    }
}

template <int PREC, int WM, int WN, int NBUF, int KS>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma4(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
             int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
             int64_t per_xcd) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  mfma4_tile<PREC, WM, WN, NBUF, KS>(W, M, K, ldw, B, ldb, N, C, ldc, alpha, beta, tile,
                                     row_blocks);
}

template <int PREC, int WM, int WN, int NBUF, int KS>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma4_batch(const BatchEntry* __restrict__ entries, const int64_t* __restrict__ tile_off,
                   int count, int64_t per_xcd) {
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= tile_off[count]) return;
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const BatchEntry& e = entries[lo];
  mfma4_tile<PREC, WM, WN, NBUF, KS>(e.W, e.M, e.K, e.ldw, static_cast<const uint16_t*>(e.B),
                                     e.ldb, e.N, e.C, e.ldc, 1.0f, 0.0f, g - tile_off[lo],
                                     e.row_blocks);
}

template <int PREC, int TB, int WM, int WN>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma3(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const void* __restrict__ Bv, int64_t ldb, int64_t N, float* __restrict__ C,
             int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
             int64_t per_xcd, int vec) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  mfma3_tile<PREC, TB, WM, WN>(W, M, K, ldw, Bv, ldb, N, C, ldc, alpha, beta, tile, row_blocks,
                               vec);
}

template <int PREC, int TB, int WM, int WN>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma3_batch(const BatchEntry* __restrict__ entries, const int64_t* __restrict__ tile_off,
                   int count, int64_t per_xcd, int vec) {
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= tile_off[count]) return;
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const BatchEntry& e = entries[lo];
  mfma3_tile<PREC, TB, WM, WN>(e.W, e.M, e.K, e.ldw, e.B, e.ldb, e.N, e.C, e.ldc, 1.0f, 0.0f,
                               g - tile_off[lo], e.row_blocks, vec);
}

template <int PREC, int TB>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
            const void* __restrict__ Bv, int64_t ldb, int64_t N, float* __restrict__ C,
            int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
            int64_t per_xcd, int vec) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  mfma_tile<PREC, TB>(W, M, K, ldw, Bv, ldb, N, C, ldc, alpha, beta, tile, row_blocks, vec);
}

// Grouped launch over a problem table (see BatchEntry); vec applies to all.
template <int PREC, int TB>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma_batch(const BatchEntry* __restrict__ entries, const int64_t* __restrict__ tile_off,
                  int count, int64_t per_xcd, int vec) {
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= tile_off[count]) return;
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const BatchEntry& e = entries[lo];
  mfma_tile<PREC, TB>(e.W, e.M, e.K, e.ldw, e.B, e.ldb, e.N, e.C, e.ldc, 1.0f, 0.0f,
                      g - tile_off[lo], e.row_blocks, vec);
}

template <int PREC, int TB>
bws_status launch_mfma(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                       int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, BM);
  const int64_t col_blocks = ceil_div(N, BN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const size_t esz = TB == 0 ? 4 : 2;
  const int vec = (ldb % (16 / esz) == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
  hipLaunchKernelGGL((k_spmm_mfma<PREC, TB>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, B, ldb, N, C, ldc, alpha, beta,
                     row_blocks, col_blocks, per_xcd, vec);
  return launch_check();
}

template <int PREC, int TB, int WM, int WN>
bws_status launch_mfma3(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                        int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, 2 * WM);
  const int64_t col_blocks = ceil_div(N, 2 * WN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const size_t esz = TB == 0 ? 4 : 2;
  const int vec = (ldb % (16 / esz) == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0);
  hipLaunchKernelGGL((k_spmm_mfma3<PREC, TB, WM, WN>), dim3((unsigned)(8 * per_xcd)), dim3(256),
                     0, s, A->words, A->rows, A->cols, A->ldw, B, ldb, N, C, ldc, alpha, beta,
                     row_blocks, col_blocks, per_xcd, vec);
  return launch_check();
}

// Tiles built per (precision, B type); mfma_tile_for() only returns these.
//   16-bit B:        128 x 128, 128 x 256, 256 x 128
//   FP32 B (BF16 / FP16): 128 x 128 (twice the staging registers)
//   BF16X2:          128 x 128, 256 x 128 (two B planes in LDS)
template <int PREC, int TB>
constexpr bool has_tile(int bm, int bn) {
  if (bm == 128 && bn == 128) return true;
  if (bm == 256 && bn == 256) return PREC != 3 && TB != 0;  // direct-to-LDS only
  if (PREC == 3) return bm == 256 && bn == 128;
  if (TB == 0) return false;
  return (bm == 128 && bn == 256) || (bm == 256 && bn == 128);
}

template <int PREC, int WM, int WN, int NBUF, int KS>
bws_status launch_mfma4(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                        int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, 2 * WM);
  const int64_t col_blocks = ceil_div(N, 2 * WN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL((k_spmm_mfma4<PREC, WM, WN, NBUF, KS>), dim3((unsigned)(8 * per_xcd)), dim3(256),
                     0, s, A->words, A->rows, A->cols, A->ldw, static_cast<const uint16_t*>(B),
                     ldb, N, C, ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

template <int PREC, int TB>
bws_status launch_mfma_cfg(const MfmaTile& t, const bws_bitmap* A, const void* B, int64_t ldb,
                           int64_t N, float* C, int64_t ldc, float alpha, float beta,
                           hipStream_t s) {
  if (t.legacy) return launch_mfma<PREC, TB>(A, B, ldb, N, C, ldc, alpha, beta, s);
  if (!has_tile<PREC, TB>(t.bm, t.bn)) return BWS_EUNSUPPORTED;
  if constexpr (TB != 0 && PREC != 3) {
    if (t.bm == 256 && t.bn == 256)  // the direct-to-LDS 128 x 128-per-wave kernel
      return launch_mfma4_big(PREC == 2 ? BWS_PREC_FP16 : BWS_PREC_BF16, A, B, ldb, N, C, ldc,
                              alpha, beta, s);
    if (t.dma) {
#define BWS_L4(WM, WN, NB, KS) return launch_mfma4<PREC, WM, WN, NB, KS>(A, B, ldb, N, C, ldc, alpha, beta, s)
      if (t.nbuf == 3) {
        if (t.bn == 256) BWS_L4(64, 128, 3, 64);
        if (t.bm == 256) BWS_L4(128, 64, 3, 64);
        BWS_L4(64, 64, 3, 64);
      }
      if (t.ks == 128) {
        if (t.bn == 256) BWS_L4(64, 128, 2, 128);
        if (t.bm == 256) BWS_L4(128, 64, 2, 128);
        BWS_L4(64, 64, 2, 128);
      }
      if (t.bn == 256) BWS_L4(64, 128, 2, 64);
      if (t.bm == 256) BWS_L4(128, 64, 2, 64);
      BWS_L4(64, 64, 2, 64);
#undef BWS_L4
    }
  }
  if constexpr (has_tile<PREC, TB>(128, 256)) {
    if (t.bn == 256) return launch_mfma3<PREC, TB, 64, 128>(A, B, ldb, N, C, ldc, alpha, beta, s);
  }
  if constexpr (has_tile<PREC, TB>(256, 128)) {
    if (t.bm == 256) return launch_mfma3<PREC, TB, 128, 64>(A, B, ldb, N, C, ldc, alpha, beta, s);
  }
  return launch_mfma3<PREC, TB, 64, 64>(A, B, ldb, N, C, ldc, alpha, beta, s);
}

template <int PREC, int TB, int WM, int WN>
void launch_mfma3_batch(dim3 grid, const BatchEntry* entries, const int64_t* tile_off, int count,
                        int64_t per_xcd, int vec, hipStream_t s) {
  hipLaunchKernelGGL((k_spmm_mfma3_batch<PREC, TB, WM, WN>), grid, dim3(256), 0, s, entries,
                     tile_off, count, per_xcd, vec);
}

template <int PREC, int TB>
bws_status launch_mfma_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                             int64_t total_tiles, const MfmaTile& t, int vec, hipStream_t s) {
  const int64_t per_xcd = ceil_div(total_tiles, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  if (total_tiles == 0) return BWS_OK;
  if (!has_tile<PREC, TB>(t.bm, t.bn)) return BWS_EUNSUPPORTED;
  const dim3 grid((unsigned)(8 * per_xcd));
  bool done = false;
  if constexpr (TB != 0 && PREC != 3) {
    if (t.dma) {
      if (t.bn == 256)
        hipLaunchKernelGGL((k_spmm_mfma4_batch<PREC, 64, 128, 2, 64>), grid, dim3(256), 0, s,
                           entries, tile_off, count, per_xcd);
      else if (t.bm == 256)
        hipLaunchKernelGGL((k_spmm_mfma4_batch<PREC, 128, 64, 2, 64>), grid, dim3(256), 0, s,
                           entries, tile_off, count, per_xcd);
      else
        hipLaunchKernelGGL((k_spmm_mfma4_batch<PREC, 64, 64, 2, 64>), grid, dim3(256), 0, s,
                           entries, tile_off, count, per_xcd);
      return launch_check();
    }
  }
  if constexpr (has_tile<PREC, TB>(128, 256)) {
    if (t.bn == 256) {
      launch_mfma3_batch<PREC, TB, 64, 128>(grid, entries, tile_off, count, per_xcd, vec, s);
      done = true;
    }
  }
  if constexpr (has_tile<PREC, TB>(256, 128)) {
    if (!done && t.bm == 256) {
      launch_mfma3_batch<PREC, TB, 128, 64>(grid, entries, tile_off, count, per_xcd, vec, s);
      done = true;
    }
  }
  if (!done) launch_mfma3_batch<PREC, TB, 64, 64>(grid, entries, tile_off, count, per_xcd, vec, s);
  return launch_check();
}

}  // namespace

// Direct-to-LDS staging applies to 16-bit B with 16-byte aligned rows and N a
// multiple of 8 (every 16-byte piece of a row is either inside the row or
// clamped to its last 8 columns).
bool dma_ok(int32_t precision, int32_t b_dtype, const void* B, int64_t ldb, int64_t N) {
  return precision != BWS_PREC_BF16X2 && b_dtype != BWS_DTYPE_F32 && N % 8 == 0 &&
         ldb % 8 == 0 && (reinterpret_cast<uintptr_t>(B) % 16) == 0;
}

MfmaTile mfma_tile_for(const bws_plan* plan, int32_t precision, int32_t b_dtype, int64_t M,
                       int64_t N) {
  MfmaTile t;
  const int variant = plan && plan->kernel == 2 ? plan->mfma_staging : BWS_MFMA_AUTO;
  if (variant == BWS_MFMA_LUT) {
    t.legacy = 1;
    return t;
  }
  t.sparse = variant == BWS_MFMA_SPARSE;
  // direct-to-LDS unless register staging is asked for; the caller clears it
  // when B does not allow it
  t.dma = variant != BWS_MFMA_REGS;
  t.nbuf = variant == BWS_MFMA_DMA3 ? 3 : 2;
  t.ks = plan && plan->kernel == 2 && plan->k_stage == 128 ? 128 : 64;
  int bm = plan && plan->kernel == 2 ? plan->rows_per_wg : 0;
  int bn = plan && plan->kernel == 2 ? plan->cols_per_wg : 0;
  if (bm == 0 || bn == 0) {
    // the larger tile when it still gives about 1.5 workgroups per CU (two fit
    // on a CU), else 128 x 128
    auto tiles = [&](int64_t a, int64_t b) { return ceil_div(M, a) * ceil_div(N, b); };
    if (tiles(128, 256) >= 384) bm = 128, bn = 256;
    else bm = 128, bn = 128;
  }
  if (precision == BWS_PREC_BF16X2) {
    bn = 128;  // two B planes
    if (!(plan && plan->kernel == 2 && plan->rows_per_wg) &&
        ceil_div(M, 256) * ceil_div(N, 128) >= 384)
      bm = 256;
  } else if (b_dtype == BWS_DTYPE_F32) {
    bm = bn = 128;
  }
  t.bm = bm;
  t.bn = bn;
  return t;
}

bool mfma_tile_valid(int bm, int bn) {
  return (bm == 0 || bm == 128 || bm == 256) && (bn == 0 || bn == 128 || bn == 256);
}

bws_status spmm_mfma_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                           int64_t total_tiles, const MfmaTile& t, int32_t b_dtype,
                           int32_t precision, int vec, hipStream_t s) {
#define BWS_MB(P, T) return launch_mfma_batch<P, T>(entries, tile_off, count, total_tiles, t, vec, s)
  if (precision == BWS_PREC_BF16) {
    if (b_dtype == BWS_DTYPE_BF16) BWS_MB(1, 1);
    if (b_dtype == BWS_DTYPE_F32) BWS_MB(1, 0);
  } else if (precision == BWS_PREC_FP16) {
    if (b_dtype == BWS_DTYPE_F16) BWS_MB(2, 2);
    if (b_dtype == BWS_DTYPE_F32) BWS_MB(2, 0);
  } else if (precision == BWS_PREC_BF16X2) {
    if (b_dtype == BWS_DTYPE_F32) BWS_MB(3, 0);
  }
#undef BWS_MB
  return BWS_EUNSUPPORTED;
}

namespace {

constexpr int kSplitSlices = 16;  // partial counts per block (8 rows each)

// Set bits of rows [128 b + 8 s, +8) of the split's row block b, slice s.
__global__ void __launch_bounds__(256)
k_split_counts(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
               int32_t* __restrict__ partial) {
  __shared__ int32_t part[4];
  const int64_t r0 = (int64_t)blockIdx.x * kSplitRows + (int64_t)blockIdx.y * (kSplitRows / kSplitSlices);
  int64_t rows = M - r0;
  rows = rows < 0 ? 0 : (rows < kSplitRows / kSplitSlices ? rows : kSplitRows / kSplitSlices);
  const int64_t wpr = words_for(K), n = rows * wpr;
  int32_t c = 0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    const int64_t r = i / wpr, w = i - r * wpr;
    c += __popc(W[(r0 + r) * ldw + w] & valid_bits(w, K));
  }
  c = wave_sum(c);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0)
    partial[blockIdx.x * kSplitSlices + blockIdx.y] = part[0] + part[1] + part[2] + part[3];
}

// One workgroup: flags[b] = the block has at least `threshold` bits, and the
// flagged blocks in order, list[0] = their count, list[1 ..] = the blocks.
__global__ void __launch_bounds__(256)
k_split_list(const int32_t* __restrict__ partial, int64_t blocks, int64_t threshold,
             uint8_t* __restrict__ flags, int32_t* __restrict__ list) {
  __shared__ int32_t wsum[4];
  __shared__ int32_t base;
  if (threadIdx.x == 0) base = 0;
  __syncthreads();
  for (int64_t b0 = 0; b0 < blocks; b0 += 256) {
    const int64_t b = b0 + threadIdx.x;
    int64_t tot = 0;
    if (b < blocks)
      for (int s = 0; s < kSplitSlices; ++s) tot += partial[b * kSplitSlices + s];
    const int f = (b < blocks && tot >= threshold) ? 1 : 0;
    if (b < blocks) flags[b] = (uint8_t)f;
    const int incl = wave_incl_scan(f);
    if ((threadIdx.x & 63) == 63) wsum[threadIdx.x >> 6] = incl;
    __syncthreads();
    int off = base;
    for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) off += wsum[w];
    if (f) list[1 + off + incl - 1] = (int32_t)b;
    __syncthreads();
    if (threadIdx.x == 0) base += wsum[0] + wsum[1] + wsum[2] + wsum[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) list[0] = base;
}

// The MFMA side of the split: the flagged row blocks' tiles, numbered column
// block by column block (tile t: column block t / count, row block list[1 +
// t % count]), handed out in contiguous ranges per XCD as the full kernel does,
// so that one XCD's L2 holds few column blocks of B.  Workgroups beyond each
// XCD's share (the count is only known on the device) exit at once.
template <int PREC, int WN>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma4_list(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                  const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
                  int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
                  const int32_t* __restrict__ list) {
  const int64_t nm = list[0];
  const int64_t total = nm * col_blocks, per_xcd = (total + 7) >> 3;
  if ((int64_t)(blockIdx.x >> 3) >= per_xcd) return;
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= total) return;
  const int64_t rb = list[1 + g % nm], cb = g / nm;
  mfma4_tile<PREC, 64, WN, 2, 64>(W, M, K, ldw, B, ldb, N, C, ldc, alpha, beta,
                                  cb * row_blocks + rb, row_blocks);
}

}  // namespace

size_t split_ws_bytes(int64_t M) {
  const int64_t blocks = ceil_div(M, kSplitRows);
  return (size_t)(ceil_div(blocks * kSplitSlices * 4, 256) + ceil_div(blocks, 256) +
                  ceil_div((blocks + 1) * 4, 256)) * 256;
}

bws_status launch_split_flags(const bws_bitmap* A, int32_t min_bits, void* ws, uint8_t** flags,
                              int32_t** list, hipStream_t s) {
  const int64_t blocks = ceil_div(A->rows, kSplitRows);
  if (blocks > 65535) return BWS_ECAPACITY;
  Carve cv(ws);
  int32_t* partial = cv.take<int32_t>((size_t)(blocks * kSplitSlices));
  *flags = cv.take<uint8_t>((size_t)blocks);
  *list = cv.take<int32_t>((size_t)(blocks + 1));
  hipLaunchKernelGGL(k_split_counts, dim3((unsigned)blocks, kSplitSlices), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, partial);
  BWS_TRY(launch_check());
  hipLaunchKernelGGL(k_split_list, dim3(1), dim3(256), 0, s, partial, blocks,
                     (int64_t)min_bits * ceil_div(A->cols, 64), *flags, *list);
  return launch_check();
}

// The MFMA side of the density split: the direct-to-LDS kernel with 128 x 128
// workgroup tiles (64 x 64 per wave) on the flagged row blocks.  The tile is
// narrower than the full kernel's 128 x 256 at large N because one tile's
// K loop is latency-bound: with half the row blocks flagged, 128 x 256 tiles
// would leave half of each CU's two workgroup slots empty.
bws_status spmm_mfma_split16(const bws_bitmap* A, const void* B, int32_t precision, int64_t ldb,
                             int64_t N, float* C, int64_t ldc, float alpha, float beta,
                             const int32_t* list, hipStream_t s) {
  static_assert(2 * 64 == kSplitRows, "MFMA row blocks are the split blocks");
  const int64_t row_blocks = ceil_div(A->rows, kSplitRows), col_blocks = ceil_div(N, 128);
  const int64_t grid = 8 * ceil_div(row_blocks * col_blocks, 8);
  if (grid > kMaxGrid) return BWS_ECAPACITY;
  const uint16_t* b = static_cast<const uint16_t*>(B);
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL((k_spmm_mfma4_list<2, 64>), dim3((unsigned)grid), dim3(256), 0, s, A->words,
                       A->rows, A->cols, A->ldw, b, ldb, N, C, ldc, alpha, beta, row_blocks,
                       col_blocks, list);
  else
    hipLaunchKernelGGL((k_spmm_mfma4_list<1, 64>), dim3((unsigned)grid), dim3(256), 0, s, A->words,
                       A->rows, A->cols, A->ldw, b, ldb, N, C, ldc, alpha, beta, row_blocks,
                       col_blocks, list);
  return launch_check();
}

bws_status spmm_mfma(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                     float* C, int64_t ldc, float alpha, float beta, int32_t precision,
                     const bws_plan* plan, hipStream_t s) {
  MfmaTile t = mfma_tile_for(plan, precision, b_dtype, A->rows, N);
  if (!dma_ok(precision, b_dtype, B, ldb, N)) t.dma = 0;
  if (t.sparse && t.dma && !A->values &&
      ((precision == BWS_PREC_BF16 && b_dtype == BWS_DTYPE_BF16) ||
       (precision == BWS_PREC_FP16 && b_dtype == BWS_DTYPE_F16)))
    return launch_smfmac(precision, A, B, ldb, N, C, ldc, alpha, beta, s);
  if (!t.dma && t.bm == 256 && t.bn == 256) t.bm = 128;  // 256 x 256 needs the direct copy
#define BWS_M(P, T) return launch_mfma_cfg<P, T>(t, A, B, ldb, N, C, ldc, alpha, beta, s)
  if (precision == BWS_PREC_BF16) {
    if (b_dtype == BWS_DTYPE_BF16) BWS_M(1, 1);
    if (b_dtype == BWS_DTYPE_F32) BWS_M(1, 0);
  } else if (precision == BWS_PREC_FP16) {
    if (b_dtype == BWS_DTYPE_F16) BWS_M(2, 2);
    if (b_dtype == BWS_DTYPE_F32) BWS_M(2, 0);
  } else if (precision == BWS_PREC_BF16X2) {
    if (b_dtype == BWS_DTYPE_F32) BWS_M(3, 0);
  }
#undef BWS_M
  return BWS_EUNSUPPORTED;
}

namespace {

// hi = bf16(b), lo = bf16(b - hi) (the split of pack_block), four columns per
// thread (N % 8 == 0); VEC: B rows 16-byte aligned, one float4 load.
template <bool VEC>
__global__ void __launch_bounds__(256)
k_split_bf16x2(const float* __restrict__ B, int64_t K, int64_t N, int64_t ldb,
               uint2* __restrict__ hi, uint2* __restrict__ lo) {
  const int64_t quads = N / 4, total = K * quads;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int64_t k = i / quads, c = 4 * (i - k * quads);
    const float* src = B + k * ldb + c;
    float e[4];
    if (VEC) {
      const float4 v = *reinterpret_cast<const float4*>(src);
      e[0] = v.x, e[1] = v.y, e[2] = v.z, e[3] = v.w;
    } else {
#pragma unroll
      for (int x = 0; x < 4; ++x) e[x] = src[x];
    }
    uint32_t h[4], l[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      h[x] = f32_to_bf16_rne(e[x]);
      l[x] = f32_to_bf16_rne(e[x] - bf16_to_f32(h[x]));
    }
    hi[i] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
    lo[i] = make_uint2(l[0] | (l[1] << 16), l[2] | (l[3] << 16));
  }
}

size_t split_plane_bytes(int64_t K, int64_t N) { return (size_t)ceil_div(K * N * 2, 256) * 256; }

// FP32 B rounded once to the precision's 16-bit type (to nearest even, as the
// register-staged kernels round it), four columns per thread (N % 8 == 0).
template <int PREC, bool VEC>
__global__ void __launch_bounds__(256)
k_round_b16(const float* __restrict__ B, int64_t K, int64_t N, int64_t ldb,
            uint2* __restrict__ out) {
  const int64_t quads = N / 4, total = K * quads;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < total;
       i += (int64_t)gridDim.x * 256) {
    const int64_t k = i / quads, c = 4 * (i - k * quads);
    const float* src = B + k * ldb + c;
    float e[4];
    if (VEC) {
      const float4 v = *reinterpret_cast<const float4*>(src);
      e[0] = v.x, e[1] = v.y, e[2] = v.z, e[3] = v.w;
    } else {
#pragma unroll
      for (int x = 0; x < 4; ++x) e[x] = src[x];
    }
    uint32_t h[4];
#pragma unroll
    for (int x = 0; x < 4; ++x) h[x] = PREC == 2 ? f32_to_f16_rne(e[x]) : f32_to_bf16_rne(e[x]);
    out[i] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
  }
}

}  // namespace

size_t spmm_split_workspace(int64_t K, int64_t N) { return 2 * split_plane_bytes(K, N); }
size_t spmm_round_workspace(int64_t K, int64_t N) { return split_plane_bytes(K, N); }

// BF16 / FP16 with FP32 B on a 0/1 A: B is rounded once into the workspace
// (row-major, ld N) and the direct-to-LDS 16-bit kernel multiplies it; the
// rounding and the K order are those of the register-staged kernel, so the
// result is the same bit for bit, without staging B through registers.
// FP32 B rounded once into out (K x N, ld N) in the precision's 16-bit type.
bws_status launch_round_b16(const float* B, int64_t K, int64_t N, int64_t ldb, int32_t precision,
                            uint16_t* Bh, hipStream_t s) {
  const int64_t quads = K * (N / 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(quads, 256), 16384);
  const bool vec = ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0;
  uint2* out = reinterpret_cast<uint2*>(Bh);
  if (precision == BWS_PREC_FP16) {
    if (vec) hipLaunchKernelGGL((k_round_b16<2, true>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
    else hipLaunchKernelGGL((k_round_b16<2, false>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
  } else {
    if (vec) hipLaunchKernelGGL((k_round_b16<1, true>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
    else hipLaunchKernelGGL((k_round_b16<1, false>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
  }
  return launch_check();
}

bws_status spmm_mfma_round(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N, float* C,
                           int64_t ldc, float alpha, float beta, int32_t precision,
                           const bws_plan* plan, void* workspace, hipStream_t s) {
  const int64_t K = A->cols;
  uint16_t* Bh = static_cast<uint16_t*>(workspace);
  const int64_t quads = K * (N / 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(quads, 256), 16384);
  const bool vec = ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0;
  uint2* out = reinterpret_cast<uint2*>(Bh);
  if (precision == BWS_PREC_FP16) {
    if (vec) hipLaunchKernelGGL((k_round_b16<2, true>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
    else hipLaunchKernelGGL((k_round_b16<2, false>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
  } else {
    if (vec) hipLaunchKernelGGL((k_round_b16<1, true>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
    else hipLaunchKernelGGL((k_round_b16<1, false>), dim3(grid), dim3(256), 0, s, B, K, N, ldb, out);
  }
  BWS_TRY(launch_check());
  return spmm_mfma(A, Bh, precision == BWS_PREC_FP16 ? BWS_DTYPE_F16 : BWS_DTYPE_BF16, N, N, C,
                   ldc, alpha, beta, precision, plan, s);
}

// Split-BF16 through the 16-bit kernels: B's hi and lo planes are written to
// the workspace (row-major, ld N) and multiplied by the BF16 MFMA kernel in
// two passes, C = alpha·A·hi + beta·C, then C = alpha·A·lo + C.  Each pass
// stages B straight into LDS, which the register-staged split cannot.
bws_status spmm_mfma_split(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N, float* C,
                           int64_t ldc, float alpha, float beta, const bws_plan* plan,
                           void* workspace, hipStream_t s) {
  const int64_t K = A->cols;
  uint16_t* hi = static_cast<uint16_t*>(workspace);
  uint16_t* lo = reinterpret_cast<uint16_t*>(static_cast<char*>(workspace) + split_plane_bytes(K, N));
  const int64_t quads = K * (N / 4);
  const unsigned grid = (unsigned)std::min<int64_t>(ceil_div(quads, 256), 16384);
  if (ldb % 4 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0)
    hipLaunchKernelGGL(k_split_bf16x2<true>, dim3(grid), dim3(256), 0, s, B, K, N, ldb,
                       reinterpret_cast<uint2*>(hi), reinterpret_cast<uint2*>(lo));
  else
    hipLaunchKernelGGL(k_split_bf16x2<false>, dim3(grid), dim3(256), 0, s, B, K, N, ldb,
                       reinterpret_cast<uint2*>(hi), reinterpret_cast<uint2*>(lo));
  BWS_TRY(launch_check());
  BWS_TRY(spmm_mfma(A, hi, BWS_DTYPE_BF16, N, N, C, ldc, alpha, beta, BWS_PREC_BF16, plan, s));
  return spmm_mfma(A, lo, BWS_DTYPE_BF16, N, N, C, ldc, alpha, 1.0f, BWS_PREC_BF16, plan, s);
}

}  // namespace bws

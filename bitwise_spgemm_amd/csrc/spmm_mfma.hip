// spmm_mfma.hip — 16-bit MFMA tile GEMM for a 0/1 bitmap A (gfx950,
// v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulate).
//
// Workgroup = 4 waves (2 x 2), C tile 128 x 128, each wave 64 x 64 (4 x 4
// MFMA tiles).  K advances 128 rows per LDS stage = four 32-wide MFMA steps =
// four bitmap words, so one barrier covers 64 MFMAs per wave.
//  - A operand: the byte of the bitmap word that a lane's 8 K-slots cover
//    indexes a 256-entry LDS table of eight 16-bit 0/1 values (one
//    ds_read_b128); no arithmetic builds the fragment.  A stage whose 64 x 128
//    A block (one wave's rows) has no set bit skips its MFMAs; otherwise the
//    MFMAs run branch-free with the fragments of K-step s+1 read while step s
//    multiplies.
//  - B operand: staged in LDS transposed (N-major, K contiguous, 272-byte rows)
//    so a lane's 8 K-consecutive values are one 16-byte read.  Each thread
//    loads an 8(K) x 8(N) block with 16-byte global loads, transposes it in
//    registers with v_perm_b32 and writes eight 16-byte LDS rows.  The next
//    stage's global loads are in flight while the current stage computes.
//  - A words (4 per row per stage) go through LDS as well.
// Operand layout used for both operands: lane l covers row/column (l % 16)
// and the K slots 8*(l/16) .. 8*(l/16)+7 of the 32-wide step; since A and B use
// the same K assignment the product does not depend on the hardware's K order.
// The accumulator layout is D[4*(l/16) + e][l % 16], e = 0..3.
//
// BF16X2: B = hi + lo with hi = bf16(b), lo = bf16(b - hi); both are staged
// and both products accumulate into the same FP32 accumulators.
#include "bws_internal.hpp"

namespace bws {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int BM = 128, BN = 128, KS = 128;
constexpr int LDK = KS + 8;  // 16-bit elements per LDS row: 272 bytes, 16-byte aligned

__device__ __forceinline__ uint32_t f32_to_bf16_rne(float x) {
  const __bf16 h = (__bf16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ uint32_t f32_to_f16_rne(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ float bf16_to_f32(uint32_t u) { return __uint_as_float(u << 16); }

// PREC: 1 = BF16, 2 = FP16, 3 = BF16X2.  TB: B element type 0 = f32, 1 = bf16, 2 = f16.
template <int PREC>
__device__ __forceinline__ f32x4 mfma(const uint4& a, const uint4& b, const f32x4& c) {
  if constexpr (PREC == 2) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(f16x8, a),
                                                  __builtin_bit_cast(f16x8, b), c, 0, 0, 0);
  } else {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                   __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
  }
}

__device__ __forceinline__ int64_t xcd_tile(int64_t g, int64_t per_xcd) {
  return (g & 7) * per_xcd + (g >> 3);
}

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

template <int PREC, int TB>
bws_status launch_mfma_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                             int64_t total_tiles, int vec, hipStream_t s) {
  const int64_t per_xcd = ceil_div(total_tiles, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  if (total_tiles == 0) return BWS_OK;
  hipLaunchKernelGGL((k_spmm_mfma_batch<PREC, TB>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0,
                     s, entries, tile_off, count, per_xcd, vec);
  return launch_check();
}

}  // namespace

bws_status spmm_mfma_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                           int64_t total_tiles, int32_t b_dtype, int32_t precision, int vec,
                           hipStream_t s) {
#define BWS_MB(P, T) return launch_mfma_batch<P, T>(entries, tile_off, count, total_tiles, vec, s)
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

bws_status spmm_mfma(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                     float* C, int64_t ldc, float alpha, float beta, int32_t precision,
                     hipStream_t s) {
  if (precision == BWS_PREC_BF16) {
    if (b_dtype == BWS_DTYPE_BF16) return launch_mfma<1, 1>(A, B, ldb, N, C, ldc, alpha, beta, s);
    if (b_dtype == BWS_DTYPE_F32) return launch_mfma<1, 0>(A, B, ldb, N, C, ldc, alpha, beta, s);
  } else if (precision == BWS_PREC_FP16) {
    if (b_dtype == BWS_DTYPE_F16) return launch_mfma<2, 2>(A, B, ldb, N, C, ldc, alpha, beta, s);
    if (b_dtype == BWS_DTYPE_F32) return launch_mfma<2, 0>(A, B, ldb, N, C, ldc, alpha, beta, s);
  } else if (precision == BWS_PREC_BF16X2) {
    if (b_dtype == BWS_DTYPE_F32) return launch_mfma<3, 0>(A, B, ldb, N, C, ldc, alpha, beta, s);
  }
  return BWS_EUNSUPPORTED;
}

}  // namespace bws

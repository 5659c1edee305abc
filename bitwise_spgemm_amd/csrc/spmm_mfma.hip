// spmm_mfma.hip — 16-bit MFMA tile GEMM for a 0/1 bitmap A (gfx950,
// v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulate).
//
// Workgroup = 4 waves, C tile 128 x 128, each wave 64 x 64 (4 x 4 MFMA tiles).
// K advances one 32-column bitmap word per step, which is exactly the K extent
// of one MFMA.  The A operand is built from the bitmap with no arithmetic: the
// byte of the word that a lane's 8 K-slots cover indexes a 256-entry LDS table
// of eight 16-bit 0/1 values (one ds_read_b128).  B is staged in LDS
// transposed (N-major, K contiguous) so that a lane's 8 K-consecutive values
// are one 16-byte read.  A 16 x 32 A tile whose bits are all zero skips its
// MFMAs (wave-uniform ballot test), so empty tiles cost nothing.
//
// Operand layout used for both operands: lane l covers row/column (l % 16)
// and the K slots 8*(l/16) .. 8*(l/16)+7; since A and B use the same K
// assignment the product is independent of the hardware's K order.  The
// accumulator layout is D[4*(l/16) + e][l % 16], e = 0..3.
//
// BF16X2: B = hi + lo with hi = bf16(b), lo = bf16(b - hi); both are staged
// and both products accumulate into the same FP32 accumulators.
#include "bws_internal.hpp"

namespace bws {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(8))) _Float16 f16x8;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr int BM = 128, BN = 128, KC = 32;
constexpr int LDK = KC + 8;  // 16-bit elements per LDS row (80 bytes, 16-byte aligned)

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float x) {
  const __bf16 h = (__bf16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ uint16_t f32_to_f16_rne(float x) {
  const _Float16 h = (_Float16)x;
  uint16_t u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}
__device__ __forceinline__ float bf16_to_f32(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

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

// Convert one B element to the 16-bit operand(s) of precision PREC.
template <int PREC, int TB>
__device__ __forceinline__ void to16(const void* __restrict__ Bv, int64_t off, bool ok,
                                     uint16_t& hi, uint16_t& lo) {
  lo = 0;
  if (!ok) {
    hi = 0;
    return;
  }
  if constexpr (TB == 0) {
    const float x = static_cast<const float*>(Bv)[off];
    if constexpr (PREC == 2) {
      hi = f32_to_f16_rne(x);
    } else {
      hi = f32_to_bf16_rne(x);
      if constexpr (PREC == 3) lo = f32_to_bf16_rne(x - bf16_to_f32(hi));
    }
  } else {
    hi = static_cast<const uint16_t*>(Bv)[off];
  }
}

__device__ __forceinline__ int64_t xcd_tile(int64_t g, int64_t per_xcd) {
  return (g & 7) * per_xcd + (g >> 3);
}

template <int PREC, int TB>
__global__ void __launch_bounds__(256)
k_spmm_mfma(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
            const void* __restrict__ Bv, int64_t ldb, int64_t N, float* __restrict__ C,
            int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
            int64_t per_xcd) {
  constexpr int NS = PREC == 3 ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t sBt[2][NS][BN][LDK];
  __shared__ uint4 sLut[256];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
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

  const int64_t nk = (K + KC - 1) / KC;
  const int krem = (int)(K % KC);
  const uint32_t last_mask = krem ? (0xFFFFFFFFu << (32 - krem)) : 0xFFFFFFFFu;

  // staging assignment: 4 x 4 block of B per thread
  const int kb = tid >> 5, nb = tid & 31;
  uint16_t sh[4][4], sl[4][4];  // [k][n]
  auto load_b = [&](int64_t kc) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t k = kc * KC + 4 * kb + r;
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const int64_t n = n0 + 4 * nb + x;
        to16<PREC, TB>(Bv, k * ldb + n, k < K && n < N, sh[r][x], sl[r][x]);
      }
    }
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const uint2 vh = make_uint2((uint32_t)sh[0][x] | ((uint32_t)sh[1][x] << 16),
                                  (uint32_t)sh[2][x] | ((uint32_t)sh[3][x] << 16));
      *reinterpret_cast<uint2*>(&sBt[buf][0][4 * nb + x][4 * kb]) = vh;
      if constexpr (NS == 2) {
        const uint2 vl = make_uint2((uint32_t)sl[0][x] | ((uint32_t)sl[1][x] << 16),
                                    (uint32_t)sl[2][x] | ((uint32_t)sl[3][x] << 16));
        *reinterpret_cast<uint2*>(&sBt[buf][NS - 1][4 * nb + x][4 * kb]) = vl;
      }
    }
  };

  // A rows of this lane: m0 + 64*wm + 16*a + j, a = 0..3
  int64_t arow[4];
  bool aok[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    arow[a] = m0 + 64 * wm + 16 * a + j;
    aok[a] = arow[a] < M;
  }
  uint32_t wcur[4], wnext[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) wcur[a] = aok[a] ? W[arow[a] * ldw] : 0u;

  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  load_b(0);
  store_b(0);
  __syncthreads();

  for (int64_t kc = 0; kc < nk; ++kc) {
    const int buf = (int)(kc & 1);
    const bool more = kc + 1 < nk;
    if (more) {
      load_b(kc + 1);
#pragma unroll
      for (int a = 0; a < 4; ++a) wnext[a] = aok[a] ? W[arow[a] * ldw + kc + 1] : 0u;
    }
    uint4 bh[4], bl[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      bh[b] = *reinterpret_cast<const uint4*>(&sBt[buf][0][64 * wn + 16 * b + j][8 * q]);
      if constexpr (NS == 2)
        bl[b] = *reinterpret_cast<const uint4*>(&sBt[buf][NS - 1][64 * wn + 16 * b + j][8 * q]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const uint32_t w = more ? wcur[a] : (wcur[a] & last_mask);
      const uint32_t byte = (w >> (24 - 8 * q)) & 0xFFu;
      if (__ballot(byte != 0u) == 0ull) continue;  // empty 16 x 32 tile
      const uint4 af = sLut[byte];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        acc[a][b] = mfma<PREC>(af, bh[b], acc[a][b]);
        if constexpr (NS == 2) acc[a][b] = mfma<PREC>(af, bl[b], acc[a][b]);
      }
    }
    if (more) {
      store_b(buf ^ 1);
#pragma unroll
      for (int a = 0; a < 4; ++a) wcur[a] = wnext[a];
    }
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
bws_status launch_mfma(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                       int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, BM);
  const int64_t col_blocks = ceil_div(N, BN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL((k_spmm_mfma<PREC, TB>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, B, ldb, N, C, ldc, alpha, beta,
                     row_blocks, col_blocks, per_xcd);
  return launch_check();
}

}  // namespace

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

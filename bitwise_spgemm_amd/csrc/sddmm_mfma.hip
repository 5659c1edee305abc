// sddmm_mfma.hip — sampled dense-dense product on the bitmap pattern with
// 16-bit MFMA (gfx950, v_mfma_f32_16x16x32_{bf16,f16}, FP32 accumulate):
//   dA[row_ptr[r] + rank(r, k)] = Σ_n dC[r, n] · B[k, n]   for every set bit (r, k).
//
// Workgroup = 4 waves, output block 128 rows x 128 K-columns of A, each wave
// 64 x 64 (4 x 4 MFMA tiles); the reduction runs over N in LDS stages of 64.
// Both operands are N-contiguous in their row-major layouts (dC rows, B rows),
// which is exactly the K-contiguous fragment each MFMA lane needs, so staging
// is a plain copy (no transpose).  A block whose 128 x 128 bitmap is empty is
// skipped as a whole.  The epilogue stores only set bits, at the value index
// row_ptr[r] + word offset + popcount of the earlier bits of the word.
// BF16X2 splits both operands (hi + lo) and accumulates hi·hi + hi·lo + lo·hi.
// Large problems with direct-to-LDS operands take the 256 x 256-block kernel
// of sddmm_mfma_big.hip.
#include "mfma_common.hpp"

namespace bws {
namespace {

constexpr int BR = 128, BK = 128, NSTG = 64;
constexpr int LDN = NSTG + 8;  // 144-byte LDS rows: 16-byte aligned, conflict-free reads

__device__ __forceinline__ uint32_t to_bf16(float x) { return f32_to_bf16_rne(x); }
__device__ __forceinline__ uint32_t to_f16(float x) { return f32_to_f16_rne(x); }

// Epilogue of a 128 x 128 block: its bitmap words and the value index of each
// word's first bit are staged in LDS (`scratch`: 6 KiB the block no longer
// needs), then every lane stores the set bits among its accumulators (rows
// 4q + e of each 16-row tile, column j of each 16-column tile).
__device__ __forceinline__ void store_set_bits_128(const f32x4 (&acc)[4][4], uint8_t* scratch,
                                                   const uint32_t* __restrict__ W, int64_t M,
                                                   int64_t K, int64_t ldw,
                                                   const int64_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ woff,
                                                   float* __restrict__ dA, int64_t r0,
                                                   int64_t k0, int wm, int wn, int j, int q) {
  const int64_t wpr = (K + 31) / 32;
  uint32_t* eW = reinterpret_cast<uint32_t*>(scratch);          // [128][4] words
  int64_t* eBase = reinterpret_cast<int64_t*>(scratch + 2048);  // [128][4] value indices
  __syncthreads();  // every wave is done with the operand buffers
  for (int t = threadIdx.x; t < BR * (BK / 32); t += 256) {
    const int64_t r = r0 + (t >> 2), w = k0 / 32 + (t & 3);
    uint32_t word = 0u;
    int64_t base = 0;
    if (r < M && w < wpr) {
      word = W[r * ldw + w] & valid_bits(w, K);
      base = row_ptr[r] + woff[r * wpr + w];
    }
    eW[t] = word;
    eBase[t] = base;
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int rr = 64 * wm + 16 * a + 4 * q + e;
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int ew = rr * (BK / 32) + 2 * wn + (b >> 1);
        const uint32_t word = eW[ew];
        const int i = 16 * (b & 1) + j;
        if (word & wbit(i)) dA[eBase[ew] + rank_before(word, i)] = acc[a][b][e];
      }
    }
}

// Load 8 consecutive elements (row, n..n+7) as raw input, zero outside.
template <int TB>
struct Raw8;
template <>
struct Raw8<0> {
  float4 a, b;
};
template <>
struct Raw8<1> {
  uint4 v;
};
template <>
struct Raw8<2> {
  uint4 v;
};

template <int TB>
__device__ __forceinline__ void load8(Raw8<TB>& r, const void* base, int64_t row, int64_t rows,
                                      int64_t ld, int64_t n, int64_t N, bool vec) {
  if constexpr (TB == 0) {
    const float* p = static_cast<const float*>(base) + row * ld + n;
    if (row < rows && vec && n + 8 <= N) {
      r.a = *reinterpret_cast<const float4*>(p);
      r.b = *reinterpret_cast<const float4*>(p + 4);
    } else {
      float e[8];
#pragma unroll
      for (int x = 0; x < 8; ++x) e[x] = (row < rows && n + x < N) ? p[x] : 0.0f;
      r.a = make_float4(e[0], e[1], e[2], e[3]);
      r.b = make_float4(e[4], e[5], e[6], e[7]);
    }
  } else {
    const uint16_t* p = static_cast<const uint16_t*>(base) + row * ld + n;
    if (row < rows && vec && n + 8 <= N) {
      r.v = *reinterpret_cast<const uint4*>(p);
    } else {
      uint32_t d[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        const uint32_t lo = (row < rows && n + 2 * x < N) ? p[2 * x] : 0u;
        const uint32_t hi = (row < rows && n + 2 * x + 1 < N) ? p[2 * x + 1] : 0u;
        d[x] = lo | (hi << 16);
      }
      r.v = make_uint4(d[0], d[1], d[2], d[3]);
    }
  }
}

template <int PREC, int TB>
__device__ __forceinline__ void pack8(const Raw8<TB>& r, uint4& hi, uint4& lo) {
  if constexpr (TB == 0) {
    const float e[8] = {r.a.x, r.a.y, r.a.z, r.a.w, r.b.x, r.b.y, r.b.z, r.b.w};
    uint32_t h[8], l[8];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
      if constexpr (PREC == 2) {
        h[x] = to_f16(e[x]);
        l[x] = 0;
      } else {
        h[x] = to_bf16(e[x]);
        l[x] = PREC == 3 ? to_bf16(e[x] - __uint_as_float(h[x] << 16)) : 0u;
      }
    }
    hi = make_uint4(h[0] | (h[1] << 16), h[2] | (h[3] << 16), h[4] | (h[5] << 16), h[6] | (h[7] << 16));
    lo = make_uint4(l[0] | (l[1] << 16), l[2] | (l[3] << 16), l[4] | (l[5] << 16), l[6] | (l[7] << 16));
  } else {
    hi = r.v;
    lo = make_uint4(0, 0, 0, 0);
  }
}

template <int PREC, int TB>
__global__ void __launch_bounds__(256, 2)
k_sddmm_mfma(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
             const void* __restrict__ dC, int64_t lddc, const void* __restrict__ Bv,
             int64_t ldb, int64_t N, float* __restrict__ dA, int64_t row_blocks,
             int64_t k_blocks, int64_t per_xcd, int vec) {
  constexpr int NS = PREC == 3 ? 2 : 1;
  __shared__ __attribute__((aligned(16))) uint16_t sC[2][NS][BR][LDN];
  __shared__ __attribute__((aligned(16))) uint16_t sB[2][NS][BK][LDN];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * k_blocks) return;
  const int64_t kbk = tile / row_blocks, rb = tile - kbk * row_blocks;
  const int64_t r0 = rb * BR, k0 = kbk * BK;
  const int64_t wpr = (K + 31) / 32;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;

  // skip an empty 128 x 128 bitmap block (all waves agree through the barrier)
  {
    uint32_t any = 0;
    const int row = tid >> 1;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int64_t w = k0 / 32 + 2 * (tid & 1) + x;
      if (r0 + row < M && w < wpr) any |= W[(r0 + row) * ldw + w] & valid_bits(w, K);
    }
    if (!__syncthreads_or(any != 0u)) return;
  }

  // staging: each thread copies 4 eight-element segments of dC and of B per stage
  // (128 rows x 64 columns = 1024 segments of each, 4 per thread)
  Raw8<TB> rc[4], rbv[4];
  auto load_stage = [&](int64_t st) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int seg = tid + 256 * x;
      const int row = seg >> 3, col = 8 * (seg & 7);
      load8<TB>(rc[x], dC, r0 + row, M, lddc, st * NSTG + col, N, vec != 0);
      load8<TB>(rbv[x], Bv, k0 + row, K, ldb, st * NSTG + col, N, vec != 0);
    }
  };
  auto store_stage = [&](int buf) {
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const int seg = tid + 256 * x;
      const int row = seg >> 3, col = 8 * (seg & 7);
      uint4 h, l;
      pack8<PREC, TB>(rc[x], h, l);
      *reinterpret_cast<uint4*>(&sC[buf][0][row][col]) = h;
      if constexpr (NS == 2) *reinterpret_cast<uint4*>(&sC[buf][NS - 1][row][col]) = l;
      pack8<PREC, TB>(rbv[x], h, l);
      *reinterpret_cast<uint4*>(&sB[buf][0][row][col]) = h;
      if constexpr (NS == 2) *reinterpret_cast<uint4*>(&sB[buf][NS - 1][row][col]) = l;
    }
  };

  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int64_t nst = (N + NSTG - 1) / NSTG;
  load_stage(0);
  store_stage(0);
  __syncthreads();
  for (int64_t st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    const bool more = st + 1 < nst;
    if (more) load_stage(st + 1);
#pragma unroll
    for (int s = 0; s < NSTG / 32; ++s) {
      uint4 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        ah[a] = *reinterpret_cast<const uint4*>(&sC[buf][0][64 * wm + 16 * a + j][32 * s + 8 * q]);
        if constexpr (NS == 2)
          al[a] = *reinterpret_cast<const uint4*>(&sC[buf][NS - 1][64 * wm + 16 * a + j][32 * s + 8 * q]);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        bh[b] = *reinterpret_cast<const uint4*>(&sB[buf][0][64 * wn + 16 * b + j][32 * s + 8 * q]);
        if constexpr (NS == 2)
          bl[b] = *reinterpret_cast<const uint4*>(&sB[buf][NS - 1][64 * wn + 16 * b + j][32 * s + 8 * q]);
      }
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          acc[a][b] = mfma<PREC>(ah[a], bh[b], acc[a][b]);
          if constexpr (NS == 2) {
            acc[a][b] = mfma<PREC>(ah[a], bl[b], acc[a][b]);
            acc[a][b] = mfma<PREC>(al[a], bh[b], acc[a][b]);
          }
        }
    }
    if (more) store_stage(buf ^ 1);
    __syncthreads();
  }

  store_set_bits_128(acc, reinterpret_cast<uint8_t*>(&sC[0][0][0][0]), W, M, K, ldw, row_ptr,
                     woff, dA, r0, k0, wm, wn, j, q);
}

// Direct-to-LDS variant for 16-bit dC and B (rows 16-byte aligned, leading
// dimensions multiples of 8, N a multiple of the 64-column stage): both
// operands are copied row-major into LDS by global_load_lds_dwordx4, 16-byte
// chunk c of row r at chunk c ^ ((r >> 1) & 7), so the 16 lanes reading one
// chunk of 16 rows hit 16 different bank groups; each fragment is one
// 16-byte read.  Two stages, 64 KiB of LDS, two workgroups per CU.
constexpr int ROWB_D = NSTG * 2;          // 128 bytes per staged row
constexpr int STAGE_D = BR * ROWB_D;      // 16 KiB per operand per stage
constexpr int NC_W = STAGE_D / 1024 / 4;  // copies per wave per operand per stage

template <int PREC>
__global__ void __launch_bounds__(256, 2)
k_sddmm_mfma_dma(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
                 const uint16_t* __restrict__ dC, int64_t lddc, const uint16_t* __restrict__ B,
                 int64_t ldb, int64_t N, float* __restrict__ dA, int64_t row_blocks,
                 int64_t k_blocks, int64_t per_xcd) {
  __shared__ __attribute__((aligned(1024))) uint8_t sC[2][STAGE_D];
  __shared__ __attribute__((aligned(1024))) uint8_t sB[2][STAGE_D];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * k_blocks) return;
  const int64_t kbk = tile / row_blocks, rb = tile - kbk * row_blocks;
  const int64_t r0 = rb * BR, k0 = kbk * BK;
  const int64_t wpr = (K + 31) / 32;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;

  {  // skip an empty 128 x 128 bitmap block (all waves agree through the barrier)
    uint32_t any = 0;
    const int row = tid >> 1;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int64_t w = k0 / 32 + 2 * (tid & 1) + x;
      if (r0 + row < M && w < wpr) any |= W[(r0 + row) * ldw + w] & valid_bits(w, K);
    }
    if (!__syncthreads_or(any != 0u)) return;
  }

  const uint16_t* csrc[NC_W];
  const uint16_t* bsrc[NC_W];
#pragma unroll
  for (int t = 0; t < NC_W; ++t) {
    const int g = wave * NC_W + t;  // 1 KiB = 8 staged rows
    const int r = 8 * g + lane / 8;
    const int c = (lane & 7) ^ ((r >> 1) & 7);
    int64_t rc = r0 + r, rk = k0 + r;
    if (rc > M - 1) rc = M - 1;
    if (rk > K - 1) rk = K - 1;
    csrc[t] = dC + rc * lddc + 8 * c;
    bsrc[t] = B + rk * ldb + 8 * c;
  }
  const uint32_t sC0 = lds_u32(&sC[0][0]), sB0 = lds_u32(&sB[0][0]);
  auto issue = [&](int64_t st) {
    const int buf = (int)(st & 1);
#pragma unroll
    for (int t = 0; t < NC_W; ++t) {
      const uint32_t off = (uint32_t)(buf * STAGE_D + (wave * NC_W + t) * 1024);
      dma16(csrc[t] + st * NSTG, sC0 + off);
      dma16(bsrc[t] + st * NSTG, sB0 + off);
    }
  };

  f32x4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int64_t nst = N / NSTG;
  issue(0);
  for (int64_t st = 0; st < nst; ++st) {
    const int buf = (int)(st & 1);
    wait_vmcnt<0>();
    __syncthreads();
    if (st + 1 < nst) issue(st + 1);
    // every fragment of the stage is requested before its first MFMA
    uint4 af[NSTG / 32][4], bf[NSTG / 32][4];
#pragma unroll
    for (int ks = 0; ks < NSTG / 32; ++ks) {
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int r = 64 * wm + 16 * a + j;
        af[ks][a] = *reinterpret_cast<const uint4*>(
            &sC[buf][r * ROWB_D + 16 * ((4 * ks + q) ^ ((r >> 1) & 7))]);
      }
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int r = 64 * wn + 16 * b + j;
        bf[ks][b] = *reinterpret_cast<const uint4*>(
            &sB[buf][r * ROWB_D + 16 * ((4 * ks + q) ^ ((r >> 1) & 7))]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < NSTG / 32; ++ks)
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = mfma<PREC>(af[ks][a], bf[ks][b], acc[a][b]);
  }

  store_set_bits_128(acc, &sC[0][0], W, M, K, ldw, row_ptr, woff, dA, r0, k0, wm, wn, j, q);
}

template <int PREC, int TB>
bws_status launch_sddmm(const bws_bitmap* A, const int32_t* woff, const void* dC, int64_t lddc,
                        const void* B, int64_t ldb, int64_t N, float* dA, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, BR);
  const int64_t k_blocks = ceil_div(A->cols, BK);
  const int64_t per_xcd = ceil_div(row_blocks * k_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const size_t esz = TB == 0 ? 4 : 2;
  const int vec = (ldb % (16 / esz) == 0) && (lddc % (16 / esz) == 0) &&
                  (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                  (reinterpret_cast<uintptr_t>(dC) % 16 == 0);
  hipLaunchKernelGGL((k_sddmm_mfma<PREC, TB>), dim3((unsigned)(8 * per_xcd)), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, A->row_ptr, woff, dC, lddc, B, ldb, N,
                     dA, row_blocks, k_blocks, per_xcd, vec);
  return launch_check();
}

}  // namespace

int64_t sddmm_f32_tiles(int64_t rows, int64_t cols);

// Workspace of bws_sddmm: per-word popcount offsets (int32, rows x words),
// then one byte per 64 x 64 tile for the FP32 tiled kernel.
size_t sddmm_woff_bytes(const bws_bitmap* A) {
  return align_up((size_t)(A->rows * words_for(A->cols)) * sizeof(int32_t), 256);
}
size_t sddmm_mfma_workspace(const bws_bitmap* A) {
  return sddmm_woff_bytes(A) + align_up((size_t)sddmm_f32_tiles(A->rows, A->cols), 256) + 256;
}

bws_status sddmm_mfma_launch(const bws_bitmap* A, const void* dC, const void* B, int32_t dtype,
                             int64_t lddc, int64_t ldb, int64_t N, float* dA, int32_t precision,
                             void* ws, hipStream_t s) {
  int32_t* woff = static_cast<int32_t*>(ws);
  BWS_TRY(launch_word_offsets(A->words, A->rows, A->cols, A->ldw, woff, s));
  // direct-to-LDS staging: 16-bit operands in the precision's own type, rows
  // 16-byte aligned
  const bool dma = ((precision == BWS_PREC_BF16 && dtype == BWS_DTYPE_BF16) ||
                    (precision == BWS_PREC_FP16 && dtype == BWS_DTYPE_F16)) &&
                   lddc % 8 == 0 && ldb % 8 == 0 && reinterpret_cast<uintptr_t>(dC) % 16 == 0 &&
                   reinterpret_cast<uintptr_t>(B) % 16 == 0;
  // 256 x 256 blocks once there are enough of them to occupy most CUs
  if (dma && N >= 64 && N % 64 == 0 && sddmm_big_tiles(A->rows, A->cols) >= 128)
    return launch_sddmm_big(precision, A, woff, dC, lddc, B, ldb, N, dA, s);
  if (dma && N >= NSTG && N % NSTG == 0) {
    const int64_t row_blocks = ceil_div(A->rows, BR), k_blocks = ceil_div(A->cols, BK);
    const int64_t per_xcd = ceil_div(row_blocks * k_blocks, 8);
    if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
    const dim3 grid((unsigned)(8 * per_xcd));
    const uint16_t* c16 = static_cast<const uint16_t*>(dC);
    const uint16_t* b16 = static_cast<const uint16_t*>(B);
    if (precision == BWS_PREC_FP16)
      hipLaunchKernelGGL(k_sddmm_mfma_dma<2>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                         A->ldw, A->row_ptr, woff, c16, lddc, b16, ldb, N, dA, row_blocks,
                         k_blocks, per_xcd);
    else
      hipLaunchKernelGGL(k_sddmm_mfma_dma<1>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                         A->ldw, A->row_ptr, woff, c16, lddc, b16, ldb, N, dA, row_blocks,
                         k_blocks, per_xcd);
    return launch_check();
  }
  if (precision == BWS_PREC_BF16) {
    if (dtype == BWS_DTYPE_BF16) return launch_sddmm<1, 1>(A, woff, dC, lddc, B, ldb, N, dA, s);
    if (dtype == BWS_DTYPE_F32) return launch_sddmm<1, 0>(A, woff, dC, lddc, B, ldb, N, dA, s);
  } else if (precision == BWS_PREC_FP16) {
    if (dtype == BWS_DTYPE_F16) return launch_sddmm<2, 2>(A, woff, dC, lddc, B, ldb, N, dA, s);
    if (dtype == BWS_DTYPE_F32) return launch_sddmm<2, 0>(A, woff, dC, lddc, B, ldb, N, dA, s);
  } else if (precision == BWS_PREC_BF16X2) {
    if (dtype == BWS_DTYPE_F32) return launch_sddmm<3, 0>(A, woff, dC, lddc, B, ldb, N, dA, s);
  }
  return BWS_EUNSUPPORTED;
}

}  // namespace bws

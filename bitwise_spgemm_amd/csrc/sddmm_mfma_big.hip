// sddmm_mfma_big.hip — the 16-bit MFMA SDDMM with a 256 x 256 output block:
//   dA[row_ptr[r] + rank(r, k)] = Σ_n dC[r, n] · B[k, n]   for every set bit (r, k).
//
// Four waves of 128 x 128 (8 x 8 MFMA tiles, 256 FP32 accumulators each, held
// in AGPRs: this file is compiled with the AGPR form of v_mfma), one wave per
// SIMD.  Both operands are staged by global_load_lds_dwordx4, 64 values of N
// per stage (128-byte rows, 16-byte chunk c of row r at c ^ ((r >> 1) & 7), so
// the 16 lanes of a fragment read hit 16 different bank groups), two stages:
// the copy of stage s + 1 is issued while stage s is multiplied (128 MFMAs
// per wave).  Per 32-wide K step a wave reads 8 + 8 fragments for 64 MFMAs
// (the 128 x 128-block kernel: 8 + 8 for 32).  Each output's N steps run in
// ascending order, 32 at a time, as in the other MFMA SDDMM kernels, so all of
// them give the same result.  A stage's 32 fragment reads are all issued
// before its first MFMA (3% faster than reading each row of tiles' A fragment
// just before its MFMAs); 32-value stages with four in flight measured 3-8%
// slower.
#include "mfma_common.hpp"

namespace bws {
namespace {

constexpr int BT = 256;               // output block: BT rows of A x BT columns of A
constexpr int NS = 64;                // N values per stage
constexpr int NBUF = 2;               // stages in LDS
constexpr int ROWB = NS * 2;          // bytes per staged row
constexpr int CPR = ROWB / 16;        // 16-byte chunks per staged row
constexpr int RPK = 1024 / ROWB;      // staged rows per 1 KiB copy
constexpr int STAGE = BT * ROWB;      // bytes per operand per stage
constexpr int NC_W = STAGE / 1024 / 4;  // 1 KiB copies per wave per operand per stage
constexpr int KSTEPS = NS / 32;       // MFMA K steps per stage

// LDS slot of 16-byte chunk c of staged row r: rows are swizzled so that the
// 16 lanes reading one chunk of 16 consecutive rows hit 16 different bank
// groups (64 banks x 4 bytes)
__device__ __forceinline__ int slot_of(int r, int c) { return c ^ ((r / (16 / CPR)) % CPR); }

template <int PREC>
__global__ void __launch_bounds__(256, 1)
k_sddmm_mfma_big(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
                 const uint16_t* __restrict__ dC, int64_t lddc, const uint16_t* __restrict__ B,
                 int64_t ldb, int64_t N, float* __restrict__ dA, int64_t row_blocks,
                 int64_t k_blocks, int64_t per_xcd) {
  __shared__ __attribute__((aligned(1024))) uint8_t sC[NBUF][STAGE];
  __shared__ __attribute__((aligned(1024))) uint8_t sB[NBUF][STAGE];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * k_blocks) return;
  const int64_t kbk = tile / row_blocks, rb = tile - kbk * row_blocks;
  const int64_t r0 = rb * BT, k0 = kbk * BT;
  const int64_t wpr = (K + 31) / 32;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;

  {  // skip an empty 256 x 256 bitmap block (all waves agree through the barrier)
    uint32_t any = 0;
    const int64_t w = k0 / 32 + (tid & 7);
#pragma unroll
    for (int x = 0; x < BT / 32; ++x) {
      const int64_t row = r0 + (tid >> 3) + 32 * x;
      if (row < M && w < wpr) any |= W[row * ldw + w] & valid_bits(w, K);
    }
    if (!__syncthreads_or(any != 0u)) return;
  }

  // per-lane copy sources: instruction g of a stage fills LDS rows RPK g ..
  // RPK g + RPK-1; lane l writes 16 bytes at slot l % CPR of row
  // RPK g + l / CPR, which holds the row's global chunk slot_of(row, l % CPR)
  // (the swizzle is its own inverse)
  const uint16_t* csrc[NC_W];
  const uint16_t* bsrc[NC_W];
#pragma unroll
  for (int t = 0; t < NC_W; ++t) {
    const int g = wave * NC_W + t;
    const int r = RPK * g + lane / CPR;
    const int c = slot_of(r, lane % CPR);
    int64_t rc = r0 + r, rk = k0 + r;
    if (rc > M - 1) rc = M - 1;
    if (rk > K - 1) rk = K - 1;
    csrc[t] = dC + rc * lddc + 8 * c;
    bsrc[t] = B + rk * ldb + 8 * c;
  }
  const uint32_t sC0 = lds_u32(&sC[0][0]), sB0 = lds_u32(&sB[0][0]);
  auto issue = [&](int64_t st) {
    const uint32_t buf = (uint32_t)(st % NBUF) * STAGE;
#pragma unroll
    for (int t = 0; t < NC_W; ++t) {
      const uint32_t off = buf + (uint32_t)((wave * NC_W + t) * 1024);
      dma16(csrc[t] + st * NS, sC0 + off);
      dma16(bsrc[t] + st * NS, sB0 + off);
    }
  };

  f32x4 acc[8][8];
#pragma unroll
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int b = 0; b < 8; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  const int64_t nst = N / NS;
#pragma unroll
  for (int p = 0; p < NBUF - 1; ++p)
    if (p < nst) issue(p);
  // fragment rows 16x + j of the wave's 128 rows (the swizzle depends only on
  // j since 16x is a multiple of 16)
  const uint32_t row_c = (uint32_t)((128 * wm + j) * ROWB);
  const uint32_t row_b = (uint32_t)((128 * wn + j) * ROWB);

  for (int64_t st = 0; st < nst; ++st) {
    // stage st has landed when at most the later stages' copies are out
    // (up to NBUF - 2 stages were issued after it)
    {
      const int64_t later = nst - 1 - st < NBUF - 2 ? nst - 1 - st : NBUF - 2;
      static_assert(NBUF >= 2 && NBUF <= 5, "two to five stages");
      if (NBUF >= 5 && later >= 3) wait_vmcnt<(NBUF >= 5 ? 3 : 0) * 2 * NC_W>();
      else if (NBUF >= 4 && later >= 2) wait_vmcnt<(NBUF >= 4 ? 2 : 0) * 2 * NC_W>();
      else if (NBUF >= 3 && later >= 1) wait_vmcnt<(NBUF >= 3 ? 1 : 0) * 2 * NC_W>();
      else wait_vmcnt<0>();
    }
    __syncthreads();  // every wave's copies of stage st; stage st-1's buffer is free
    if (st + NBUF - 1 < nst) issue(st + NBUF - 1);
    const uint8_t* bufC = &sC[0][0] + (uint32_t)(st % NBUF) * STAGE;
    const uint8_t* bufB = &sB[0][0] + (uint32_t)(st % NBUF) * STAGE;
    // every fragment of the stage is requested before its first MFMA: with
    // one wave per SIMD nothing else hides an LDS read's latency, and left to
    // itself the compiler reloads one A fragment register per row of tiles
    uint4 af[KSTEPS][8], bf[KSTEPS][8];
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      const uint32_t chunk = (uint32_t)(16 * slot_of(j, 4 * ks + q));
#pragma unroll
      for (int a = 0; a < 8; ++a)
        af[ks][a] = *reinterpret_cast<const uint4*>(bufC + row_c + chunk + a * 16 * ROWB);
#pragma unroll
      for (int b = 0; b < 8; ++b)
        bf[ks][b] = *reinterpret_cast<const uint4*>(bufB + row_b + chunk + b * 16 * ROWB);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks)
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int b = 0; b < 8; ++b) acc[a][b] = mfma<PREC>(af[ks][a], bf[ks][b], acc[a][b]);
  }

  // Epilogue.  The block's bitmap words and the value index of each word's
  // first bit are staged in LDS (over the operand buffers), so the stores
  // below only read LDS: lane (j, q) holds rows 4q + e of each 16-row tile
  // and column j of each 16-column tile; only set bits are stored.
  __syncthreads();  // every wave is done with the operand buffers
  uint32_t* eW = reinterpret_cast<uint32_t*>(&sC[0][0]);        // [BT][BT / 32] words
  int64_t* eBase = reinterpret_cast<int64_t*>(&sC[1][0]);       // [BT][BT / 32] value indices
  for (int t = tid; t < BT * (BT / 32); t += 256) {
    const int64_t r = r0 + (t >> 3), w = k0 / 32 + (t & 7);
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
  for (int a = 0; a < 8; ++a)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int rr = 128 * wm + 16 * a + 4 * q + e;
#pragma unroll
      for (int b = 0; b < 8; ++b) {
        const int ew = rr * (BT / 32) + 4 * wn + (b >> 1);
        const uint32_t word = eW[ew];
        const int i = 16 * (b & 1) + j;
        if (word & wbit(i)) dA[eBase[ew] + rank_before(word, i)] = acc[a][b][e];
      }
    }
}

}  // namespace

// 16-bit dC and B with 16-byte aligned rows, leading dimensions multiples of 8
// and N a multiple of 64 (the caller checks); PREC 1 = BF16, 2 = FP16.
bws_status launch_sddmm_big(int32_t precision, const bws_bitmap* A, const int32_t* woff,
                            const void* dC, int64_t lddc, const void* B, int64_t ldb, int64_t N,
                            float* dA, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, BT), k_blocks = ceil_div(A->cols, BT);
  const int64_t per_xcd = ceil_div(row_blocks * k_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const dim3 grid((unsigned)(8 * per_xcd));
  const uint16_t* c16 = static_cast<const uint16_t*>(dC);
  const uint16_t* b16 = static_cast<const uint16_t*>(B);
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL(k_sddmm_mfma_big<2>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                       A->ldw, A->row_ptr, woff, c16, lddc, b16, ldb, N, dA, row_blocks, k_blocks,
                       per_xcd);
  else
    hipLaunchKernelGGL(k_sddmm_mfma_big<1>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                       A->ldw, A->row_ptr, woff, c16, lddc, b16, ldb, N, dA, row_blocks, k_blocks,
                       per_xcd);
  return launch_check();
}

int64_t sddmm_big_tiles(int64_t rows, int64_t cols) {
  return ceil_div(rows, BT) * ceil_div(cols, BT);
}

}  // namespace bws

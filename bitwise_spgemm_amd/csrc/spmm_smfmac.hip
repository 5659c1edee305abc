// spmm_smfmac.hip — 16-bit SpMM of a 0/1 bitmap on the sparse matrix cores
// (gfx950 v_smfmac_f32_16x16x64_{bf16,f16}, FP32 accumulate).
//
// The sparse MFMA multiplies a 16 x 64 A block stored 4:2-compressed (of
// every 4 consecutive K values at most 2 are kept, with a 2-bit position
// each) by a dense 64 x 16 B block: the work of two dense 16x16x32 MFMAs in
// one instruction.  A 0/1 bitmap at the densities of this library almost
// always fits: a nibble of the mask (4 K columns of one row) with at most 2
// bits becomes two kept values (1 or 0) and their positions.  Nibbles with 3
// or 4 bits keep their first two; the rest (the overflow, at 5% density one
// row-nibble in 2000) is multiplied by a second sparse MFMA, issued only for
// the 16-row tiles and 64-row stages of a wave that have any overflow.
//
// Operand layout of v_smfmac_f32_16x16x64 (determined on the device with a
// probe of the builtin, one wave, in-bounds):
//  - B: lane l holds column l % 16 and K rows 16*(l/16) .. +15 (16 values);
//  - A: lane l holds row l % 16 and 8 kept values; values 2p and 2p + 1 come
//    from K group G(q, p) = 4*(2*(q%2) + p/2) + 2*(q/2) + p%2 (q = l/16, K
//    rows 4G .. 4G+3), value s at position (idx >> 2s) & 3 of its group;
//  - D: lane l holds column l % 16 and rows 4*(l/16) .. +3.
// With the stage mask bit-reversed (bit k = K row k), lane group q's groups
// are the nibbles of bytes 4*(q%2) + q/2 and that + 2.  A 256-entry LDS
// table of 32-bit words turns a mask byte into the positions of its two
// nibbles' kept values, whether each kept value is 1 and the overflow bits;
// the 16-bit operand values are built from those bits in registers.
//
// Tiles and B staging are those of the direct-to-LDS dense kernel
// (mfma4_tile.hpp): 4 waves, wave tile WM x WN, 64 K rows per stage copied
// HBM -> LDS by global_load_lds_dwordx4 (swizzled for these reads), B fragments by
// the transposing LDS read (four per 16 K rows).  Stages whose WM x 64 mask
// block is empty are skipped.  The summation order over K differs from the
// dense kernel's, so results agree with it to FP32 rounding, not bit for bit.
#include "bws_internal.hpp"
#include "mfma_common.hpp"

namespace bws {
namespace {

typedef __attribute__((ext_vector_type(16))) __bf16 bf16x16;
typedef __attribute__((ext_vector_type(16))) _Float16 f16x16;

struct U8x2 {
  uint4 lo, hi;
};

template <int PREC>
__device__ __forceinline__ f32x4 smfmac(const uint4& a, const U8x2& b, const f32x4& c, int idx) {
  if constexpr (PREC == 2) {
    return __builtin_amdgcn_smfmac_f32_16x16x64_f16(__builtin_bit_cast(f16x8, a),
                                                    __builtin_bit_cast(f16x16, b), c, idx, 0, 0);
  } else {
    return __builtin_amdgcn_smfmac_f32_16x16x64_bf16(__builtin_bit_cast(bf16x8, a),
                                                     __builtin_bit_cast(bf16x16, b), c, idx, 0, 0);
  }
}

// Table entry of a mask byte x (bit k = K column k of its 8): bits 0-7 = the
// four positions of the kept values (2 bits each, low nibble first), bits
// 8-15 = the bits of x not kept (at most 2 per nibble), bits 16-19 = whether
// the kept values are 1 (low nibble's first, second, high nibble's first,
// second).  Per nibble v the first kept value is 1 at the position of v's
// lowest bit (0 at position 0 when v is empty), the second 1 at the position
// of v's next bit (0, at the position after the first, when v has one bit).
__device__ __forceinline__ uint32_t lut_entry(uint32_t x) {
  uint32_t pos = 0, ovf = 0, one = 0;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const uint32_t v = (x >> (4 * h)) & 15u;
    const uint32_t r = v & (v - 1u);  // v without its lowest bit
    const uint32_t ia = v ? (uint32_t)__builtin_ctz(v) : 0u;
    const uint32_t ib = r ? (uint32_t)__builtin_ctz(r) : ((ia + 1u) & 3u);
    pos |= (ia | (ib << 2)) << (4 * h);
    ovf |= (r & (r - 1u)) << (4 * h);
    one |= ((v ? 1u : 0u) | (r ? 2u : 0u)) << (2 * h);
  }
  return pos | (ovf << 8) | (one << 16);
}

// The kept values of the two nibbles of a table entry e as two 32-bit words
// of 16-bit values (first kept value in the low half): with t = e's bits
// 16-19 and z = t | (t << 15), bit 0 and bit 16 of z are the low nibble's
// values and bits 2 and 18 the high nibble's, so (z & 0x10001) * ONE and
// (z & 0x40004) * (ONE >> 2) place 1.0 where they are set (ONE >> 2 is exact:
// ONE has at least 7 trailing zero bits).
template <int PREC>
__device__ __forceinline__ uint2 lut_values(uint32_t e) {
  constexpr uint32_t ONE = PREC == 2 ? 0x3C00u : 0x3F80u;
  const uint32_t t = __builtin_amdgcn_ubfe(e, 16, 4);
  const uint32_t z = t | (t << 15);
  return make_uint2((uint32_t)__umul24(z & 0x00010001u, ONE),
                    (uint32_t)__umul24(z & 0x00040004u, ONE >> 2));
}

template <int WM, int WN, int NBUF>
struct CfgS {
  static constexpr int BM = 2 * WM, BN = 2 * WN, KS = 64, KW = 2;
  static constexpr int TM = WM / 16, TN = WN / 16;
  static constexpr int ROWB = BN * 2;           // bytes per B row in LDS
  static constexpr int SB = KS * ROWB;          // bytes per B stage
  static constexpr int NB_W = SB / 1024 / 4;    // B DMA instructions per wave per stage
  static constexpr int NA_W = BM * KW / 64 / 4; // word DMA instructions per wave
  static constexpr int PER_STAGE = NB_W + NA_W;
  static constexpr int LPR = ROWB / 16;         // DMA lanes per B row
};

template <int PREC, int WM, int WN, int NBUF>
__global__ void __launch_bounds__(256, 2)
k_spmm_smfmac(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
              const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
              int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
              int64_t per_xcd) {
  using G = CfgS<WM, WN, NBUF>;
  constexpr int BM = G::BM, BN = G::BN, KS = G::KS, KW = G::KW, TM = G::TM, TN = G::TN;
  constexpr int ROWB = G::ROWB, SB = G::SB, NB_W = G::NB_W, NA_W = G::NA_W, LPR = G::LPR;
  static_assert(NB_W >= 1 && NA_W >= 1, "stage must give every wave DMA work");
  __shared__ __attribute__((aligned(1024))) uint8_t sB[NBUF][SB];
  __shared__ __attribute__((aligned(256))) uint32_t sA[NBUF][BM * KW];
  __shared__ uint32_t lut[256];

  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t m0 = rb * BM, n0 = cb * BN;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;
  const int64_t nst = (K + KS - 1) / KS;
  const int64_t wpr = (K + 31) / 32;

  lut[tid] = lut_entry((uint32_t)tid);  // visible after the first stage's barrier

  // per-lane DMA sources (as in the dense direct-to-LDS kernel): B row
  // r0 + t * (1024 / ROWB) of the stage, column bcol[t] (swizzled per row)
  const int r0 = wave * NB_W * (1024 / ROWB) + lane / LPR;
  int32_t bcol[NB_W];
#pragma unroll
  for (int t = 0; t < NB_W; ++t) {
    const int r = r0 + t * (1024 / ROWB);
    const int u = lane % LPR;
    const int slot = u >> 1, half = u & 1;
    const int c = slot ^ ((r & 3) | ((r >> 2) & 4));
    int64_t col = n0 + 16 * c + 8 * half;
    if (col + 8 > N) col = N - 8;
    bcol[t] = (int32_t)col;
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
  auto buf_of = [](int64_t st) -> int {
    return (NBUF & (NBUF - 1)) == 0 ? (int)(st & (NBUF - 1)) : (int)(st % NBUF);
  };
  auto issue = [&](int64_t st) {
    const int buf = buf_of(st);
#pragma unroll
    for (int t = 0; t < NB_W; ++t) {
      int64_t k = st * KS + r0 + t * (1024 / ROWB);
      if (k > K - 1) k = K - 1;
      dma16(B + k * ldb + bcol[t], sB0 + (uint32_t)(buf * SB + (wave * NB_W + t) * 1024));
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

  // transposing reads: lane group q reads K rows 16q + 4h + i4 (h = 0..3),
  // 8-byte piece j % 4 of the 32-byte unit of its column block; the unit of
  // row k sits at slot c ^ f(k), f(k) = (k & 3) | ((k >> 2) & 4), which for
  // these rows is i4 | 4*(q % 2): the 32 lanes of a read (q = 0, 1 or 2, 3)
  // touch 8 different slots, all banks
  const int i4 = j >> 2;
  const int fsw = i4 | ((q & 1) << 2);
  const uint32_t rd_lane = (uint32_t)((16 * q + i4) * ROWB + 8 * (j & 3));
  const int c0 = (WN * wn) / 16;
  // the mask bytes of lane group q: byte 4*(q%2) + q/2 and that + 2
  const int wsel = q & 1, bsh = 8 * (q >> 1);

  for (int64_t st = 0; st < nst; ++st) {
    const int buf = buf_of(st);
    if (st + NBUF - 2 < nst) wait_vmcnt<(NBUF - 2) * G::PER_STAGE>();
    else wait_vmcnt<0>();
    __syncthreads();
    if (st + NBUF - 1 < nst) issue(st + NBUF - 1);

    uint32_t wr[TM];  // this lane's mask word of each row tile
    uint32_t any = 0;
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const uint2 x = *reinterpret_cast<const uint2*>(&sA[buf][(WM * wm + 16 * a + j) * KW]);
      const uint32_t w0 = __brev(x.x & valid_bits(st * KW, K));
      const uint32_t w1 = __brev(x.y & valid_bits(st * KW + 1, K));
      any |= w0 | w1;
      wr[a] = wsel ? w1 : w0;
    }
    if (__ballot(any != 0u) == 0ull) continue;  // this wave's mask block is empty

    // the row tiles' A operands from the table first, then per half of the
    // column tiles their B fragments (32 registers) and the sparse MFMAs; the
    // overflow of a row tile, when any lane has one, is a second sparse MFMA
    // on the same B fragments
    uint4 af[TM];
    int ia[TM];
    uint32_t ov[TM];
#pragma unroll
    for (int a = 0; a < TM; ++a) {
      const uint32_t e0 = lut[__builtin_amdgcn_ubfe(wr[a], bsh, 8)];
      const uint32_t e2 = lut[__builtin_amdgcn_ubfe(wr[a], bsh + 16, 8)];
      const uint2 v0 = lut_values<PREC>(e0), v2 = lut_values<PREC>(e2);
      af[a] = make_uint4(v0.x, v0.y, v2.x, v2.y);
      ia[a] = (int)((e0 & 0xFFu) | ((e2 & 0xFFu) << 8));
      ov[a] = ((e0 >> 8) & 0xFFu) | (e2 & 0xFF00u);  // overflow bytes: x0's low, x2's high
    }
    __builtin_amdgcn_s_setprio(1);
    const uint32_t sbase = sB0 + (uint32_t)(buf * SB) + rd_lane;
    constexpr int TH = TN >= 4 ? TN / 2 : TN;
#pragma unroll
    for (int b0 = 0; b0 < TN; b0 += TH) {
      U8x2 bf[TH];
#pragma unroll
      for (int b = 0; b < TH; ++b) {
        uint2 v[4];
#pragma unroll
        for (int h = 0; h < 4; ++h)
          v[h] = tr_read(sbase + (uint32_t)(4 * h * ROWB) + (uint32_t)(32 * ((c0 + b0 + b) ^ fsw)));
        bf[b].lo = make_uint4(v[0].x, v[0].y, v[1].x, v[1].y);
        bf[b].hi = make_uint4(v[2].x, v[2].y, v[3].x, v[3].y);
      }
#pragma unroll
      for (int a = 0; a < TM; ++a) {
#pragma unroll
        for (int b = 0; b < TH; ++b)
          acc[a][b0 + b] = smfmac<PREC>(af[a], bf[b], acc[a][b0 + b], ia[a]);
        if (__ballot(ov[a] != 0u) != 0ull) {  // nibbles with 3 or 4 bits in this row tile
          const uint32_t f0 = lut[ov[a] & 0xFFu], f2 = lut[ov[a] >> 8];
          const uint2 g0 = lut_values<PREC>(f0), g2 = lut_values<PREC>(f2);
          const uint4 ag = make_uint4(g0.x, g0.y, g2.x, g2.y);
          const int ig = (int)((f0 & 0xFFu) | ((f2 & 0xFFu) << 8));
#pragma unroll
          for (int b = 0; b < TH; ++b)
            acc[a][b0 + b] = smfmac<PREC>(ag, bf[b], acc[a][b0 + b], ig);
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
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

// The sparse-MFMA product for 16-bit B (BF16 or FP16 by precision): N and ldb
// multiples of 8, B 16-byte aligned (checked by the caller).
template <int PREC, int WM, int WN, int NBUF>
bws_status launch_smfmac_t(const bws_bitmap* A, const uint16_t* B, int64_t ldb, int64_t N,
                           float* C, int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, 2 * WM);
  const int64_t col_blocks = ceil_div(N, 2 * WN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL((k_spmm_smfmac<PREC, WM, WN, NBUF>), dim3((unsigned)(8 * per_xcd)), dim3(256),
                     0, s, A->words, A->rows, A->cols, A->ldw, B, ldb, N, C, ldc, alpha, beta,
                     row_blocks, col_blocks, per_xcd);
  return launch_check();
}

#ifndef BWS_SMFMAC_NO_PRODUCT
bws_status launch_smfmac(int32_t precision, const bws_bitmap* A, const void* B, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         hipStream_t s) {
  const uint16_t* b = static_cast<const uint16_t*>(B);
  if (precision == BWS_PREC_FP16)
    return launch_smfmac_t<2, 64, 128, 2>(A, b, ldb, N, C, ldc, alpha, beta, s);
  return launch_smfmac_t<1, 64, 128, 2>(A, b, ldb, N, C, ldc, alpha, beta, s);
}
#endif

}  // namespace bws

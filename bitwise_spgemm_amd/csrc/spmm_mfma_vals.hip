// spmm_mfma_vals.hip — 16-bit MFMA SpMM for a value-carrying bitmap A and for
// an int8 A held as bit-planes (gfx950, v_mfma_f32_16x16x32_{bf16,f16}, FP32
// accumulate).
//
// A 0/1 operand is built from the bitmap bits inside the MFMA kernels; a
// value-carrying A needs each set bit's own value in the fragment, a gather
// by popcount rank.  Here that gather runs once, as a pass that expands the
// bitmap and its values into a dense 16-bit copy of A in the caller's
// workspace (values rounded to nearest even, zeros elsewhere, rows padded to a
// multiple of 64 columns); the product then streams both operands straight
// into LDS:
//  - A stages (128 rows x 64 K, row-major, 128-byte rows) are copied by
//    global_load_lds_dwordx4; 16-byte chunk c of row r sits at chunk
//    c ^ ((r >> 1) & 7), so the 16 lanes of a fragment read (rows j, the same
//    chunk) fall in 16 different bank groups.  A fragment is one 16-byte read.
//  - B exactly as in the direct-to-LDS 0/1 kernel (swizzled rows, transposing
//    reads).
// Workgroup = 4 waves (2 x 2), C tile 128 x 128, two LDS stages, so two
// workgroups fit on a CU.
#include "mfma_common.hpp"

namespace bws {
namespace {

constexpr int VBM = 128, VBN = 128, VKS = 64, VWM = 64, VWN = 64;
constexpr int VTM = VWM / 16, VTN = VWN / 16;
constexpr int A_ROWB = VKS * 2;                  // 128 bytes per A row in LDS
constexpr int A_STAGE = VBM * A_ROWB;            // 16 KiB
constexpr int B_ROWB = VBN * 2;                  // 256 bytes per B row in LDS
constexpr int B_STAGE = VKS * B_ROWB;            // 16 KiB
constexpr int NA_W = A_STAGE / 1024 / 4;         // A copies per wave per stage
constexpr int NB_W = B_STAGE / 1024 / 4;         // B copies per wave per stage
constexpr int PER_STAGE = NA_W + NB_W;
constexpr int NBUF = 2;

// Dense 16-bit A: one thread per word of the padded row (lda / 32 words).
//  - A wave's 64 words are consecutive (row, word) positions and values are
//    stored row by row, so the wave's values are one contiguous run, from its
//    first word with a set bit to the end of its last (at most 2048).  The
//    wave copies that run into LDS with coalesced loads, rounding each value
//    once, and every lane reads its word's values from there by rank.
//  - Since lda = 32 x (words per padded row), thread t's 64 output bytes sit
//    at byte 64 t of Ad: the workgroup's output is one contiguous 16 KiB run,
//    staged in LDS (over the values) and stored 4 KiB per instruction.
template <int PREC>
__global__ void __launch_bounds__(256)
k_expand_values(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
                const float* __restrict__ values, uint16_t* __restrict__ Ad, int64_t lda) {
  __shared__ uint4 stage[256 * 4];
  const int64_t wpr = words_for(K);
  const int64_t wpad = lda / 32;
  const int64_t n = M * wpad;
  const int64_t stride = (int64_t)gridDim.x * 256;
  const int tid = threadIdx.x, lane = tid & 63;
  uint16_t* sv = reinterpret_cast<uint16_t*>(stage) + 2048 * (tid >> 6);  // this wave's values
  for (int64_t t0 = (int64_t)blockIdx.x * 256; t0 < n; t0 += stride) {
    const int64_t t = t0 + tid;
    const int64_t r = t / wpad, w = t - r * wpad;
    uint32_t m = 0u;
    int64_t e = 0;
    if (t < n && w < wpr) {
      m = W[r * ldw + w] & valid_bits(w, K);
      e = row_ptr[r] + woff[r * wpr + w];
    }
    const uint64_t act = __ballot(m != 0u);
    int64_t e0 = 0;
    __syncthreads();  // the previous run's stores have read the stage
    if (act) {
      const int first = __ffsll((unsigned long long)act) - 1;
      const int last = 63 - __clzll((long long)act);
      e0 = ((int64_t)__builtin_amdgcn_readlane((int)(e >> 32), first) << 32) |
           (uint32_t)__builtin_amdgcn_readlane((int)e, first);
      const int64_t e_end = e + __popc(m);
      const int64_t e1 = ((int64_t)__builtin_amdgcn_readlane((int)(e_end >> 32), last) << 32) |
                         (uint32_t)__builtin_amdgcn_readlane((int)e_end, last);
      const int cnt = e1 - e0 < 2048 ? (int)(e1 - e0) : 2048;  // 2048 unless row_ptr is not A's
      for (int k = lane; k < cnt; k += 64)
        sv[k] = (uint16_t)(PREC == 2 ? f32_to_f16_rne(values[e0 + k])
                                     : f32_to_bf16_rne(values[e0 + k]));
    }
    __syncthreads();
    // bit i's value sits at (set bits before i) past the word's first; every
    // position reads (a clear bit at an index clamped into the word's own
    // values, 0 for an empty word) and the clear ones are zeroed afterwards
    uint32_t d[16];
    {
      const int base = m ? (int)(e - e0) : 0;
      const int top = m ? __popc(m) - 1 : 0;
      auto at = [&](int i) -> uint32_t {
        const int before = i == 0 ? 0 : __popc(m >> (32 - i));
        return sv[base + (before < top ? before : top)];
      };
#pragma unroll
      for (int i = 0; i < 16; ++i) d[i] = at(2 * i) | (at(2 * i + 1) << 16);
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const uint32_t keep = ((m & (0x80000000u >> (2 * i))) ? 0x0000FFFFu : 0u) |
                              ((m & (0x40000000u >> (2 * i))) ? 0xFFFF0000u : 0u);
        d[i] &= keep;
      }
    }
    __syncthreads();  // every wave has read its values (the stage overlays them)
#pragma unroll
    for (int i = 0; i < 4; ++i)
      stage[4 * tid + i] = make_uint4(d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]);
    __syncthreads();
    const int64_t chunks = 4 * (n - t0 < 256 ? n - t0 : 256);
    uint4* dst = reinterpret_cast<uint4*>(Ad) + 4 * t0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int c = 256 * i + tid;
      if (c < chunks) dst[c] = stage[c];
    }
  }
}

// Dense 16-bit A from int8 bit-planes P[8][M][ldw]: a thread per (row, word):
// value(c) = Σ_b w_b·bit_b(c) with w = [1, 2, 4, ..., 64, -128] (exact in
// BF16 and FP16: integers of magnitude <= 128), written as 32 16-bit values.
template <int PREC>
__global__ void __launch_bounds__(256)
k_expand_planes(const uint32_t* __restrict__ P, int64_t M, int64_t K, int64_t ldw,
                uint16_t* __restrict__ Ad, int64_t lda) {
  const int64_t wpr = words_for(K);
  const int64_t wpad = lda / 32;
  const int64_t n = M * wpad;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / wpad, w = t - r * wpad;
    uint32_t pl[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      pl[b] = w < wpr ? P[((int64_t)b * M + r) * ldw + w] & valid_bits(w, K) : 0u;
    uint32_t d[16];
#pragma unroll
    for (int i = 0; i < 32; ++i) {
      int v = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) v += ((pl[b] >> (31 - i)) & 1u) ? (b == 7 ? -128 : (1 << b)) : 0;
      const uint32_t h = PREC == 2 ? f32_to_f16_rne((float)v) : f32_to_bf16_rne((float)v);
      if (i & 1) d[i >> 1] |= h << 16;
      else d[i >> 1] = h;
    }
    uint4* dst = reinterpret_cast<uint4*>(Ad + r * lda + 32 * w);
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[i] = make_uint4(d[4 * i], d[4 * i + 1], d[4 * i + 2], d[4 * i + 3]);
  }
}

template <int PREC>
__device__ __forceinline__ void vals_tile(const uint16_t* __restrict__ Ad, int64_t M, int64_t K,
                                          int64_t lda, const uint16_t* __restrict__ B, int64_t ldb,
                                          int64_t N, float* __restrict__ C, int64_t ldc,
                                          float alpha, float beta, int64_t tile,
                                          int64_t row_blocks) {
  __shared__ __attribute__((aligned(1024))) uint8_t sA[NBUF][A_STAGE];
  __shared__ __attribute__((aligned(1024))) uint8_t sB[NBUF][B_STAGE];
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t m0 = rb * VBM, n0 = cb * VBN;
  const int wm = wave >> 1, wn = wave & 1;
  const int j = lane & 15, q = lane >> 4;
  const int64_t nst = (K + VKS - 1) / VKS;

  // per-lane copy sources: A row / chunk, B row / column (stage moves the K offset)
  const uint16_t* asrc[NA_W];
#pragma unroll
  for (int t = 0; t < NA_W; ++t) {
    const int g = wave * NA_W + t;     // 1 KiB = 8 A rows of the stage
    const int r = 8 * g + lane / 8;
    const int slot = lane & 7;
    const int c = slot ^ ((r >> 1) & 7);
    int64_t row = m0 + r;
    if (row > M - 1) row = M - 1;
    asrc[t] = Ad + row * lda + 8 * c;
  }
  const uint16_t* bsrc[NB_W];
  int brow[NB_W];
#pragma unroll
  for (int t = 0; t < NB_W; ++t) {
    const int g = wave * NB_W + t;     // 1 KiB = 4 B rows of 256 bytes
    const int r = 4 * g + lane / 16;
    const int u = lane & 15, slot = u >> 1, half = u & 1;
    const int c = slot ^ ((r & 3) | ((r >> 1) & 4));
    int64_t col = n0 + 16 * c + 8 * half;
    if (col + 8 > N) col = N - 8;
    bsrc[t] = B + col;
    brow[t] = r;
  }
  const uint32_t sA0 = lds_u32(&sA[0][0]), sB0 = lds_u32(&sB[0][0]);

  auto issue = [&](int64_t st) {
    const int buf = (int)(st % NBUF);
#pragma unroll
    for (int t = 0; t < NA_W; ++t)
      dma16(asrc[t] + st * VKS, sA0 + (uint32_t)(buf * A_STAGE + (wave * NA_W + t) * 1024));
#pragma unroll
    for (int t = 0; t < NB_W; ++t) {
      int64_t k = st * VKS + brow[t];
      if (k > K - 1) k = K - 1;
      dma16(bsrc[t] + k * ldb, sB0 + (uint32_t)(buf * B_STAGE + (wave * NB_W + t) * 1024));
    }
  };

  f32x4 acc[VTM][VTN];
#pragma unroll
  for (int a = 0; a < VTM; ++a)
#pragma unroll
    for (int b = 0; b < VTN; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};

  issue(0);
  // transposing B reads: row 8q + i/4 (+ 4h + 32 ks), 8-byte piece i%4 of the slot
  const int i4 = j >> 2;
  const int fsw = i4 | ((q & 1) << 2);
  const uint32_t rdb = (uint32_t)((8 * q + i4) * B_ROWB + 8 * (j & 3));
  const int c0 = (VWN * wn) / 16;

  for (int64_t st = 0; st < nst; ++st) {
    const int buf = (int)(st % NBUF);
    wait_vmcnt<0>();
    __syncthreads();
    if (st + 1 < nst) issue(st + 1);
    const uint32_t bbase = sB0 + (uint32_t)(buf * B_STAGE) + rdb;
    // every fragment of the stage is requested before its first MFMA
    constexpr int VKSTEPS = VKS / 32;
    uint4 bf[VKSTEPS][VTN], af[VKSTEPS][VTM];
#pragma unroll
    for (int ks = 0; ks < VKSTEPS; ++ks) {
#pragma unroll
      for (int b = 0; b < VTN; ++b) {
        const uint32_t a0 = bbase + (uint32_t)(32 * ((c0 + b) ^ fsw)) + (uint32_t)(32 * ks * B_ROWB);
        const uint2 lo = tr_read(a0);
        const uint2 hi = tr_read(a0 + 4 * B_ROWB);
        bf[ks][b] = make_uint4(lo.x, lo.y, hi.x, hi.y);
      }
#pragma unroll
      for (int a = 0; a < VTM; ++a) {
        const int r = VWM * wm + 16 * a + j;
        const int slot = (4 * ks + q) ^ ((r >> 1) & 7);
        af[ks][a] = *reinterpret_cast<const uint4*>(&sA[buf][r * A_ROWB + 16 * slot]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int ks = 0; ks < VKSTEPS; ++ks)
#pragma unroll
      for (int a = 0; a < VTM; ++a)
#pragma unroll
        for (int b = 0; b < VTN; ++b) acc[a][b] = mfma<PREC>(af[ks][a], bf[ks][b], acc[a][b]);
  }

#pragma unroll
  for (int a = 0; a < VTM; ++a)
#pragma unroll
    for (int b = 0; b < VTN; ++b) {
      const int64_t col = n0 + VWN * wn + 16 * b + j;
      if (col >= N) continue;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int64_t row = m0 + VWM * wm + 16 * a + 4 * q + e;
        if (row < M) {
          float* cp = C + row * ldc + col;
          const float v = alpha * acc[a][b][e];
          *cp = beta != 0.0f ? fmaf(beta, *cp, v) : v;
        }
      }
    }
}

template <int PREC>
__global__ void __launch_bounds__(256, 2)
k_spmm_mfma_vals(const uint16_t* __restrict__ Ad, int64_t M, int64_t K, int64_t lda,
                 const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
                 int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
                 int64_t per_xcd) {
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= row_blocks * col_blocks) return;
  // groups of 8 row blocks, column blocks in order within a group, row blocks
  // fastest: the ~64 tiles an XCD runs together are 8 row blocks x 8 column
  // blocks, so each panel of the dense A and of B comes from HBM for 8 tiles
  // (column blocks outermost, every XCD read all of A: 272 MiB at 4096² x
  // 2048 against 80 MiB of operands)
  const int64_t gsz = 8 * col_blocks, grp = g / gsz, first = grp * 8;
  const int64_t rows_g = row_blocks - first < 8 ? row_blocks - first : 8;
  const int64_t in = g - grp * gsz;
  const int64_t rb = first + in % rows_g, cb = in / rows_g;
  vals_tile<PREC>(Ad, M, K, lda, B, ldb, N, C, ldc, alpha, beta, cb * row_blocks + rb,
                  row_blocks);
}

int64_t padded_lda(int64_t K) { return ceil_div(K, VKS) * VKS; }

}  // namespace

size_t spmm_vals_workspace(const bws_bitmap* A) {
  const int64_t lda = padded_lda(A->cols);
  Carve cv(nullptr);
  cv.take<uint16_t>((size_t)(A->rows * lda));
  cv.take<int32_t>((size_t)(A->rows * words_for(A->cols)));
  return cv.used + 256;
}

size_t spmm_planes_workspace(int64_t M, int64_t K) {
  return align_up((size_t)(M * padded_lda(K)) * 2, 256) + 256;
}

// The 16-bit MFMA product of a dense 16-bit A (rows padded to lda) already in
// the workspace.
static bws_status launch_vals_product(const uint16_t* Ad, int64_t M, int64_t K, int64_t lda,
                                      const void* B, int32_t precision, int64_t ldb, int64_t N,
                                      float* C, int64_t ldc, float alpha, float beta,
                                      hipStream_t s) {
  const int64_t row_blocks = ceil_div(M, VBM), col_blocks = ceil_div(N, VBN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const dim3 grid((unsigned)(8 * per_xcd));
  const uint16_t* Bh = static_cast<const uint16_t*>(B);
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL(k_spmm_mfma_vals<2>, grid, dim3(256), 0, s, Ad, M, K, lda, Bh, ldb, N, C,
                       ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  else
    hipLaunchKernelGGL(k_spmm_mfma_vals<1>, grid, dim3(256), 0, s, Ad, M, K, lda, Bh, ldb, N, C,
                       ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

// int8 bit-planes on MFMA in one pass over B: the planes folded into a dense
// 16-bit A (exact), then one product.  The caller checked precision, B
// (dma_ok) and the workspace size.
bws_status spmm_mfma_planes(const uint32_t* P, int64_t M, int64_t K, int64_t ldw, const void* B,
                            int32_t precision, int64_t ldb, int64_t N, float* C, int64_t ldc,
                            float alpha, float beta, void* workspace, hipStream_t s) {
  const int64_t lda = padded_lda(K);
  uint16_t* Ad = static_cast<uint16_t*>(workspace);
  int64_t g = ceil_div(M * (lda / 32), 256);
  if (g > 65536) g = 65536;
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL(k_expand_planes<2>, dim3((unsigned)g), dim3(256), 0, s, P, M, K, ldw, Ad,
                       lda);
  else
    hipLaunchKernelGGL(k_expand_planes<1>, dim3((unsigned)g), dim3(256), 0, s, P, M, K, ldw, Ad,
                       lda);
  BWS_TRY(launch_check());
  return launch_vals_product(Ad, M, K, lda, B, precision, ldb, N, C, ldc, alpha, beta, s);
}

// Value-carrying A on MFMA; the caller checked precision (BF16 / FP16), a
// 16-bit B that dma_ok() accepts, and the workspace size.
bws_status spmm_mfma_vals(const bws_bitmap* A, const void* B, int32_t precision, int64_t ldb,
                          int64_t N, float* C, int64_t ldc, float alpha, float beta,
                          void* workspace, hipStream_t s) {
  const int64_t M = A->rows, K = A->cols;
  const int64_t lda = padded_lda(K);
  Carve cv(workspace);
  uint16_t* Ad = cv.take<uint16_t>((size_t)(M * lda));
  int32_t* woff = cv.take<int32_t>((size_t)(M * words_for(K)));
  BWS_TRY(launch_word_offsets(A->words, M, K, A->ldw, woff, s));
  int64_t g = ceil_div(M * (lda / 32), 256);
  if (g > 65536) g = 65536;
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL(k_expand_values<2>, dim3((unsigned)g), dim3(256), 0, s, A->words, M, K,
                       A->ldw, A->row_ptr, woff, A->values, Ad, lda);
  else
    hipLaunchKernelGGL(k_expand_values<1>, dim3((unsigned)g), dim3(256), 0, s, A->words, M, K,
                       A->ldw, A->row_ptr, woff, A->values, Ad, lda);
  BWS_TRY(launch_check());
  return launch_vals_product(Ad, M, K, lda, B, precision, ldb, N, C, ldc, alpha, beta, s);
}

}  // namespace bws
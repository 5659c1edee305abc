// spmm_mfma_big.hip — the direct-to-LDS 16-bit MFMA SpMM with a 256 x 256
// workgroup tile: 4 waves of 128 x 128 (256 FP32 accumulators each, held in
// AGPRs: this file is compiled with the AGPR form of v_mfma), one wave per
// SIMD.  Each K step a wave reads 8 B fragments for 64 MFMAs, half the LDS
// traffic per MFMA of the 64 x 128 wave tile.  There is no empty-stage skip:
// a branch around the MFMAs would make the compiler move the accumulators
// between register files, so this tile is for masks whose stages all have
// bits (it is chosen by plan or by the autotuner).  Grouped launches do not
// use it: there the per-problem lookup pushes the wave past its register
// budget, so they run a 256 x 256 request as 128 x 256.
#include "mfma4_tile.hpp"

namespace bws {
namespace {

template <int PREC>
__global__ void __launch_bounds__(256, 1)
k_spmm_mfma4_big(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
                 int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
                 int64_t per_xcd) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  mfma4_tile<PREC, 128, 128, 2, 64, false>(W, M, K, ldw, B, ldb, N, C, ldc, alpha, beta, tile,
                                           row_blocks);
}

}  // namespace

// 16-bit B that dma_ok() accepts; PREC 1 = BF16, 2 = FP16.
bws_status launch_mfma4_big(int32_t precision, const bws_bitmap* A, const void* B, int64_t ldb,
                            int64_t N, float* C, int64_t ldc, float alpha, float beta,
                            hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, 256), col_blocks = ceil_div(N, 256);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  const dim3 grid((unsigned)(8 * per_xcd));
  const uint16_t* Bh = static_cast<const uint16_t*>(B);
  if (precision == BWS_PREC_FP16)
    hipLaunchKernelGGL(k_spmm_mfma4_big<2>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                       A->ldw, Bh, ldb, N, C, ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  else
    hipLaunchKernelGGL(k_spmm_mfma4_big<1>, grid, dim3(256), 0, s, A->words, A->rows, A->cols,
                       A->ldw, Bh, ldb, N, C, ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

}  // namespace bws

// spmm.hip — C-ABI dispatch of the grouped tile GEMM family: argument
// validation on the host, kernel selection, and the small helper kernels.
#include "bws_internal.hpp"

namespace bws {

bws_status spmm_valu(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                     float* C, int64_t ldc, float alpha, float beta, const ValuCfg& cfg,
                     hipStream_t s);

bool valu_cfg_valid(const ValuCfg& c) {
  static const int32_t ok[][3] = {{256, 8, 64}, {256, 16, 64}, {256, 16, 32}, {128, 8, 64},
                                  {128, 8, 32},  {64, 8, 32},   {128, 16, 32}, {128, 16, 64},
                                  {64, 16, 32}};
  for (const auto& v : ok)
    if (c.rows == v[0] && c.waves == v[1] && c.kstage == v[2]) return true;
  return false;
}

// Default vector-add tile for an M x N output: the tallest row block that
// still gives ~192+ workgroups (one per CU: the B stages take 128 KiB of LDS),
// always 16 waves.  Measured on 4096 x 4096 x {1024, 2048, 4096}, 2048² x 4096,
// 8192 x 4096 x 1024 and 16384 x 4096 x 512 with 5% clustered and 10% iid
// masks: 16 waves beat 8 waves on the same rows by 25-45% (four waves per
// SIMD hide the LDS and bit-scan latency of the per-row chains).  Value masks
// stop at 8 rows per wave: with 16 rows the per-row values of two stages
// exceed the 128-VGPR budget of a 1024-lane workgroup and spill, and 16 x 8
// measured 1-3% faster than that on 4096³ (2%, 10%), 8192 x 4096 x 2048 and
// 16384 x 4096 x 1024 value masks.
ValuCfg default_valu_cfg(int64_t M, int64_t N, bool vals) {
  ValuCfg c;
  c.waves = 16;
  const int64_t tiles256 = ceil_div(M, 256) * ceil_div(N, 256);
  if (tiles256 >= 192 && !vals) return c;  // 16 waves x 16 rows
  const int64_t tiles128 = ceil_div(M, 128) * ceil_div(N, 256);
  if (tiles128 >= 192) {
    c.rows = 128;
    return c;
  }
  c.rows = 64;
  c.kstage = 32;
  return c;
}

// The plan's vector-add tile, or the default where the plan leaves it 0.
ValuCfg plan_valu_cfg(const bws_plan* plan, int64_t M, int64_t N, bool vals) {
  ValuCfg c = default_valu_cfg(M, N, vals);
  if (plan && plan->rows_per_wg) {
    c.rows = plan->rows_per_wg;
    c.waves = plan->waves_per_wg ? plan->waves_per_wg : 16;
    c.kstage = plan->k_stage ? plan->k_stage : (c.rows == 64 ? 32 : 64);
  }
  return c;
}
bws_status spmm_mfma(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                     float* C, int64_t ldc, float alpha, float beta, int32_t precision,
                     const bws_plan* plan, hipStream_t s);
size_t spmm_vals_workspace(const bws_bitmap* A);
size_t spmm_split_workspace(int64_t K, int64_t N);
size_t spmm_round_workspace(int64_t K, int64_t N);
bws_status spmm_mfma_round(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N, float* C,
                           int64_t ldc, float alpha, float beta, int32_t precision,
                           const bws_plan* plan, void* workspace, hipStream_t s);
bws_status spmm_mfma_split(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N, float* C,
                           int64_t ldc, float alpha, float beta, const bws_plan* plan,
                           void* workspace, hipStream_t s);
bws_status spmm_mfma_vals(const bws_bitmap* A, const void* B, int32_t precision, int64_t ldb,
                          int64_t N, float* C, int64_t ldc, float alpha, float beta,
                          void* workspace, hipStream_t s);
bws_status spmm_groups_launch(const bws_bitmap* A, const int32_t* group_id, const float* gv,
                              int32_t T, int32_t G, const float* B, int64_t ldb, int64_t N,
                              float* C, int64_t ldc, hipStream_t s);
bool sddmm_f32_tile_ok(const float* dC, int64_t lddc, const float* B, int64_t ldb, int64_t N);
bws_status launch_sddmm_f32_tile(const bws_bitmap* A, const int32_t* woff, const float* dC,
                                 int64_t lddc, const float* B, int64_t ldb, int64_t N, float* dA,
                                 uint8_t* dense_tile, hipStream_t s);
bws_status sddmm_f32_launch(const bws_bitmap* A, const float* dC, int64_t lddc, const float* B,
                            int64_t ldb, int64_t N, float* dA, const uint8_t* dense_tile,
                            hipStream_t s);
bws_status sddmm_mfma_launch(const bws_bitmap* A, const void* dC, const void* B, int32_t dtype,
                             int64_t lddc, int64_t ldb, int64_t N, float* dA, int32_t precision,
                             void* ws, hipStream_t s);
size_t sddmm_mfma_workspace(const bws_bitmap* A);
size_t sddmm_woff_bytes(const bws_bitmap* A);

namespace {

__global__ void __launch_bounds__(256)
k_scale(float* __restrict__ C, int64_t M, int64_t N, int64_t ldc, float beta) {
  const int64_t n = M * N;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * 256) {
    const int64_t r = e / N, c = e - r * N;
    float* p = C + r * ldc + c;
    *p = beta != 0.0f ? beta * *p : 0.0f;
  }
}

bool dtype_ok(int32_t d) { return d == BWS_DTYPE_F32 || d == BWS_DTYPE_BF16 || d == BWS_DTYPE_F16; }
bool prec_ok(int32_t p) { return p >= BWS_PREC_FP32 && p <= BWS_PREC_BF16X2; }

// precision/b_dtype combinations accepted by bws_spmm
bool combo_ok(int32_t precision, int32_t b_dtype) {
  switch (precision) {
    case BWS_PREC_FP32: return true;  // 16-bit inputs widen exactly
    case BWS_PREC_BF16: return b_dtype == BWS_DTYPE_BF16 || b_dtype == BWS_DTYPE_F32;
    case BWS_PREC_FP16: return b_dtype == BWS_DTYPE_F16 || b_dtype == BWS_DTYPE_F32;
    case BWS_PREC_BF16X2: return b_dtype == BWS_DTYPE_F32;
    default: return false;
  }
}

size_t esize(int32_t d) { return d == BWS_DTYPE_F32 ? 4 : 2; }

// The streamed FP32 kernel: 16-byte rows of B and C, and enough 256 x 256
// tiles to occupy the chip (one workgroup per CU).
bool stream_ok(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, const float* C,
               int64_t ldc) {
  if (N < 4 || N % 4 != 0 || ldb % 4 != 0 || ldc % 4 != 0) return false;
  if ((reinterpret_cast<uintptr_t>(B) & 15u) || (reinterpret_cast<uintptr_t>(C) & 15u)) return false;
  return ceil_div(A->rows, kStreamTileRows) * ceil_div(N, kStreamTileCols) >= 192;
}

}  // namespace

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s) {
  const int64_t M = A->rows, K = A->cols;
  if (M == 0 || N == 0) return BWS_OK;
  if (K == 0) {
    int64_t g = ceil_div(M * N, 256);
    if (g > 65536) g = 65536;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)g), dim3(256), 0, s, C, M, N, ldc, beta);
    return launch_check();
  }
  int32_t kernel = plan ? plan->kernel : 0;
  if (precision == BWS_PREC_FP32) {
    if (kernel == 2) return BWS_EUNSUPPORTED;
    kernel = 1;
  } else if (kernel == 0) {
    // MFMA for 0/1 masks; value-carrying masks take the FP32 vector path,
    // which is at least as accurate as the requested 16-bit precision.
    kernel = A->values ? 1 : 2;
  }
  if (kernel == 2) {
    if (A->values) return BWS_EUNSUPPORTED;
    return spmm_mfma(A, B, b_dtype, ldb, N, C, ldc, alpha, beta, precision, plan, s);
  }
  const ValuCfg cfg = plan_valu_cfg(plan, M, N, A->values != nullptr);
  return spmm_valu(A, B, b_dtype, ldb, N, C, ldc, alpha, beta, cfg, s);
}

}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_spmm_workspace_size(const bws_bitmap* A, int64_t N, int32_t precision,
                               const bws_plan* plan) {
  (void)N;
  if (!A || A->rows < 0 || A->cols < 0 || N < 0) return 0;
  if (precision == BWS_PREC_FP32 && !A->values && N % 4 == 0 && plan && plan->kernel == 1 &&
      plan->rows_per_wg == 0 && plan->k_stage == kSubsetKStage)
    return subset_stream_bytes(A->rows, A->cols);
  if (precision == BWS_PREC_FP32 && !A->values && N % 4 == 0 && (!plan || plan->rows_per_wg == 0))
    return mask_stream_bytes(A->rows, A->cols);
  if (plan && plan->kernel == 1) return 0;
  // the density split: the row-block flags and room to round an FP32 B
  if (!A->values && plan && plan->mfma_min_bits > 0 && N % 8 == 0 &&
      (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16))
    return split_ws_bytes(A->rows) + spmm_round_workspace(A->cols, N);
  if (!A->values) {
    if (N % 8 != 0) return 0;
    if (precision == BWS_PREC_BF16X2) return spmm_split_workspace(A->cols, N);
    // BF16 / FP16: room to round an FP32 B once (unused for 16-bit B)
    if (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16)
      return spmm_round_workspace(A->cols, N);
    return 0;
  }
  if (precision != BWS_PREC_BF16 && precision != BWS_PREC_FP16) return 0;
  return spmm_vals_workspace(A);
}

bws_status bws_spmm(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                    float* C, int64_t ldc, float alpha, float beta, int32_t precision,
                    const bws_plan* plan, void* workspace, size_t workspace_bytes,
                    bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (N < 0 || !dtype_ok(b_dtype) || !prec_ok(precision)) return BWS_EINVAL;
  if (!combo_ok(precision, b_dtype)) return BWS_EUNSUPPORTED;
  const bool subset = plan && plan->kernel == 1 && plan->rows_per_wg == 0 &&
                      plan->k_stage == kSubsetKStage;
  if (plan && (plan->kernel < 0 || plan->kernel > 2 || plan->rows_per_wg < 0 ||
               plan->mfma_min_bits < 0 ||
               (plan->kernel != 2 && plan->rows_per_wg != 0 &&
                !valu_cfg_valid(plan_valu_cfg(plan, A->rows, N, A->values != nullptr))) ||
               (plan->kernel == 2 && !mfma_tile_valid(plan->rows_per_wg, plan->cols_per_wg))))
    return BWS_EINVAL;
  if (A->rows == 0 || N == 0) return BWS_OK;
  if (!C || ldc < N) return BWS_EINVAL;
  if (A->cols > 0 && (!B || ldb < N)) return BWS_EINVAL;
  if ((reinterpret_cast<uintptr_t>(C) & 3u) ||
      (reinterpret_cast<uintptr_t>(B) & (esize(b_dtype) - 1)))
    return BWS_EINVAL;
  // FP32 on a 0/1 A, the plan asking for the subset-sum kernel (dense masks)
  // and its mask stream in the workspace
  if (subset && precision == BWS_PREC_FP32 && !A->values && A->cols > 0 &&
      b_dtype == BWS_DTYPE_F32 && N >= 4 && N % 4 == 0 && ldb % 4 == 0 && ldc % 4 == 0 &&
      reinterpret_cast<uintptr_t>(B) % 16 == 0 && reinterpret_cast<uintptr_t>(C) % 16 == 0 &&
      ws_ok(workspace, workspace_bytes, subset_stream_bytes(A->rows, A->cols)))
    return launch_spmm_subset(A, static_cast<uint32_t*>(workspace), static_cast<const float*>(B),
                              ldb, N, C, ldc, alpha, beta, stream);
  // FP32 on a 0/1 A with the mask stream in the workspace: the streamed
  // kernel (row masks walked on the scalar unit), for problems that fill the
  // chip with 256 x 256 tiles
  if (precision == BWS_PREC_FP32 && !A->values && A->cols > 0 && b_dtype == BWS_DTYPE_F32 &&
      (!plan || plan->rows_per_wg == 0) && stream_ok(A, B, ldb, N, C, ldc) &&
      ws_ok(workspace, workspace_bytes, mask_stream_bytes(A->rows, A->cols))) {
    uint32_t* S = static_cast<uint32_t*>(workspace);
    BWS_TRY(launch_mask_stream(A, S, stream));
    return launch_spmm_stream(S, nullptr, A->rows, A->cols, static_cast<const float*>(B), ldb, N,
                              C, ldc, alpha, beta, stream);
  }
  // A value-carrying A on MFMA needs the dense 16-bit copy of A in the
  // workspace; without it (or for B layouts the direct-to-LDS kernel cannot
  // take) the FP32 vector path runs, which is at least as accurate.
  if (A->values && A->cols > 0 && (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16) &&
      (!plan || plan->kernel != 1) && b_dtype != BWS_DTYPE_F32 &&
      dma_ok(precision, b_dtype, B, ldb, N) && workspace &&
      workspace_bytes >= spmm_vals_workspace(A) &&
      (reinterpret_cast<uintptr_t>(workspace) & 255u) == 0)
    return spmm_mfma_vals(A, B, precision, ldb, N, C, ldc, alpha, beta, workspace, stream);
  // BF16 / FP16 on a 0/1 A with the plan's density split (mfma_min_bits > 0):
  // the 128-row blocks with fewer than mfma_min_bits set bits per 64 K columns
  // take the vector kernel, the others the direct-to-LDS MFMA kernel (MFMA
  // only where the block is dense enough to be a real contraction); the block
  // flags, and an FP32 B rounded once, live in the workspace.
  if (!A->values && A->cols > 0 && (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16) &&
      plan && plan->mfma_min_bits > 0 && plan->kernel != 1 && N % 8 == 0) {
    const size_t fb = split_ws_bytes(A->rows);
    const bool b16 = b_dtype != BWS_DTYPE_F32;
    if (ws_ok(workspace, workspace_bytes, fb + (b16 ? 0 : spmm_round_workspace(A->cols, N)))) {
      const void* Bh = B;
      int64_t ldbh = ldb;
      int32_t bd = b_dtype;
      if (!b16) {
        uint16_t* r = reinterpret_cast<uint16_t*>(static_cast<uint8_t*>(workspace) + fb);
        BWS_TRY(launch_round_b16(static_cast<const float*>(B), A->cols, N, ldb, precision, r,
                                 stream));
        Bh = r;
        ldbh = N;
        bd = precision == BWS_PREC_FP16 ? BWS_DTYPE_F16 : BWS_DTYPE_BF16;
      }
      if (dma_ok(precision, bd, Bh, ldbh, N)) {
        uint8_t* flags = nullptr;
        int32_t* list = nullptr;
        BWS_TRY(launch_split_flags(A, plan->mfma_min_bits, workspace, &flags, &list, stream));
        BWS_TRY(spmm_mfma_split16(A, Bh, precision, ldbh, N, C, ldc, alpha, beta, list, stream));
        return spmm_valu_split16(A, Bh, bd, ldbh, N, C, ldc, alpha, beta, flags, stream);
      }
    }
  }
  // BF16 / FP16 with FP32 B on a 0/1 A: with room for the rounded B, round it
  // once and run the direct-to-LDS kernel; otherwise round in registers.
  if (!A->values && A->cols > 0 && (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16) &&
      b_dtype == BWS_DTYPE_F32 && (!plan || plan->kernel != 1) && N % 8 == 0 &&
      ws_ok(workspace, workspace_bytes, spmm_round_workspace(A->cols, N)))
    return spmm_mfma_round(A, static_cast<const float*>(B), ldb, N, C, ldc, alpha, beta, precision,
                           plan, workspace, stream);
  // Split-BF16 on a 0/1 A: with room for B's two BF16 planes, split B once and
  // run the direct-to-LDS BF16 kernel twice; otherwise split in registers.
  if (!A->values && A->cols > 0 && precision == BWS_PREC_BF16X2 && (!plan || plan->kernel != 1) &&
      N % 8 == 0 && ws_ok(workspace, workspace_bytes, spmm_split_workspace(A->cols, N)))
    return spmm_mfma_split(A, static_cast<const float*>(B), ldb, N, C, ldc, alpha, beta, plan,
                           workspace, stream);
  return spmm_dispatch(A, B, b_dtype, ldb, N, C, ldc, alpha, beta, precision, plan, stream);
}

bws_status bws_spmm_tA(const bws_bitmap* AT, const void* dC, int32_t dc_dtype, int64_t lddc,
                       int64_t N, float* dB, int64_t lddb, float alpha, float beta,
                       int32_t precision, const bws_plan* plan, void* workspace,
                       size_t workspace_bytes, bws_stream stream) {
  return bws_spmm(AT, dC, dc_dtype, lddc, N, dB, lddb, alpha, beta, precision, plan, workspace,
                  workspace_bytes, stream);
}

bws_status bws_spmm_groups(const bws_bitmap* A, const int32_t* group_id, const float* group_vals,
                           int32_t tile_rows, int32_t max_groups, const float* B, int64_t ldb,
                           int64_t N, float* C, int64_t ldc, bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (N < 0 || tile_rows < 1 || tile_rows > 65536 || max_groups < 1 || max_groups > 32)
    return BWS_EINVAL;
  if (A->rows == 0 || N == 0) return BWS_OK;
  if (!C || ldc < N) return BWS_EINVAL;
  if (A->cols > 0 && (!B || ldb < N || !group_id || !group_vals)) return BWS_EINVAL;
  return spmm_groups_launch(A, group_id, group_vals, tile_rows, max_groups, B, ldb, N, C, ldc,
                            stream);
}

size_t bws_sddmm_workspace_size(const bws_bitmap* A) {
  if (!A || A->rows < 0 || A->cols < 0) return 0;
  return sddmm_mfma_workspace(A);
}

bws_status bws_sddmm(const bws_bitmap* A, const void* dC, const void* B, int32_t dtype,
                     int64_t lddc, int64_t ldb, int64_t N, float* dA_values, int32_t precision,
                     void* workspace, size_t workspace_bytes, bws_stream stream) {
  BWS_TRY(check_bitmap(A, true));
  if (N < 0 || !dtype_ok(dtype) || !prec_ok(precision)) return BWS_EINVAL;
  if (precision == BWS_PREC_FP32 && dtype != BWS_DTYPE_F32) return BWS_EUNSUPPORTED;
  if (!combo_ok(precision, dtype)) return BWS_EUNSUPPORTED;
  if (A->rows == 0 || A->cols == 0) return BWS_OK;
  if (!dA_values) return BWS_EINVAL;
  if (N > 0 && (!dC || !B || lddc < N || ldb < N)) return BWS_EINVAL;
  if (precision == BWS_PREC_FP32) {
    const float* dc32 = static_cast<const float*>(dC);
    const float* b32 = static_cast<const float*>(B);
    // 64 x 64 tiles with both operands staged in LDS for the tiles with at
    // least 64 set bits, when the workspace holds A's per-word offsets and the
    // tile flags and the rows allow 16-byte copies; then the wave-per-row
    // kernel takes the bits of the other tiles
    if (sddmm_f32_tile_ok(dc32, lddc, b32, ldb, N) &&
        ws_ok(workspace, workspace_bytes, sddmm_mfma_workspace(A))) {
      int32_t* woff = static_cast<int32_t*>(workspace);
      uint8_t* dense = static_cast<uint8_t*>(workspace) + sddmm_woff_bytes(A);
      BWS_TRY(launch_word_offsets(A->words, A->rows, A->cols, A->ldw, woff, stream));
      BWS_TRY(launch_sddmm_f32_tile(A, woff, dc32, lddc, b32, ldb, N, dA_values, dense, stream));
      return sddmm_f32_launch(A, dc32, lddc, b32, ldb, N, dA_values, dense, stream);
    }
    return sddmm_f32_launch(A, dc32, lddc, b32, ldb, N, dA_values, nullptr, stream);
  }
  if (!ws_ok(workspace, workspace_bytes, sddmm_mfma_workspace(A))) return BWS_EWORKSPACE;
  return sddmm_mfma_launch(A, dC, B, dtype, lddc, ldb, N, dA_values, precision, workspace, stream);
}

}  // extern "C"

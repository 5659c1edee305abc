// symbolic.hip — the symbolic phase of pre_spgemm (boolean product of two
// bitmaps), the numeric sparse x sparse product onto that pattern, and the
// int8 bit-plane SpMM.
#include "bws_internal.hpp"

#include <algorithm>

namespace bws {

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s);
size_t spmm_planes_workspace(int64_t M, int64_t K);
bws_status spmm_mfma_planes(const uint32_t* P, int64_t M, int64_t K, int64_t ldw, const void* B,
                            int32_t precision, int64_t ldb, int64_t N, float* C, int64_t ldc,
                            float alpha, float beta, void* workspace, hipStream_t s);

// Workspace of the one-pass FP32 bit-plane product: the planes' OR, its row
// pointers and word offsets, and room for one value per element.
static size_t planes_fp32_workspace(int64_t M, int64_t K) {
  const int64_t wpr = words_for(K);
  Carve cv(nullptr);
  cv.take<uint32_t>((size_t)(M * wpr));
  cv.take<int64_t>((size_t)M + 1);
  cv.take<int32_t>((size_t)(M * wpr));
  cv.take<float>((size_t)(M * K));
  cv.take<uint8_t>(scan_workspace_bytes(M));
  return cv.used + 256;
}

namespace {

// WC[r, w] = OR of WB[k, w] over the set bits k of row r of WA.  One wave per
// row; lanes own output words; the row's bits are walked wave-uniformly.
__global__ void __launch_bounds__(256)
k_symbolic_or(const uint32_t* __restrict__ WA, int64_t M, int64_t K, int64_t ldwa,
              const uint32_t* __restrict__ WB, int64_t N, int64_t ldwb, uint32_t* __restrict__ WC,
              int64_t ldwc) {
  const int lane = lane_id();
  const int64_t wpa = words_for(K), wpn = words_for(N);
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < M;
       r += (int64_t)gridDim.x * 4) {
    for (int64_t w0 = 0; w0 < ldwc; w0 += 64) {
      const int64_t w = w0 + lane;
      uint32_t acc = 0;
      for (int64_t x = 0; x < wpa; ++x) {
        uint32_t m = WA[r * ldwa + x] & valid_bits(x, K);
        while (m) {
          const int i = __clz((int)m);
          m &= ~wbit(i);
          if (w < wpn) acc |= WB[(32 * x + i) * ldwb + w] & valid_bits(w, N);
        }
      }
      if (w < ldwc) WC[r * ldwc + w] = acc;
    }
  }
}

// OR of the 8 int8 bit-planes: the nonzero pattern of the int8 matrix.
__global__ void __launch_bounds__(256)
k_planes_or(const uint32_t* __restrict__ P, int64_t M, int64_t K, int64_t ldw,
            uint32_t* __restrict__ Wor, int64_t wpr) {
  const int64_t n = M * wpr;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / wpr, w = t - r * wpr;
    uint32_t m = 0u;
#pragma unroll
    for (int b = 0; b < 8; ++b) m |= P[((int64_t)b * M + r) * ldw + w];
    Wor[t] = m & valid_bits(w, K);
  }
}

// The int8 values of the pattern's set bits, in (row, column) order:
// Σ_b w_b·bit_b with w = [1, 2, ..., 64, -128] (the plane recombination of
// source/src/main.cu:889-912).  A thread per word.
__global__ void __launch_bounds__(256)
k_planes_values(const uint32_t* __restrict__ P, int64_t M, int64_t ldw,
                const uint32_t* __restrict__ Wor, int64_t wpr, const int64_t* __restrict__ rp,
                const int32_t* __restrict__ woff, float* __restrict__ values) {
  const int64_t n = M * wpr;
  for (int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (int64_t)gridDim.x * 256) {
    const int64_t r = t / wpr, w = t - r * wpr;
    uint32_t m = Wor[t];
    if (!m) continue;
    uint32_t pl[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) pl[b] = P[((int64_t)b * M + r) * ldw + w];
    int64_t e = rp[r] + woff[t];
    while (m) {
      const int i = __clz((int)m);
      m &= ~wbit(i);
      int v = 0;
#pragma unroll
      for (int b = 0; b < 8; ++b) v += (pl[b] & wbit(i)) ? (b == 7 ? -128 : (1 << b)) : 0;
      values[e++] = (float)v;
    }
  }
}

constexpr int kSlab = 4096;       // output columns per LDS accumulator slab
constexpr int kSpWaves = 4;       // output rows (waves) per workgroup

// Numeric product onto a given output pattern: C.values[rpC[r] + j] = Σ_k
// A[r, k]·B[k, c_j] for the j-th set bit c_j of row r of WC.  A wave per
// output row and 4096-column slab: the slab's accumulators live in LDS
// (16 KiB per wave); A's set bits are walked in ascending k (wave-uniform),
// and for each, lane l adds a·B[k, c] for the set bits c of B's word l of the
// slab (distinct columns per lane, so no two lanes touch one accumulator);
// then lane l writes the accumulators of the set bits of C's word l.  FP32
// FMA in ascending k for every output entry.
__global__ void __launch_bounds__(kSpWaves * 64)
k_spgemm(const uint32_t* __restrict__ WA, int64_t M, int64_t K, int64_t ldwa,
         const int64_t* __restrict__ rpA, const float* __restrict__ vA,
         const uint32_t* __restrict__ WB, int64_t N, int64_t ldwb, const int64_t* __restrict__ rpB,
         const float* __restrict__ vB, const int32_t* __restrict__ woffB,
         const uint32_t* __restrict__ WC, int64_t ldwc, const int64_t* __restrict__ rpC,
         const int32_t* __restrict__ woffC, float alpha, float* __restrict__ vC) {
  __shared__ float acc[kSpWaves][kSlab];
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * kSpWaves + wave;
  if (r >= M) return;
  const int64_t wpa = words_for(K), wpn = words_for(N);
  float* a_s = acc[wave];
  for (int64_t c0 = 0; c0 < N; c0 += kSlab) {
    const int64_t w0 = c0 / 32, wn = std::min<int64_t>(wpn - w0, kSlab / 32);
    for (int i = lane; i < kSlab; i += 64) a_s[i] = 0.0f;
    __builtin_amdgcn_wave_barrier();
    int64_t ea = vA ? rpA[r] : 0;  // value index of A's next set bit
    for (int64_t x = 0; x < wpa; ++x) {
      uint32_t m = WA[r * ldwa + x] & valid_bits(x, K);
      while (m) {
        const int i = __clz((int)m);
        m &= ~wbit(i);
        const int64_t k = 32 * x + i;
        const float a = vA ? vA[ea] : 1.0f;
        ++ea;
        for (int64_t q = lane; q < wn; q += 64) {
          const int64_t w = w0 + q;
          uint32_t b = WB[k * ldwb + w] & valid_bits(w, N);
          int64_t e = vB ? rpB[k] + woffB[k * wpn + w] : 0;
          while (b) {
            const int j = __clz((int)b);
            b &= ~wbit(j);
            a_s[32 * q + j] = fmaf(a, vB ? vB[e] : 1.0f, a_s[32 * q + j]);
            ++e;
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    const int64_t wpc = wpn;
    for (int64_t q = lane; q < wn; q += 64) {
      const int64_t w = w0 + q;
      uint32_t cm = WC[r * ldwc + w] & valid_bits(w, N);
      int64_t e = rpC[r] + woffC[r * wpc + w];
      while (cm) {
        const int j = __clz((int)cm);
        cm &= ~wbit(j);
        vC[e++] = alpha * a_s[32 * q + j];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

bws_status bws_symbolic_or(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa,
                           const uint32_t* WB, int64_t N, int64_t ldwb, uint32_t* WC,
                           int64_t ldwc, bws_stream stream) {
  if (!dims_ok(M, K) || N < 0 || !ldw_ok(K, ldwa) || !ldw_ok(N, ldwb) || !ldw_ok(N, ldwc))
    return BWS_EINVAL;
  if (M == 0) return BWS_OK;
  if (!WC || (K > 0 && (!WA || !WB))) return BWS_EINVAL;
  int64_t g = ceil_div(M, 4);
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(k_symbolic_or, dim3((unsigned)g), dim3(256), 0, stream, WA, M, K, ldwa, WB, N,
                     ldwb, WC, ldwc);
  return launch_check();
}

size_t bws_spgemm_workspace_size(int64_t M, int64_t K, int64_t N) {
  if (M < 0 || K < 0 || N < 0) return 0;
  const int64_t wpn = words_for(N);
  return align_up((size_t)(K * wpn) * 4, 256) + align_up((size_t)(M * wpn) * 4, 256) + 256;
}

bws_status bws_spgemm(const bws_bitmap* A, const bws_bitmap* B, const bws_bitmap* C,
                      float alpha, float* C_values, void* workspace, size_t workspace_bytes,
                      bws_stream stream) {
  BWS_TRY(check_bitmap(A, A && A->values != nullptr));
  BWS_TRY(check_bitmap(B, B && B->values != nullptr));
  BWS_TRY(check_bitmap(C, true));
  if (A->cols != B->rows || C->rows != A->rows || C->cols != B->cols) return BWS_EINVAL;
  if (C->rows == 0 || C->cols == 0) return BWS_OK;
  if (!C_values) return BWS_EINVAL;
  const int64_t M = A->rows, K = A->cols, N = B->cols;
  if (!ws_ok(workspace, workspace_bytes, bws_spgemm_workspace_size(M, K, N)))
    return BWS_EWORKSPACE;
  Carve cv(workspace);
  int32_t* woffB = cv.take<int32_t>((size_t)(K * words_for(N)));
  int32_t* woffC = cv.take<int32_t>((size_t)(M * words_for(N)));
  BWS_TRY(launch_word_offsets(B->words, K, N, B->ldw, woffB, stream));
  BWS_TRY(launch_word_offsets(C->words, M, N, C->ldw, woffC, stream));
  const int64_t g = ceil_div(M, kSpWaves);
  if (g > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_spgemm, dim3((unsigned)g), dim3(kSpWaves * 64), 0, stream, A->words, M, K,
                     A->ldw, A->row_ptr, A->values, B->words, N, B->ldw, B->row_ptr, B->values,
                     woffB, C->words, C->ldw, C->row_ptr, woffC, alpha, C_values);
  return launch_check();
}

size_t bws_spmm_bitplanes_workspace_size(int64_t M, int64_t K, int32_t precision) {
  if (M < 0 || K < 0) return 0;
  if (precision == BWS_PREC_FP32) return planes_fp32_workspace(M, K);
  if (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16) return spmm_planes_workspace(M, K);
  return 0;
}

bws_status bws_spmm_bitplanes_i8(const uint32_t* P, int64_t M, int64_t K, int64_t ldw,
                                 const void* B, int32_t b_dtype, int64_t ldb, int64_t N, float* C,
                                 int64_t ldc, float alpha, float beta, int32_t precision,
                                 const bws_plan* plan, void* workspace, size_t workspace_bytes,
                                 bws_stream stream) {
  static const float w8[8] = {1.f, 2.f, 4.f, 8.f, 16.f, 32.f, 64.f, -128.f};
  if (!dims_ok(M, K) || N < 0 || !ldw_ok(K, ldw)) return BWS_EINVAL;
  if (M == 0 || N == 0) return BWS_OK;
  if (K > 0 && !P) return BWS_EINVAL;
  const bool one_pass_plan = !plan || plan->kernel == 0;
  // one pass over B, FP32: the planes as a value-carrying bitmap of int8 values
  if (K > 0 && one_pass_plan && precision == BWS_PREC_FP32 && b_dtype == BWS_DTYPE_F32 && B && C &&
      ws_ok(workspace, workspace_bytes, planes_fp32_workspace(M, K))) {
    const int64_t wpr = words_for(K);
    Carve cv(workspace);
    uint32_t* Wor = cv.take<uint32_t>((size_t)(M * wpr));
    int64_t* rp = cv.take<int64_t>((size_t)M + 1);
    int32_t* woff = cv.take<int32_t>((size_t)(M * wpr));
    float* vals = cv.take<float>((size_t)(M * K));
    void* scan_ws = cv.take<uint8_t>(scan_workspace_bytes(M));
    int64_t g = std::min<int64_t>(ceil_div(M * wpr, 256), 65536);
    hipLaunchKernelGGL(k_planes_or, dim3((unsigned)g), dim3(256), 0, stream, P, M, K, ldw, Wor, wpr);
    BWS_TRY(launch_check());
    BWS_TRY(launch_row_ptr(Wor, M, K, wpr, rp, scan_ws, stream));
    BWS_TRY(launch_word_offsets(Wor, M, K, wpr, woff, stream));
    hipLaunchKernelGGL(k_planes_values, dim3((unsigned)g), dim3(256), 0, stream, P, M, ldw, Wor,
                       wpr, rp, woff, vals);
    BWS_TRY(launch_check());
    bws_bitmap Av{M, K, wpr, Wor, rp, vals};
    return bws_spmm(&Av, B, b_dtype, ldb, N, C, ldc, alpha, beta, BWS_PREC_FP32, plan, nullptr, 0,
                    stream);
  }
  // one pass over B, BF16 / FP16: the planes folded into one exact 16-bit A
  if (K > 0 && one_pass_plan && (precision == BWS_PREC_BF16 || precision == BWS_PREC_FP16) &&
      b_dtype != BWS_DTYPE_F32 && C && dma_ok(precision, b_dtype, B, ldb, N) &&
      ldc >= N && (reinterpret_cast<uintptr_t>(C) & 3u) == 0 &&
      ws_ok(workspace, workspace_bytes, spmm_planes_workspace(M, K)))
    return spmm_mfma_planes(P, M, K, ldw, B, precision, ldb, N, C, ldc, alpha, beta, workspace,
                            stream);
  // otherwise eight weighted 0/1 products
  for (int b = 0; b < 8; ++b) {
    bws_bitmap A{M, K, ldw, P ? P + (int64_t)b * M * ldw : nullptr, nullptr, nullptr};
    BWS_TRY(bws_spmm(&A, B, b_dtype, ldb, N, C, ldc, alpha * w8[b], b == 0 ? beta : 1.0f,
                     precision, plan, nullptr, 0, stream));
  }
  return BWS_OK;
}

}  // extern "C"

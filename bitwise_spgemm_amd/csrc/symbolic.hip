// symbolic.hip — the symbolic phase of pre_spgemm (boolean product of two
// bitmaps) and the int8 bit-plane SpMM built on the 0/1 kernel.
#include "bws_internal.hpp"

namespace bws {

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s);

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

bws_status bws_spmm_bitplanes_i8(const uint32_t* P, int64_t M, int64_t K, int64_t ldw,
                                 const void* B, int32_t b_dtype, int64_t ldb, int64_t N, float* C,
                                 int64_t ldc, float alpha, float beta, int32_t precision,
                                 const bws_plan* plan, bws_stream stream) {
  static const float w8[8] = {1.f, 2.f, 4.f, 8.f, 16.f, 32.f, 64.f, -128.f};
  for (int b = 0; b < 8; ++b) {
    bws_bitmap A{M, K, ldw, P ? P + (int64_t)b * M * ldw : nullptr, nullptr, nullptr};
    BWS_TRY(bws_spmm(&A, B, b_dtype, ldb, N, C, ldc, alpha * w8[b], b == 0 ? beta : 1.0f,
                     precision, plan, nullptr, 0, stream));
  }
  return BWS_OK;
}

}  // extern "C"

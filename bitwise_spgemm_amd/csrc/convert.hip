// convert.hip — COO / CSR <-> packed bitmap converters for gfx950.
//
// These replace the reference's serial tiling pass csr2tiledcsr (<<<1,1>>>,
// source/src/main.cu:499-588) and its use of cuSPARSE DenseToSparse
// (main.cu:630-690) for getting a sparse matrix into and out of the bitmap
// format: every step here is a data-parallel pass.
//
// Into a bitmap (entries in any order):
//  1. clear the words; 2. one thread per entry sets its bit with atomicOr
//     (an entry outside the matrix, or a second entry at a position already
//     set, is counted in *n_invalid and otherwise ignored);
//  3. row pointers (popcount prefix) and per-row word offsets;
//  4. one thread per entry stores its value at
//     row_ptr[r] + word offset + popcount of the earlier bits of its word,
//     i.e. in the bitmap's (row, column) order.
// Out of a bitmap: row pointers and word offsets, then one thread per word
// writes the column index (and value) of each set bit.
#include "bws_internal.hpp"

namespace bws {
namespace {

constexpr int kBlock = 256;

unsigned grid_of(int64_t n) {
  int64_t g = ceil_div(n, kBlock);
  if (g < 1) g = 1;
  if (g > (int64_t)1 << 20) g = (int64_t)1 << 20;
  return (unsigned)g;
}

__global__ void __launch_bounds__(kBlock)
k_clear(uint32_t* __restrict__ W, int64_t n, unsigned long long* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) W[i] = 0u;
  if (bad && blockIdx.x == 0 && threadIdx.x == 0) *bad = 0ull;
}

// bit of column c in its word (MSB-first)
__device__ __forceinline__ uint32_t col_bit(int64_t c) { return 0x80000000u >> (c & 31); }

__device__ __forceinline__ void set_entry(int64_t r, int64_t c, int64_t M, int64_t K,
                                          uint32_t* __restrict__ W, int64_t ldw,
                                          unsigned long long* __restrict__ bad) {
  if (r < 0 || r >= M || c < 0 || c >= K) {
    atomicAdd(bad, 1ull);
    return;
  }
  const uint32_t b = col_bit(c);
  const uint32_t old = atomicOr(&W[r * ldw + (c >> 5)], b);
  if (old & b) atomicAdd(bad, 1ull);
}

__global__ void __launch_bounds__(kBlock)
k_coo_set(const int32_t* __restrict__ ri, const int32_t* __restrict__ ci, int64_t nnz, int64_t M,
          int64_t K, uint32_t* __restrict__ W, int64_t ldw, unsigned long long* __restrict__ bad) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nnz; i += stride)
    set_entry(ri[i], ci[i], M, K, W, ldw, bad);
}

// The entries row r of a CSR matrix may read: [rp[r], rp[r+1]) clamped to
// [0, nnz) and to non-decreasing bounds.  A row pointer that is negative,
// decreasing or past nnz (and rp[0] != 0) is an invalid input; the kernel
// that sets the bits counts one per such row into *bad.
__device__ __forceinline__ void csr_range(const int64_t* __restrict__ rp, int64_t r, int64_t nnz,
                                          int64_t& lo, int64_t& hi, bool& ok) {
  const int64_t a = rp[r], b = rp[r + 1];
  ok = a >= 0 && b >= a && b <= nnz && (r != 0 || a == 0);
  lo = a < 0 ? 0 : (a > nnz ? nnz : a);
  hi = b < lo ? lo : (b > nnz ? nnz : b);
}

// CSR: one wave per row.
__global__ void __launch_bounds__(kBlock)
k_csr_set(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci, int64_t nnz, int64_t M,
          int64_t K, uint32_t* __restrict__ W, int64_t ldw, unsigned long long* __restrict__ bad) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t r = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); r < M; r += waves) {
    int64_t lo, hi;
    bool ok;
    csr_range(rp, r, nnz, lo, hi, ok);
    if (!ok && lane == 0) atomicAdd(bad, 1ull);
    for (int64_t i = lo + lane; i < hi; i += 64) set_entry(r, ci[i], M, K, W, ldw, bad);
  }
}

// Value of entry (r, c) to its bitmap-order position in out [cap] (invalid
// CSR row pointers can give more bits than entries: positions past cap are
// not written).
__device__ __forceinline__ void place(int64_t r, int64_t c, float v, int64_t M, int64_t K,
                                      const uint32_t* __restrict__ W, int64_t ldw, int64_t wpr,
                                      const int64_t* __restrict__ row_ptr,
                                      const int32_t* __restrict__ woff, float* __restrict__ out,
                                      int64_t cap) {
  if (r < 0 || r >= M || c < 0 || c >= K) return;
  const int64_t w = c >> 5;
  const uint32_t before = W[r * ldw + w] & ~(0xFFFFFFFFu >> (c & 31));
  const int64_t pos = row_ptr[r] + woff[r * wpr + w] + __popc(before);
  if (pos < cap) out[pos] = v;
}

__global__ void __launch_bounds__(kBlock)
k_coo_place(const int32_t* __restrict__ ri, const int32_t* __restrict__ ci,
            const float* __restrict__ vals, int64_t nnz, int64_t M, int64_t K,
            const uint32_t* __restrict__ W, int64_t ldw, int64_t wpr,
            const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
            float* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < nnz; i += stride)
    place(ri[i], ci[i], vals[i], M, K, W, ldw, wpr, row_ptr, woff, out, nnz);
}

__global__ void __launch_bounds__(kBlock)
k_csr_place(const int64_t* __restrict__ rp, const int32_t* __restrict__ ci,
            const float* __restrict__ vals, int64_t nnz, int64_t M, int64_t K,
            const uint32_t* __restrict__ W, int64_t ldw, int64_t wpr,
            const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
            float* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const int64_t waves = (int64_t)gridDim.x * (kBlock / 64);
  for (int64_t r = (int64_t)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6); r < M; r += waves) {
    int64_t lo, hi;
    bool ok;
    csr_range(rp, r, nnz, lo, hi, ok);
    for (int64_t i = lo + lane; i < hi; i += 64)
      place(r, ci[i], vals[i], M, K, W, ldw, wpr, row_ptr, woff, out, nnz);
  }
}

// One thread per word: column indices (and values) of its set bits.
__global__ void __launch_bounds__(kBlock)
k_to_csr(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw, int64_t wpr,
         const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
         int32_t* __restrict__ col_idx, const float* __restrict__ vin, float* __restrict__ vout) {
  const int64_t n = M * wpr;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  for (int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x; t < n; t += stride) {
    const int64_t r = t / wpr, w = t - r * wpr;
    uint32_t m = W[r * ldw + w] & valid_bits(w, K);
    int64_t pos = row_ptr[r] + woff[t];
    while (m) {
      const int i = __clz((int)m);
      m &= ~(0x80000000u >> i);
      if (col_idx) col_idx[pos] = (int32_t)(32 * w + i);
      if (vout) vout[pos] = vin ? vin[pos] : 1.0f;
      ++pos;
    }
  }
}

struct ConvWs {
  int32_t* woff;
  int64_t* row_ptr;
  void* scan;
  size_t total;
};

ConvWs carve_conv(void* base, int64_t rows, int64_t cols) {
  ConvWs c{};
  Carve cv(base);
  c.woff = cv.take<int32_t>((size_t)(rows * words_for(cols)));
  c.row_ptr = cv.take<int64_t>((size_t)(rows + 1));
  c.scan = cv.take<char>(scan_workspace_bytes(rows));
  c.total = cv.used;
  return c;
}

bool shape_ok(int64_t rows, int64_t cols, int64_t ldw) {
  return rows >= 0 && cols >= 0 && rows <= INT32_MAX && cols <= INT32_MAX &&
         ldw >= words_for(cols) && (rows == 0 || ldw >= 1);
}

// Shared tail of both "into a bitmap" entry points: row pointers, word
// offsets, then the values in bitmap order.
template <class Place>
bws_status finish_into(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                       int64_t* row_ptr, float* values_out, const ConvWs& c, Place&& place_values,
                       hipStream_t s) {
  int64_t* rp = row_ptr ? row_ptr : c.row_ptr;
  if (row_ptr || values_out) BWS_TRY(launch_row_ptr(W, rows, cols, ldw, rp, c.scan, s));
  if (values_out && cols > 0) {
    BWS_TRY(launch_word_offsets(W, rows, cols, ldw, c.woff, s));
    place_values(rp, c.woff);
    BWS_TRY(launch_check());
  }
  return BWS_OK;
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_convert_workspace_size(int64_t rows, int64_t cols) {
  if (rows < 0 || cols < 0) return 0;
  return carve_conv(nullptr, rows, cols).total + 256;
}

bws_status bws_coo_to_bitmap(const int32_t* row_idx, const int32_t* col_idx, const float* values,
                             int64_t nnz, int64_t rows, int64_t cols, uint32_t* W, int64_t ldw,
                             int64_t* row_ptr, float* values_out, unsigned long long* n_invalid,
                             void* workspace, size_t workspace_bytes, bws_stream stream) {
  if (!shape_ok(rows, cols, ldw) || nnz < 0 || !n_invalid) return BWS_EINVAL;
  if (nnz > 0 && (!row_idx || !col_idx)) return BWS_EINVAL;
  if (values_out && nnz > 0 && !values) return BWS_EINVAL;
  if (rows > 0 && !W) return BWS_EINVAL;
  if (!ws_ok(workspace, workspace_bytes, bws_convert_workspace_size(rows, cols)))
    return BWS_EWORKSPACE;
  hipStream_t s = stream;
  const ConvWs c = carve_conv(workspace, rows, cols);
  hipLaunchKernelGGL(k_clear, dim3(grid_of(rows * ldw)), dim3(kBlock), 0, s, W, rows * ldw,
                     n_invalid);
  BWS_TRY(launch_check());
  if (nnz > 0) {
    hipLaunchKernelGGL(k_coo_set, dim3(grid_of(nnz)), dim3(kBlock), 0, s, row_idx, col_idx, nnz,
                       rows, cols, W, ldw, n_invalid);
    BWS_TRY(launch_check());
  }
  const int64_t wpr = words_for(cols);
  return finish_into(W, rows, cols, ldw, row_ptr, values_out, c,
                     [&](const int64_t* rp, const int32_t* woff) {
                       if (nnz > 0)
                         hipLaunchKernelGGL(k_coo_place, dim3(grid_of(nnz)), dim3(kBlock), 0, s,
                                            row_idx, col_idx, values, nnz, rows, cols, W, ldw, wpr,
                                            rp, woff, values_out);
                     },
                     s);
}

bws_status bws_csr_to_bitmap(const int64_t* csr_row_ptr, const int32_t* col_idx,
                             const float* values, int64_t nnz, int64_t rows, int64_t cols,
                             uint32_t* W, int64_t ldw, int64_t* row_ptr, float* values_out,
                             unsigned long long* n_invalid, void* workspace,
                             size_t workspace_bytes, bws_stream stream) {
  if (!shape_ok(rows, cols, ldw) || !n_invalid || nnz < 0) return BWS_EINVAL;
  if (rows > 0 && (!W || !csr_row_ptr)) return BWS_EINVAL;
  if (nnz > 0 && !col_idx) return BWS_EINVAL;
  if (values_out && nnz > 0 && !values) return BWS_EINVAL;
  if (!ws_ok(workspace, workspace_bytes, bws_convert_workspace_size(rows, cols)))
    return BWS_EWORKSPACE;
  hipStream_t s = stream;
  const ConvWs c = carve_conv(workspace, rows, cols);
  hipLaunchKernelGGL(k_clear, dim3(grid_of(rows * ldw)), dim3(kBlock), 0, s, W, rows * ldw,
                     n_invalid);
  BWS_TRY(launch_check());
  if (rows > 0) {
    hipLaunchKernelGGL(k_csr_set, dim3(grid_of(rows * 64)), dim3(kBlock), 0, s, csr_row_ptr,
                       col_idx, nnz, rows, cols, W, ldw, n_invalid);
    BWS_TRY(launch_check());
  }
  const int64_t wpr = words_for(cols);
  return finish_into(W, rows, cols, ldw, row_ptr, values_out, c,
                     [&](const int64_t* rp, const int32_t* woff) {
                       if (rows > 0 && values)
                         hipLaunchKernelGGL(k_csr_place, dim3(grid_of(rows * 64)), dim3(kBlock),
                                            0, s, csr_row_ptr, col_idx, values, nnz, rows, cols,
                                            W, ldw, wpr, rp, woff, values_out);
                     },
                     s);
}

bws_status bws_bitmap_to_csr(const bws_bitmap* A, int64_t* row_ptr, int32_t* col_idx,
                             float* values_out, void* workspace, size_t workspace_bytes,
                             bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (!row_ptr) return BWS_EINVAL;
  if (A->rows > INT32_MAX || A->cols > INT32_MAX) return BWS_ECAPACITY;
  if (!ws_ok(workspace, workspace_bytes, bws_convert_workspace_size(A->rows, A->cols)))
    return BWS_EWORKSPACE;
  hipStream_t s = stream;
  const ConvWs c = carve_conv(workspace, A->rows, A->cols);
  BWS_TRY(launch_row_ptr(A->words, A->rows, A->cols, A->ldw, row_ptr, c.scan, s));
  const int64_t wpr = words_for(A->cols);
  if (A->rows == 0 || wpr == 0 || (!col_idx && !values_out)) return BWS_OK;
  BWS_TRY(launch_word_offsets(A->words, A->rows, A->cols, A->ldw, c.woff, s));
  hipLaunchKernelGGL(k_to_csr, dim3(grid_of(A->rows * wpr)), dim3(kBlock), 0, s, A->words,
                     A->rows, A->cols, A->ldw, wpr, row_ptr, c.woff, col_idx, A->values,
                     values_out);
  return launch_check();
}

}  // extern "C"

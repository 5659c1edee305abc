// bws_internal.hpp — shared host/device helpers of libbws (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "bws.h"

namespace bws {

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void set_last_hip_error(int32_t e);

inline bws_status hip_check(hipError_t e) {
  if (e != hipSuccess) {
    set_last_hip_error((int32_t)e);
    return BWS_EHIP;
  }
  return BWS_OK;
}

// Status of the launches just enqueued (launch-configuration errors only).
inline bws_status launch_check() { return hip_check(hipGetLastError()); }

__host__ __device__ inline int64_t words_for(int64_t cols) { return (cols + 31) / 32; }
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

#define BWS_TRY(expr)                    \
  do {                                   \
    bws_status _s = (expr);              \
    if (_s != BWS_OK) return _s;         \
  } while (0)

// Validation helpers (host only, no HIP calls).
inline bool dims_ok(int64_t rows, int64_t cols) { return rows >= 0 && cols >= 0; }
inline bool ldw_ok(int64_t cols, int64_t ldw) { return ldw >= words_for(cols) && ldw >= 1; }
inline bool ld_ok(int64_t cols, int64_t ld) { return ld >= cols && ld >= 1; }
// Caller workspaces: present, large enough and 256-byte aligned (the
// alignment hipMalloc gives; sub-buffers are carved at 256-byte offsets).
inline bool ws_ok(const void* ws, size_t have, size_t need) {
  return ws && have >= need && (reinterpret_cast<uintptr_t>(ws) & 255u) == 0;
}
bws_status check_bitmap(const bws_bitmap* A, bool need_row_ptr);

// Maximum grid dimension used for 1-D launches (blocks are looped over).
constexpr int64_t kMaxGrid = 1 << 30;

// Vector-add path tile configuration (rows of C per workgroup, waves per
// workgroup, K rows per LDS stage).
struct ValuCfg {
  int32_t rows = 256;
  int32_t waves = 8;
  int32_t kstage = 64;
};
bool valu_cfg_valid(const ValuCfg& c);

// MFMA workgroup tile: bm x bn of C (128 or 256 each); legacy selects the
// 64 x 64-per-wave kernel with table-lookup A fragments.
struct MfmaTile {
  int32_t bm = 128;
  int32_t bn = 128;
  int32_t legacy = 0;
  int32_t dma = 0;   // direct-to-LDS staging of 16-bit B
  int32_t nbuf = 2;  // LDS stages of the direct-to-LDS kernel
  int32_t ks = 64;   // K rows per stage of the direct-to-LDS kernel (64 or 128)
  int32_t sparse = 0; // the sparse-MFMA kernel (spmm_smfmac.hip)
};
bool dma_ok(int32_t precision, int32_t b_dtype, const void* B, int64_t ldb, int64_t N);
MfmaTile mfma_tile_for(const bws_plan* plan, int32_t precision, int32_t b_dtype, int64_t M,
                       int64_t N);
bool mfma_tile_valid(int bm, int bn);

// One problem of a grouped launch, as the kernels read it from the workspace.
struct BatchEntry {
  const uint32_t* W;
  const int64_t* row_ptr;
  const float* values;
  const void* B;
  float* C;
  int64_t M, K, ldw, ldb, N, ldc;
  int64_t row_blocks, col_blocks;
};

// the sparse-MFMA kernel for a 0/1 A and 16-bit B (spmm_smfmac.hip)
bws_status launch_smfmac(int32_t precision, const bws_bitmap* A, const void* B, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         hipStream_t s);
// the 256 x 256-tile direct-to-LDS kernel (spmm_mfma_big.hip)
bws_status launch_mfma4_big(int32_t precision, const bws_bitmap* A, const void* B, int64_t ldb,
                            int64_t N, float* C, int64_t ldc, float alpha, float beta,
                            hipStream_t s);
// the 256 x 256-block 16-bit MFMA SDDMM (sddmm_mfma_big.hip)
bws_status launch_sddmm_big(int32_t precision, const bws_bitmap* A, const int32_t* woff,
                            const void* dC, int64_t lddc, const void* B, int64_t ldb, int64_t N,
                            float* dA, hipStream_t s);
int64_t sddmm_big_tiles(int64_t rows, int64_t cols);

// The streamed FP32 0/1 product (spmm_stream.hip): the mask stream of a
// plain bitmap or of a row-clustering plan, and the product over it.
size_t mask_stream_bytes(int64_t M, int64_t K);
size_t cluster_stream_bytes(int64_t M, int64_t K);
bws_status launch_mask_stream(const bws_bitmap* A, uint32_t* S, hipStream_t s);
bws_status launch_cluster_stream(const bws_bitmap* A, const int32_t* slot_row,
                                 const uint32_t* chunk_mask, uint32_t* S, hipStream_t s);
bws_status launch_spmm_stream(const uint32_t* S, const int32_t* slot_row, int64_t M, int64_t K,
                              const float* B, int64_t ldb, int64_t N, float* C, int64_t ldc,
                              float alpha, float beta, hipStream_t s);
// The subset-sum product for dense masks (16-row sub-stages, 4-row groups):
// its 32-row mask stream and the product (mask stream pass included).
size_t subset_stream_bytes(int64_t M, int64_t K);
bws_status launch_spmm_subset(const bws_bitmap* A, uint32_t* S, const float* B, int64_t ldb,
                              int64_t N, float* C, int64_t ldc, float alpha, float beta,
                              hipStream_t s);
// K rows per sub-stage of the subset-sum product (its plan: kernel 1, k_stage 16)
constexpr int32_t kSubsetKStage = 16;
// Rows and columns of C per workgroup of the streamed product.
constexpr int64_t kStreamTileRows = 256;
constexpr int64_t kStreamTileCols = 256;

// MFMA density split (plan mfma_min_bits > 0, 16-bit precision, 0/1 A): the
// rows in blocks of kSplitRows; the flags pass (per-block bit counts, a byte
// per block: 1 = MFMA, and the list of flagged blocks, list[0] = their count)
// in a workspace of split_ws_bytes, and the two sides of the product.
constexpr int64_t kSplitRows = 128;
size_t split_ws_bytes(int64_t M);
bws_status launch_split_flags(const bws_bitmap* A, int32_t min_bits, void* ws, uint8_t** flags,
                              int32_t** list, hipStream_t s);
bws_status spmm_valu_split16(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                             int64_t N, float* C, int64_t ldc, float alpha, float beta,
                             const uint8_t* mfma_rows, hipStream_t s);
bws_status launch_round_b16(const float* B, int64_t K, int64_t N, int64_t ldb, int32_t precision,
                            uint16_t* Bh, hipStream_t s);
bws_status spmm_mfma_split16(const bws_bitmap* A, const void* B, int32_t precision, int64_t ldb,
                             int64_t N, float* C, int64_t ldc, float alpha, float beta,
                             const int32_t* list, hipStream_t s);

// Device scratch carving: a linear allocator over the caller's workspace.
struct Carve {
  char* base;
  size_t used = 0;
  explicit Carve(void* p) : base(static_cast<char*>(p)) {}
  template <class T>
  T* take(size_t n) {
    used = align_up(used, 256);
    T* p = reinterpret_cast<T*>(base ? base + used : nullptr);
    used += n * sizeof(T);
    return p;
  }
};

// Internal launches shared between translation units.
bws_status launch_word_offsets(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                               int32_t* woff, hipStream_t s);
bws_status launch_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                          int64_t* row_ptr, void* ws, hipStream_t s);
bws_status launch_exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws,
                                     hipStream_t s);
size_t scan_workspace_bytes(int64_t n);

// ---------------------------------------------------------------------------
// device side
// ---------------------------------------------------------------------------
constexpr int kWave = 64;

__device__ __forceinline__ uint32_t wbit(int i) { return 0x80000000u >> i; }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Bits of word w that correspond to columns < cols (MSB-first).
__device__ __forceinline__ uint32_t valid_bits(int64_t w, int64_t cols) {
  const int64_t rem = cols - 32 * w;
  return rem >= 32 ? 0xFFFFFFFFu : (rem <= 0 ? 0u : (0xFFFFFFFFu << (32 - rem)));
}

// Number of set bits of `m` at columns < i (bits 31 .. 32-i).
__device__ __forceinline__ int rank_before(uint32_t m, int i) {
  return i == 0 ? 0 : __popc(m & ~(0xFFFFFFFFu >> i));
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wave.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T u = __shfl_up(v, o, kWave);
    if (l >= o) v += u;
  }
  return v;
}

__device__ __forceinline__ uint32_t lo32(uint64_t x) { return (uint32_t)x; }
__device__ __forceinline__ uint32_t hi32(uint64_t x) { return (uint32_t)(x >> 32); }

}  // namespace bws

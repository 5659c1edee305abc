// layout.hip — bitmap layout transforms (pack, bit-planes, popcount prefix,
// value packing, bit-matrix transpose) for gfx950.
//
// The pack is a wave64 ballot: one wave reads 64 consecutive elements of a row
// per load (256 contiguous bytes, coalesced), __ballot turns "!= 0" into a
// 64-bit LSB-first lane mask, and __brev of each 32-bit half gives the
// reference's MSB-first word (bit 31-i <-> column i, source/src/main.cu:619).
#include "bws_internal.hpp"

namespace bws {
namespace {

constexpr int kBlock = 256;  // 4 waves

// ---------------------------------------------------------------------------
// pack: one wave per (row, 256-column chunk); 4 loads in flight per lane.
// chunks_per_row = ceil(ldw / 8) so padding words up to ldw are written as 0.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_pack_bitmap_f32(const float* __restrict__ D, int64_t rows, int64_t cols, int64_t ld,
                  uint32_t* __restrict__ W, int64_t ldw, int64_t chunks_per_row,
                  int64_t total_waves) {
  const int lane = lane_id();
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < total_waves;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t r = gw / chunks_per_row;
    const int64_t c0 = (gw - r * chunks_per_row) * 256;
    const float* row = D + r * ld;
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (c0 < cols) {  // wave-uniform
      // unconditional loads (columns past the row read its last one and are
      // zeroed afterwards), so that all four are in flight together
      float x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t c = c0 + 64 * u + lane;
        x[u] = row[c < cols ? c : cols - 1];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (c0 + 64 * u + lane < cols) v[u] = x[u];
    }
    uint64_t b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b[u] = __ballot(v[u] != 0.0f);
    if (lane < 8) {
      const int64_t w = c0 / 32 + lane;
      if (w < ldw) {
        const int u = lane >> 1;
        const uint64_t bb = u == 0 ? b[0] : u == 1 ? b[1] : u == 2 ? b[2] : b[3];
        const uint32_t half = (lane & 1) ? hi32(bb) : lo32(bb);
        W[r * ldw + w] = __brev(half);
      }
    }
  }
}

// Vector pack for 16-byte aligned rows (ld and cols multiples of 4): one
// wave per (row, 1024-column span), four 16-byte loads in flight per lane
// (4 KiB per wave; the scalar kernel above keeps 1 KiB in flight, which
// leaves HBM half idle).  Lane l's float4 u holds columns 256u + 4l .. +3, so
// word 8u + i of the span takes bit m of byte i of the ballot of component c
// for column 4m + c: the four bytes are interleaved (bit m -> bit 4m + c) and
// bit-reversed into the MSB-first word by lane 8u + i.
__device__ __forceinline__ uint32_t spread4(uint32_t x) {  // bit m -> bit 4m, m < 8
  x = (x | (x << 12)) & 0x000F000Fu;
  x = (x | (x << 6)) & 0x03030303u;
  x = (x | (x << 3)) & 0x11111111u;
  return x;
}

__global__ void __launch_bounds__(kBlock)
k_pack_bitmap_f32_vec(const float* __restrict__ D, int64_t rows, int64_t cols, int64_t ld,
                      uint32_t* __restrict__ W, int64_t ldw, int64_t spans_per_row,
                      int64_t total_waves) {
  const int lane = lane_id();
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < total_waves;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t r = gw / spans_per_row;
    const int64_t c0 = (gw - r * spans_per_row) * 1024;
    const float* row = D + r * ld;
    float4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (c0 < cols) {  // wave-uniform; then cols >= 4
      // unconditional loads (columns past the row read its last four and are
      // zeroed afterwards): a load under a per-lane branch waits where the
      // branch joins, which would leave one load in flight instead of four
      float4 x[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int64_t c = c0 + 256 * u + 4 * lane;
        x[u] = *reinterpret_cast<const float4*>(row + (c < cols ? c : cols - 4));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (c0 + 256 * u + 4 * lane < cols) v[u] = x[u];
    }
    uint64_t b[4][4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      b[u][0] = __ballot(v[u].x != 0.0f);
      b[u][1] = __ballot(v[u].y != 0.0f);
      b[u][2] = __ballot(v[u].z != 0.0f);
      b[u][3] = __ballot(v[u].w != 0.0f);
    }
    if (lane < 32) {
      const int64_t w = c0 / 32 + lane;
      if (w < ldw) {
        const int u = lane >> 3, i = lane & 7;
        uint32_t word = 0u;
#pragma unroll
        for (int uu = 0; uu < 4; ++uu)
          if (uu == u) {
#pragma unroll
            for (int c = 0; c < 4; ++c)
              word |= spread4((uint32_t)(b[uu][c] >> (8 * i)) & 0xFFu) << c;
          }
        W[r * ldw + w] = __brev(word);
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_unpack_bitmap_f32(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t ldw,
                    float* __restrict__ D, int64_t ld) {
  const int64_t n = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / cols, c = e - r * cols;
    D[r * ld + c] = (W[r * ldw + (c >> 5)] & wbit((int)(c & 31))) ? 1.0f : 0.0f;
  }
}

// layout_trans / uint8_layout_trans (backup/layout_trans.cu:75-97)
__global__ void __launch_bounds__(kBlock)
k_unpack_bitplanes_f32(const uint32_t* __restrict__ src, int64_t n, int bit_width,
                       int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t v = src[i];
    for (int b = 0; b < bit_width; ++b) dst[(int64_t)b * n + i] = (int32_t)((v >> b) & 1u);
  }
}

__global__ void __launch_bounds__(kBlock)
k_unpack_bitplanes_u8(const uint8_t* __restrict__ src, int64_t n, int32_t* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock) {
    const uint32_t v = src[i];
#pragma unroll
    for (int b = 0; b < 8; ++b) dst[(int64_t)b * n + i] = (int32_t)((v >> b) & 1u);
  }
}

// Packed int8 planes: one wave per (row, 64-column chunk), 8 ballots.
__global__ void __launch_bounds__(kBlock)
k_pack_bitplanes_i8(const int8_t* __restrict__ D, int64_t rows, int64_t cols, int64_t ld,
                    uint32_t* __restrict__ P, int64_t ldw, int64_t chunks_per_row,
                    int64_t total_waves) {
  const int lane = lane_id();
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < total_waves;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t r = gw / chunks_per_row;
    const int64_t c0 = (gw - r * chunks_per_row) * 64;
    const int64_t c = c0 + lane;
    const uint32_t v = c < cols ? (uint32_t)(uint8_t)D[r * ld + c] : 0u;
    uint32_t mine = 0;  // lane 2b+h keeps half h of plane b
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const uint64_t bb = __ballot((v >> b) & 1u);
      if ((lane >> 1) == b) mine = (lane & 1) ? hi32(bb) : lo32(bb);
    }
    if (lane < 16) {
      const int b = lane >> 1;
      const int64_t w = c0 / 32 + (lane & 1);
      if (w < ldw) P[(int64_t)b * rows * ldw + r * ldw + w] = __brev(mine);
    }
  }
}

// Per-plane per-row popcounts: one wave per (plane, row).
__global__ void __launch_bounds__(kBlock)
k_bitplane_row_counts(const uint32_t* __restrict__ P, int64_t rows, int64_t wpr, int64_t ldw,
                      int64_t total_waves, int32_t* __restrict__ counts) {
  const int lane = lane_id();
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < total_waves;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t b = gw / rows, r = gw - b * rows;
    const uint32_t* row = P + b * rows * ldw + r * ldw;
    int n = 0;
    for (int64_t w = lane; w < wpr; w += 64) n += __popc(row[w]);
    n = wave_sum(n);
    if (lane == 0) counts[gw] = n;
  }
}

// ---------------------------------------------------------------------------
// Exclusive scans.  Phase 1 writes block-local exclusive prefixes and block
// totals; phase 2 scans the totals in one block (looping); phase 3 adds them.
// ---------------------------------------------------------------------------
template <class T>
__device__ T block_excl_scan(T x, T* lds_wave, T* total) {
  const int lane = lane_id(), wave = threadIdx.x >> 6;
  const T incl = wave_incl_scan(x);
  if (lane == 63) lds_wave[wave] = incl;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kBlock / 64; ++i) {
    const T v = lds_wave[i];
    if (i < wave) before += v;
    all += v;
  }
  __syncthreads();
  *total = all;
  return before + incl - x;
}

// Row popcounts (one thread per row) + block-local scan.  A thread's word
// loads are independent, so a row's words are in flight together; the
// earlier layout (a wave per row, 64 rows in turn) waited for one row's
// loads before the next and took ~50 µs at 4096 x 4096.
__global__ void __launch_bounds__(kBlock)
k_row_ptr_phase1(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t wpr,
                 int64_t ldw, int64_t* __restrict__ row_ptr, int64_t* __restrict__ block_tot) {
  __shared__ int64_t wsum[kBlock / 64];
  const int64_t r = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int64_t n = 0;
  if (r < rows) {
    const uint32_t* row = W + r * ldw;
    const int64_t full = cols / 32;  // words with 32 valid bits
    int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    int64_t w = 0;
    for (; w + 4 <= full; w += 4) {
      c0 += __popc(row[w]);
      c1 += __popc(row[w + 1]);
      c2 += __popc(row[w + 2]);
      c3 += __popc(row[w + 3]);
    }
    for (; w < wpr; ++w) c0 += __popc(row[w] & valid_bits(w, cols));
    n = (int64_t)c0 + c1 + c2 + c3;
  }
  int64_t tot;
  const int64_t ex = block_excl_scan<int64_t>(n, wsum, &tot);
  if (r < rows) row_ptr[r] = ex;
  if (threadIdx.x == 0) block_tot[blockIdx.x] = tot;
}

// Single-block exclusive scan of n block totals in place; writes the grand
// total to *grand.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_scan_totals(T* __restrict__ tot, int64_t n, T* __restrict__ grand) {
  __shared__ T wsum[kBlock / 64];
  T carry = 0;
  for (int64_t base = 0; base < n; base += kBlock) {
    const int64_t i = base + threadIdx.x;
    const T x = i < n ? tot[i] : T(0);
    T all;
    const T ex = block_excl_scan<T>(x, wsum, &all);
    if (i < n) tot[i] = carry + ex;
    carry += all;
  }
  if (threadIdx.x == 0) *grand = carry;
}

template <class T>
__global__ void __launch_bounds__(kBlock)
k_add_block_offsets(T* __restrict__ out, int64_t n, const T* __restrict__ tot) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) out[i] += tot[blockIdx.x];
}

// Generic int32 phase 1 (input array instead of row popcounts).
__global__ void __launch_bounds__(kBlock)
k_scan_i32_phase1(const int32_t* __restrict__ in, int64_t n, int32_t* __restrict__ out,
                  int32_t* __restrict__ block_tot) {
  __shared__ int32_t wsum[kBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  int32_t tot;
  const int32_t ex = block_excl_scan<int32_t>(i < n ? in[i] : 0, wsum, &tot);
  if (i < n) out[i] = ex;
  if (threadIdx.x == 0) block_tot[blockIdx.x] = tot;
}

// Per-row exclusive prefix of word popcounts: woff[r*wpr + w].  One wave per row.
__global__ void __launch_bounds__(kBlock)
k_word_offsets(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t wpr,
               int64_t ldw, int32_t* __restrict__ woff) {
  const int lane = lane_id();
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < rows;
       r += (int64_t)gridDim.x * 4) {
    int32_t carry = 0;
    for (int64_t w0 = 0; w0 < wpr; w0 += 64) {
      const int64_t w = w0 + lane;
      const int32_t p = w < wpr ? __popc(W[r * ldw + w] & valid_bits(w, cols)) : 0;
      const int32_t incl = wave_incl_scan(p);
      if (w < wpr) woff[r * wpr + w] = carry + incl - p;
      carry += __shfl(incl, 63, kWave);
    }
  }
}

// values[row_ptr[r] + j] = D[r, c_j]: one wave per row, lanes own words.
__global__ void __launch_bounds__(kBlock)
k_pack_values_f32(const float* __restrict__ D, int64_t rows, int64_t cols, int64_t ld,
                  const uint32_t* __restrict__ W, int64_t wpr, int64_t ldw,
                  const int64_t* __restrict__ row_ptr, float* __restrict__ values) {
  const int lane = lane_id();
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < rows;
       r += (int64_t)gridDim.x * 4) {
    int64_t carry = row_ptr[r];
    for (int64_t w0 = 0; w0 < wpr; w0 += 64) {
      const int64_t w = w0 + lane;
      uint32_t m = w < wpr ? (W[r * ldw + w] & valid_bits(w, cols)) : 0u;
      const int32_t p = __popc(m);
      const int32_t incl = wave_incl_scan(p);
      int64_t e = carry + incl - p;
      while (m) {
        const int i = __clz(m);
        m &= ~wbit(i);
        values[e++] = D[r * ld + 32 * w + i];
      }
      carry += __shfl(incl, 63, kWave);
    }
  }
}

// Bit-matrix transpose of 32x32 blocks: each half-wave owns one block; 32
// ballots transpose it (ballot bit j = row j of the block).
__global__ void __launch_bounds__(kBlock)
k_transpose_bitmap(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t ldw,
                   uint32_t* __restrict__ WT, int64_t ldwt, int64_t wpr, int64_t nblocks) {
  const int lane = lane_id();
  const int h = lane >> 5, j = lane & 31;
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); 2 * gw < nblocks;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t blk = 2 * gw + h;
    const bool valid = blk < nblocks;
    const int64_t rb = valid ? blk / wpr : 0, w = valid ? blk - rb * wpr : 0;
    const int64_t r = 32 * rb + j;
    const uint32_t word = (valid && r < rows) ? W[r * ldw + w] : 0u;
    uint32_t out = 0;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
      const uint64_t bal = __ballot((word & wbit(c)) != 0u);
      const uint32_t mine = h ? hi32(bal) : lo32(bal);
      if (j == c) out = __brev(mine);
    }
    const int64_t col = 32 * w + j;
    if (valid && col < cols) WT[col * ldwt + rb] = out;
  }
}

// Values of Aᵀ: one thread per word of A.  A word's set bits are taken up to
// four at a time so that their (independent) reads of WT, the word offsets
// and row pointers of Aᵀ are in flight together.
// PERM: write the source index (perm_T[dst] = src) instead of the value.
template <bool PERM>
__global__ void __launch_bounds__(kBlock)
k_transpose_values(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t wpr,
                   int64_t ldw,
                   const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woffA,
                   const float* __restrict__ values, const uint32_t* __restrict__ WT,
                   int64_t wprT, int64_t ldwt, const int64_t* __restrict__ row_ptr_T,
                   const int32_t* __restrict__ woffT, float* __restrict__ values_T,
                   int64_t* __restrict__ perm_T) {
  const int64_t n = rows * wpr;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / wpr, w = e - r * wpr;
    uint32_t m = W[r * ldw + w] & valid_bits(w, cols);
    if (!m) continue;
    int64_t src = row_ptr[r] + woffA[r * wpr + w];
    const int64_t x = r >> 5;
    const int rb = (int)(r & 31);
    while (m) {
      int64_t c[4];
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (m) {
          const int i = __clz(m);
          m &= ~wbit(i);
          c[j] = 32 * w + i;
          cnt = j + 1;
        } else {
          c[j] = -1;
        }
      }
      uint32_t tw[4];
      int64_t base[4];
      float v[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j < cnt) {
          tw[j] = WT[c[j] * ldwt + x];
          base[j] = row_ptr_T[c[j]] + woffT[c[j] * wprT + x];
          if (!PERM) v[j] = values ? values[src + j] : 1.0f;
        }
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        if (j >= cnt) continue;
        const int64_t dst = base[j] + rank_before(tw[j], rb);
        if (PERM) perm_T[dst] = src + j;
        else values_T[dst] = v[j];
      }
      src += cnt;
    }
  }
}

// dst[i] = src[idx[i]].
__global__ void __launch_bounds__(kBlock)
k_gather_f32(const float* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
             float* __restrict__ dst) {
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[idx[i]];
}

inline int grid_for(int64_t items, int64_t per_block) {
  int64_t g = ceil_div(items, per_block);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

}  // namespace

// ---------------------------------------------------------------------------
// internal launches
// ---------------------------------------------------------------------------
size_t scan_workspace_bytes(int64_t n) {
  return align_up((size_t)(ceil_div(n, kBlock) + 1) * sizeof(int64_t), 256) + 256;
}

bws_status launch_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                          int64_t* row_ptr, void* ws, hipStream_t s) {
  const int64_t nb = ceil_div(rows, kBlock);
  int64_t* tot = static_cast<int64_t*>(ws);
  if (rows > 0) {
    hipLaunchKernelGGL(k_row_ptr_phase1, dim3((unsigned)nb), dim3(kBlock), 0, s, W, rows, cols,
                       words_for(cols), ldw, row_ptr, tot);
    hipLaunchKernelGGL(k_scan_totals<int64_t>, dim3(1), dim3(kBlock), 0, s, tot, nb,
                       row_ptr + rows);
    hipLaunchKernelGGL(k_add_block_offsets<int64_t>, dim3((unsigned)nb), dim3(kBlock), 0, s,
                       row_ptr, rows, tot);
  } else {
    hipLaunchKernelGGL(k_scan_totals<int64_t>, dim3(1), dim3(kBlock), 0, s, tot, (int64_t)0,
                       row_ptr);
  }
  return launch_check();
}

bws_status launch_exclusive_scan_i32(const int32_t* in, int64_t n, int32_t* out, void* ws,
                                     hipStream_t s) {
  // out has n+1 entries; out[n] = total.
  const int64_t nb = ceil_div(n, kBlock);
  int32_t* tot = static_cast<int32_t*>(ws);
  if (n > 0) {
    hipLaunchKernelGGL(k_scan_i32_phase1, dim3((unsigned)nb), dim3(kBlock), 0, s, in, n, out,
                       tot);
    hipLaunchKernelGGL(k_scan_totals<int32_t>, dim3(1), dim3(kBlock), 0, s, tot, nb, out + n);
    hipLaunchKernelGGL(k_add_block_offsets<int32_t>, dim3((unsigned)nb), dim3(kBlock), 0, s,
                       out, n, tot);
  } else {
    hipLaunchKernelGGL(k_scan_totals<int32_t>, dim3(1), dim3(kBlock), 0, s, tot, (int64_t)0,
                       out);
  }
  return launch_check();
}

bws_status launch_word_offsets(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                               int32_t* woff, hipStream_t s) {
  const int64_t wpr = words_for(cols);
  if (rows <= 0 || wpr <= 0) return BWS_OK;
  hipLaunchKernelGGL(k_word_offsets, dim3(grid_for(rows, 4)), dim3(kBlock), 0, s, W, rows, cols,
                     wpr, ldw, woff);
  return launch_check();
}

}  // namespace bws

using namespace bws;

extern "C" {

bws_status bws_pack_bitmap_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                               uint32_t* W, int64_t ldw, bws_stream stream) {
  if (!dims_ok(rows, cols) || !ld_ok(cols, ld) || !ldw_ok(cols, ldw)) return BWS_EINVAL;
  if (rows == 0) return BWS_OK;
  if (!W || (cols > 0 && !D)) return BWS_EINVAL;
  if (cols % 4 == 0 && ld % 4 == 0 && reinterpret_cast<uintptr_t>(D) % 16 == 0) {
    const int64_t spr = ceil_div(ldw, 32);
    const int64_t waves = rows * spr;
    hipLaunchKernelGGL(k_pack_bitmap_f32_vec, dim3(grid_for(waves, 4)), dim3(kBlock), 0, stream,
                       D, rows, cols, ld, W, ldw, spr, waves);
    return launch_check();
  }
  const int64_t cpr = ceil_div(ldw, 8);
  const int64_t waves = rows * cpr;
  hipLaunchKernelGGL(k_pack_bitmap_f32, dim3(grid_for(waves, 4)), dim3(kBlock), 0, stream, D,
                     rows, cols, ld, W, ldw, cpr, waves);
  return launch_check();
}

bws_status bws_unpack_bitmap_f32(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                                 float* D, int64_t ld, bws_stream stream) {
  if (!dims_ok(rows, cols) || !ld_ok(cols, ld) || !ldw_ok(cols, ldw)) return BWS_EINVAL;
  if (rows == 0 || cols == 0) return BWS_OK;
  if (!W || !D) return BWS_EINVAL;
  hipLaunchKernelGGL(k_unpack_bitmap_f32, dim3(grid_for(rows * cols, kBlock)), dim3(kBlock), 0,
                     stream, W, rows, cols, ldw, D, ld);
  return launch_check();
}

bws_status bws_unpack_bitplanes_f32(const float* src, int64_t n, int32_t bit_width, int32_t* dst,
                                    bws_stream stream) {
  if (n < 0 || bit_width < 1 || bit_width > 32) return BWS_EINVAL;
  if (n == 0) return BWS_OK;
  if (!src || !dst) return BWS_EINVAL;
  hipLaunchKernelGGL(k_unpack_bitplanes_f32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream,
                     reinterpret_cast<const uint32_t*>(src), n, (int)bit_width, dst);
  return launch_check();
}

bws_status bws_unpack_bitplanes_u8(const uint8_t* src, int64_t n, int32_t* dst,
                                   bws_stream stream) {
  if (n < 0) return BWS_EINVAL;
  if (n == 0) return BWS_OK;
  if (!src || !dst) return BWS_EINVAL;
  hipLaunchKernelGGL(k_unpack_bitplanes_u8, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream,
                     src, n, dst);
  return launch_check();
}

bws_status bws_pack_bitplanes_i8(const int8_t* D, int64_t rows, int64_t cols, int64_t ld,
                                 uint32_t* P, int64_t ldw, bws_stream stream) {
  if (!dims_ok(rows, cols) || !ld_ok(cols, ld) || !ldw_ok(cols, ldw)) return BWS_EINVAL;
  if (rows == 0) return BWS_OK;
  if (!P || (cols > 0 && !D)) return BWS_EINVAL;
  const int64_t cpr = ceil_div(ldw, 2);
  const int64_t waves = rows * cpr;
  hipLaunchKernelGGL(k_pack_bitplanes_i8, dim3(grid_for(waves, 4)), dim3(kBlock), 0, stream, D,
                     rows, cols, ld, P, ldw, cpr, waves);
  return launch_check();
}

bws_status bws_bitplane_weights(float* w8) {
  if (!w8) return BWS_EINVAL;
  for (int b = 0; b < 7; ++b) w8[b] = (float)(1 << b);
  w8[7] = -128.0f;
  return BWS_OK;
}

bws_status bws_bitplane_row_counts(const uint32_t* P, int32_t planes, int64_t rows, int64_t cols,
                                   int64_t ldw, int32_t* counts, bws_stream stream) {
  if (planes < 1 || !dims_ok(rows, cols) || !ldw_ok(cols, ldw)) return BWS_EINVAL;
  if (rows == 0) return BWS_OK;
  if (!P || !counts) return BWS_EINVAL;
  const int64_t waves = (int64_t)planes * rows;
  hipLaunchKernelGGL(k_bitplane_row_counts, dim3(grid_for(waves, 4)), dim3(kBlock), 0, stream, P,
                     rows, words_for(cols), ldw, waves, counts);
  return launch_check();
}

size_t bws_row_ptr_workspace_size(int64_t rows) {
  return scan_workspace_bytes(rows < 0 ? 0 : rows);
}

bws_status bws_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                       int64_t* row_ptr, void* workspace, size_t workspace_bytes,
                       bws_stream stream) {
  if (!dims_ok(rows, cols) || !ldw_ok(cols, ldw) || !row_ptr) return BWS_EINVAL;
  if (rows > 0 && !W) return BWS_EINVAL;
  if (!ws_ok(workspace, workspace_bytes, bws_row_ptr_workspace_size(rows))) return BWS_EWORKSPACE;
  return launch_row_ptr(W, rows, cols, ldw, row_ptr, workspace, stream);
}

bws_status bws_pack_values_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                               const uint32_t* W, int64_t ldw, const int64_t* row_ptr,
                               float* values, bws_stream stream) {
  if (!dims_ok(rows, cols) || !ld_ok(cols, ld) || !ldw_ok(cols, ldw)) return BWS_EINVAL;
  if (rows == 0 || cols == 0) return BWS_OK;
  if (!D || !W || !row_ptr || !values) return BWS_EINVAL;
  hipLaunchKernelGGL(k_pack_values_f32, dim3(grid_for(rows, 4)), dim3(kBlock), 0, stream, D,
                     rows, cols, ld, W, words_for(cols), ldw, row_ptr, values);
  return launch_check();
}

bws_status bws_transpose_bitmap(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                                uint32_t* WT, int64_t ldwt, bws_stream stream) {
  if (!dims_ok(rows, cols) || !ldw_ok(cols, ldw) || !ldw_ok(rows, ldwt)) return BWS_EINVAL;
  if (cols == 0) return BWS_OK;
  if (!WT || (rows > 0 && !W)) return BWS_EINVAL;
  const int64_t wpr = words_for(cols);
  const int64_t nblocks = ldwt * wpr;  // row blocks up to ldwt: padding words of WT become 0
  hipLaunchKernelGGL(k_transpose_bitmap, dim3(grid_for(ceil_div(nblocks, 2), 4)), dim3(kBlock), 0,
                     stream, W, rows, cols, ldw, WT, ldwt, wpr, nblocks);
  return launch_check();
}

size_t bws_transpose_workspace_size(int64_t rows, int64_t cols) {
  if (rows < 0 || cols < 0) return 0;
  const int64_t wpr = words_for(cols), wprT = words_for(rows);
  Carve c(nullptr);
  c.take<int32_t>((size_t)(rows * wpr));
  c.take<int32_t>((size_t)(cols * wprT));
  c.take<char>(scan_workspace_bytes(cols));
  return c.used + 256;
}

// Shared by bws_transpose_values and bws_transpose_perm: AT's words and row
// pointers, then (values_T or perm_T non-null) its values or their sources.
static bws_status transpose_impl(const bws_bitmap* A, uint32_t* WT, int64_t ldwt, int64_t* row_ptr_T,
                          float* values_T, int64_t* perm_T, void* workspace,
                          size_t workspace_bytes, hipStream_t stream) {
  BWS_TRY(check_bitmap(A, true));
  const int64_t rows = A->rows, cols = A->cols;
  if (!ldw_ok(rows, ldwt) || !WT || !row_ptr_T) return BWS_EINVAL;
  if (!ws_ok(workspace, workspace_bytes, bws_transpose_workspace_size(rows, cols)))
    return BWS_EWORKSPACE;
  const int64_t wpr = words_for(cols), wprT = words_for(rows);
  Carve c(workspace);
  int32_t* woffA = c.take<int32_t>((size_t)(rows * wpr));
  int32_t* woffT = c.take<int32_t>((size_t)(cols * wprT));
  void* scan_ws = c.take<char>(scan_workspace_bytes(cols));
  BWS_TRY(bws_transpose_bitmap(A->words, rows, cols, A->ldw, WT, ldwt, stream));
  BWS_TRY(launch_row_ptr(WT, cols, rows, ldwt, row_ptr_T, scan_ws, stream));
  if ((!values_T && !perm_T) || rows == 0 || cols == 0) return BWS_OK;
  BWS_TRY(launch_word_offsets(A->words, rows, cols, A->ldw, woffA, stream));
  BWS_TRY(launch_word_offsets(WT, cols, rows, ldwt, woffT, stream));
  const dim3 grid(grid_for(rows * wpr, kBlock));
  if (perm_T)
    hipLaunchKernelGGL(k_transpose_values<true>, grid, dim3(kBlock), 0, stream, A->words, rows,
                       cols, wpr, A->ldw, A->row_ptr, woffA, nullptr, WT, wprT, ldwt, row_ptr_T,
                       woffT, nullptr, perm_T);
  else
    hipLaunchKernelGGL(k_transpose_values<false>, grid, dim3(kBlock), 0, stream, A->words, rows,
                       cols, wpr, A->ldw, A->row_ptr, woffA, A->values, WT, wprT, ldwt,
                       row_ptr_T, woffT, values_T, nullptr);
  return launch_check();
}

bws_status bws_transpose_values(const bws_bitmap* A, uint32_t* WT, int64_t ldwt,
                                int64_t* row_ptr_T, float* values_T, void* workspace,
                                size_t workspace_bytes, bws_stream stream) {
  return transpose_impl(A, WT, ldwt, row_ptr_T, values_T, nullptr, workspace, workspace_bytes,
                        stream);
}

bws_status bws_transpose_perm(const bws_bitmap* A, uint32_t* WT, int64_t ldwt,
                              int64_t* row_ptr_T, int64_t* perm_T, void* workspace,
                              size_t workspace_bytes, bws_stream stream) {
  if (!perm_T) return BWS_EINVAL;
  return transpose_impl(A, WT, ldwt, row_ptr_T, nullptr, perm_T, workspace, workspace_bytes,
                        stream);
}

bws_status bws_gather_f32(const float* src, const int64_t* idx, int64_t n, float* dst,
                          bws_stream stream) {
  if (n < 0) return BWS_EINVAL;
  if (n == 0) return BWS_OK;
  if (!src || !idx || !dst) return BWS_EINVAL;
  hipLaunchKernelGGL(k_gather_f32, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, stream, src, idx,
                     n, dst);
  return launch_check();
}

}  // extern "C"

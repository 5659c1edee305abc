// group.hip — row grouping by bitmask (first-fit disjoint groups per tile of
// tile_rows rows x one 32-column word), spilled-row CSR, un-grouping.
//
// One wave (its own workgroup, so that the tiles spread over every CU and no
// two placement chains share a CU's scalar unit) owns one tile and walks its
// rows 64 at a time (lane l <-> row j0 + l).  Per chunk one vector pass finds
// the lanes that fit some group against the masks as the chunk starts; the
// other non-zero lanes must spill (group masks only grow), and the candidates
// are placed in ascending row order on the scalar unit against the current
// masks.  The result equals the serial ascending-row first fit exactly.
#include "bws_internal.hpp"

namespace bws {
namespace {

constexpr int kBlock = 256;
constexpr int kGroupPF = 32;  // 64-row chunks whose words a grouping wave loads at once

__global__ void k_zero_i64(int64_t* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = 0;
}

// A tile's group masks.  GC > 0: the group count is GC, known at compile
// time, and the masks sit in scalar registers, so the placement loop tests
// them and adds to them without moving them between lanes and scalars.
// GC == 0 (any count up to 32): lane g < G of one vector register holds mask
// g, read back with v_readlane.
template <int GC>
struct GroupMasks {
  uint32_t m[GC];
  __device__ GroupMasks() {
#pragma unroll
    for (int g = 0; g < GC; ++g) m[g] = 0u;
  }
  __device__ uint32_t get(int g) const {  // g wave-uniform
    uint32_t v = 0u;
#pragma unroll
    for (int i = 0; i < GC; ++i)
      if (i == g) v = m[i];
    return v;
  }
  // first group that word w fits (GC if none)
  __device__ int fit(uint32_t w, int) const {
    int f = GC;
#pragma unroll
    for (int g = GC - 1; g >= 0; --g)
      if ((m[g] & w) == 0u) f = g;
    return f;
  }
  // gj, mj wave-uniform; gj >= GC adds nothing
  __device__ void add(int gj, uint32_t mj, int) {
#pragma unroll
    for (int g = 0; g < GC; ++g) m[g] |= g == gj ? mj : 0u;
  }
  __device__ uint32_t lane_value(int lane) const {
    uint32_t v = 0u;
#pragma unroll
    for (int g = 0; g < GC; ++g)
      if (lane == g) v = m[g];
    return v;
  }
};

template <>
struct GroupMasks<0> {
  uint32_t v = 0u;
  __device__ uint32_t get(int g) const {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, g);
  }
  __device__ int fit(uint32_t w, int G) const {
    int f = G;
    for (int g = 0; g < G; ++g)
      if (f == G && (get(g) & w) == 0u) f = g;
    return f;
  }
  __device__ void add(int gj, uint32_t mj, int lane) {  // gj >= G adds nothing
    if (lane == gj) v |= mj;
  }
  __device__ uint32_t lane_value(int) const { return v; }
};

template <bool HAS_VALUES, int GC>
__global__ void __launch_bounds__(64)
k_group_rows(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t wpr,
             int64_t ldw,
             const int64_t* __restrict__ row_ptr, const float* __restrict__ values, int T, int G,
             int64_t tiles, int32_t* __restrict__ group_id, uint32_t* __restrict__ group_mask,
             float* __restrict__ group_vals, int32_t* __restrict__ member_row,
             int32_t* __restrict__ spill_cnt, int32_t* __restrict__ spill_nnz,
             int32_t* __restrict__ spill_table, int32_t* __restrict__ spill_rev,
             unsigned long long* __restrict__ grouped_nnz, int64_t per_xcd) {
  // the words of a group of chunks, then each row's result packed as
  // (slot + 1) << 8 | (group + 1) (group -1..31, slot -1..65535)
  __shared__ uint32_t s_word[kGroupPF * 64];
  __shared__ uint32_t s_res[kGroupPF * 64];
  const int lane = lane_id();
  // XCD-contiguous tile ranges: the tiles of neighbouring words of the same
  // rows share an XCD, so the 128-byte lines of group_id they fill together
  // are merged in one L2 before they are written
  for (int64_t b = blockIdx.x; b < 8 * per_xcd; b += gridDim.x) {
    const int64_t t = (b & 7) * per_xcd + (b >> 3);
    if (t >= tiles) continue;
    const int64_t rt = t / wpr, w = t - rt * wpr;
    const int64_t row0 = rt * T;
    const int64_t rem = rows - row0;
    const int nrows = rem < T ? (int)rem : T;
    GroupMasks<GC> gms;
    // 0/1 masks with a compile-time group count: lane i < 32 holds the member
    // row of bit i of each group, written once per tile
    constexpr bool kRegMembers = !HAS_VALUES && GC > 0;
    int32_t mem[GC > 0 ? GC : 1];
#pragma unroll
    for (int g = 0; g < (GC > 0 ? GC : 1); ++g) mem[g] = -1;
    int32_t cnt = 0;       // spilled rows so far (wave-uniform)
    int32_t snnz_lane = 0;
    // one chunk of 64 rows, in row order (the first-fit order).  The lanes
    // that fit some group as the chunk starts are the candidates; the group
    // masks only grow, so every other non-zero lane spills.  The candidates
    // are placed one by one in row order on the scalar unit (the lane's word
    // read with v_readlane, first fit against the current masks), which
    // gives the serial ascending-row first fit exactly.
    auto chunk = [&](int j0, uint32_t m, int32_t& gid, int32_t& slot) {
      gid = 0;
      slot = -1;
      const uint64_t nz = __ballot(m != 0u);
      uint64_t cand = __ballot(m != 0u && gms.fit(m, G) < G);
      uint64_t joined = 0;
      // straight-line per candidate (a taken branch costs a lone wave more
      // than the few selects that replace it)
      while (cand) {
        const int jl = (int)__builtin_ctzll(cand);
        cand &= cand - 1ull;
        const uint32_t mj = (uint32_t)__builtin_amdgcn_readlane((int)m, jl);
        const int gj = gms.fit(mj, G);
        const bool ok = gj < G;
        joined |= (uint64_t)ok << jl;
        gid = lane == jl && ok ? gj : gid;
        gms.add(gj, mj, lane);
        if constexpr (kRegMembers) {
          const bool mine = (mj & wbit(lane & 31)) != 0u;
#pragma unroll
          for (int g = 0; g < (GC > 0 ? GC : 1); ++g) mem[g] = mine && g == gj ? j0 + jl : mem[g];
        } else if (ok) {
          int64_t vbase = 0;
          if (HAS_VALUES) {
            const int64_t r = row0 + j0 + jl;
            int64_t part = 0;
            for (int64_t x = lane; x < w; x += 64) part += __popc(W[r * ldw + x]);  // full words
            vbase = row_ptr[r] + wave_sum(part);
          }
          if (lane < 32 && (mj & wbit(lane))) {
            const int64_t idx = (t * G + gj) * 32 + lane;
            if (member_row) member_row[idx] = j0 + jl;
            group_vals[idx] = HAS_VALUES ? values[vbase + rank_before(mj, lane)] : 1.0f;
          }
        }
      }
      const uint64_t spillers = nz & ~joined;
      if ((spillers >> lane) & 1ull) {
        slot = cnt + __popcll(spillers & ((1ull << lane) - 1ull));
        gid = -1;
        snnz_lane += __popc(m);
      }
      cnt += __popcll(spillers);
    };
    // the words of PF chunks are loaded before the first of them is placed:
    // the placement is sequential, the loads need not be (one load latency
    // per PF chunks instead of one per chunk).  The words and the chunks'
    // per-row results pass through LDS, so that the chunk loop is not
    // unrolled (its scalar state fits in SGPRs), and the results are stored
    // after the last chunk is placed: a store issued between the loads and
    // their use would make that use wait for the store too (one memory
    // counter covers both).
    const uint32_t vb = valid_bits(w, cols);
    for (int b0 = 0; b0 < nrows; b0 += 64 * kGroupPF) {
      const int nch = min(kGroupPF, (nrows - b0 + 63) / 64);
      // loads four chunks at a time under a wave-uniform test, so that a short
      // tile loads only its own chunks; within a quad the loads are
      // unconditional (rows past the tile read its last row and are zeroed
      // afterwards)
      uint32_t mm[kGroupPF];
#pragma unroll
      for (int q = 0; q < kGroupPF; q += 4) {
        if (q < nch) {
#pragma unroll
          for (int u = q; u < q + 4; ++u) {
            const int j = b0 + 64 * u + lane;
            mm[u] = W[(row0 + (j < nrows ? j : nrows - 1)) * ldw + w];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < kGroupPF; q += 4) {
        if (q < nch) {
#pragma unroll
          for (int u = q; u < q + 4; ++u)
            s_word[u * 64 + lane] = b0 + 64 * u + lane < nrows ? mm[u] & vb : 0u;
        }
      }
#pragma unroll 1
      for (int u = 0; u < nch; ++u) {
        int32_t gid, slot;
        chunk(b0 + 64 * u, s_word[u * 64 + lane], gid, slot);
        s_res[u * 64 + lane] = ((uint32_t)(slot + 1) << 8) | (uint32_t)(gid + 1);
      }
#pragma unroll 1
      for (int u = 0; u < nch; ++u) {
        const int j = b0 + 64 * u + lane;
        if (j < nrows) {
          const uint32_t res = s_res[u * 64 + lane];
          const int32_t slot = (int32_t)(res >> 8) - 1;
          group_id[(row0 + j) * ldw + w] = (int32_t)(res & 0xffu) - 1;
          if (spill_rev) spill_rev[t * T + j] = slot;
          if (spill_table && slot >= 0) spill_table[t * T + slot] = j;
        }
      }
    }
    if (spill_rev)
      for (int j = nrows + lane; j < T; j += 64) spill_rev[t * T + j] = -1;
    if (spill_table)
      for (int s = cnt + lane; s < T; s += 64) spill_table[t * T + s] = 0;
    if constexpr (kRegMembers) {
#pragma unroll
      for (int g = 0; g < (GC > 0 ? GC : 1); ++g) {
        if (lane < 32) {
          const bool set = (gms.get(g) & wbit(lane)) != 0u;
          const int64_t idx = (t * G + g) * 32 + lane;
          group_vals[idx] = set ? 1.0f : 0.0f;
          if (member_row) member_row[idx] = set ? mem[g] : -1;
        }
      }
    } else {
      for (int g = 0; g < G; ++g) {
        const uint32_t gmg = gms.get(g);
        if (lane < 32 && !(gmg & wbit(lane))) {
          const int64_t idx = (t * G + g) * 32 + lane;
          group_vals[idx] = 0.0f;
          if (member_row) member_row[idx] = -1;
        }
      }
    }
    const uint32_t gm_lane = gms.lane_value(lane);
    if (lane < G) group_mask[t * G + lane] = gm_lane;
    const int32_t snnz = wave_sum(snnz_lane);
    const int32_t tile_nnz = wave_sum(lane < G ? __popc(gm_lane) : 0);
    if (lane == 0) {
      spill_cnt[t] = cnt;
      spill_nnz[t] = snnz;
      if (grouped_nnz) atomicAdd(grouped_nnz, (unsigned long long)tile_nnz);
    }
  }
}

template <bool HAS_VALUES>
__global__ void __launch_bounds__(kBlock)
k_group_spilled_csr(const uint32_t* __restrict__ W, int64_t cols, int64_t wpr, int64_t ldw,
                    const int64_t* __restrict__ row_ptr, const float* __restrict__ values,
                    const int32_t* __restrict__ woff, int T, int64_t tiles,
                    const int32_t* __restrict__ spill_cnt, const int32_t* __restrict__ spill_table,
                    const int32_t* __restrict__ row_off, const int32_t* __restrict__ nnz_off,
                    int32_t* __restrict__ row_end, int32_t* __restrict__ col,
                    float* __restrict__ val) {
  const int lane = lane_id();
  for (int64_t t = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); t < tiles;
       t += (int64_t)gridDim.x * 4) {
    const int64_t rt = t / wpr, w = t - rt * wpr;
    const int cnt = spill_cnt[t];
    const int32_t roff = row_off[t], noff = nnz_off[t];
    int32_t carry = 0;
    for (int s0 = 0; s0 < cnt; s0 += 64) {
      const int s = s0 + lane;
      const bool valid = s < cnt;
      const int64_t r = rt * T + (valid ? spill_table[t * T + s] : 0);
      uint32_t m = valid ? (W[r * ldw + w] & valid_bits(w, cols)) : 0u;
      const int32_t p = __popc(m);
      const int32_t incl = wave_incl_scan(p);
      int32_t e = noff + carry + incl - p;
      if (valid) row_end[roff + s] = carry + incl;
      int64_t vb = HAS_VALUES && valid ? row_ptr[r] + woff[r * wpr + w] : 0;
      while (m) {
        const int i = __clz(m);
        m &= ~wbit(i);
        col[e] = i;
        val[e] = HAS_VALUES ? values[vb++] : 1.0f;
        ++e;
      }
      carry += __builtin_amdgcn_readlane(incl, 63);
    }
  }
}

template <bool GATHER>
__global__ void __launch_bounds__(kBlock)
k_group_scatter(const uint32_t* __restrict__ W, int64_t rows, int64_t cols, int64_t wpr,
                int64_t ldw,
                const int32_t* __restrict__ group_id, float* __restrict__ group_vals, int T,
                int G, float* __restrict__ D, int64_t ld) {
  const int64_t n = rows * wpr;
  for (int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * kBlock) {
    const int64_t r = e / wpr, w = e - r * wpr;
    const int32_t g = group_id[r * ldw + w];
    if (g < 0) continue;
    const int64_t t = (r / T) * wpr + w;
    uint32_t m = W[r * ldw + w] & valid_bits(w, cols);
    while (m) {
      const int i = __clz(m);
      m &= ~wbit(i);
      float* gv = &group_vals[(t * G + g) * 32 + i];
      float* d = &D[r * ld + 32 * w + i];
      if (GATHER) *gv = *d;
      else *d = *gv;
    }
  }
}

inline int grid_for(int64_t items, int64_t per_block) {
  int64_t g = ceil_div(items, per_block);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

int64_t bws_group_tiles(int64_t rows, int64_t cols, int32_t tile_rows) {
  if (rows < 0 || cols < 0 || tile_rows < 1) return -1;
  return ceil_div(rows, tile_rows) * words_for(cols);
}

bws_status bws_group_rows(const bws_bitmap* A, int32_t tile_rows, int32_t max_groups,
                          const bws_groups* out, bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (tile_rows < 1 || tile_rows > 65536 || max_groups < 1 || max_groups > 32 || !out)
    return BWS_EINVAL;
  const int64_t tiles = bws_group_tiles(A->rows, A->cols, tile_rows);
  if (tiles > 0 && (!out->group_id || !out->group_mask || !out->group_vals || !out->spill_cnt ||
                    !out->spill_nnz))
    return BWS_EINVAL;
  if (out->grouped_nnz) {
    hipLaunchKernelGGL(k_zero_i64, dim3(1), dim3(64), 0, stream, out->grouped_nnz);
  }
  if (tiles == 0) return launch_check();
  const int64_t wpr = words_for(A->cols);
  auto* gn = reinterpret_cast<unsigned long long*>(out->grouped_nnz);
  const int64_t per_xcd = ceil_div(tiles, 8);
  // a multiple of 8 (each workgroup keeps to one XCD's range), at most 65536
  const int grid = (int)(8 * per_xcd < 65536 ? 8 * per_xcd : 65536);
#define BWS_GROUP(HV, GC)                                                                   \
  hipLaunchKernelGGL((k_group_rows<HV, GC>), dim3(grid), dim3(64), 0, stream,                     \
                     A->words, A->rows, A->cols, wpr, A->ldw, A->row_ptr, A->values,            \
                     (int)tile_rows, (int)max_groups, tiles, out->group_id, out->group_mask,     \
                     out->group_vals, out->member_row, out->spill_cnt, out->spill_nnz,          \
                     out->spill_table, out->spill_rev, gn, per_xcd)
#define BWS_GROUP_G(HV)              \
  switch (max_groups) {              \
    case 1: BWS_GROUP(HV, 1); break; \
    case 2: BWS_GROUP(HV, 2); break; \
    case 3: BWS_GROUP(HV, 3); break; \
    case 4: BWS_GROUP(HV, 4); break; \
    default: BWS_GROUP(HV, 0);       \
  }
  if (A->values) {
    BWS_GROUP_G(true);
  } else {
    BWS_GROUP_G(false);
  }
#undef BWS_GROUP_G
#undef BWS_GROUP
  return launch_check();
}

size_t bws_group_spilled_workspace_size(int64_t rows, int64_t cols, int32_t tile_rows) {
  if (rows < 0 || cols < 0 || tile_rows < 1) return 0;
  const int64_t tiles = bws_group_tiles(rows, cols, tile_rows);
  Carve c(nullptr);
  c.take<int32_t>((size_t)(rows * words_for(cols)));
  c.take<char>(scan_workspace_bytes(tiles));
  return c.used + 256;
}

bws_status bws_group_spilled_csr(const bws_bitmap* A, int32_t tile_rows,
                                 const int32_t* spill_cnt, const int32_t* spill_nnz,
                                 const int32_t* spill_table, int32_t* spill_row_off,
                                 int32_t* spill_nnz_off, int32_t* row_end, int32_t* col,
                                 float* val, void* workspace, size_t workspace_bytes,
                                 bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (tile_rows < 1 || tile_rows > 65536) return BWS_EINVAL;
  const int64_t tiles = bws_group_tiles(A->rows, A->cols, tile_rows);
  if (!spill_row_off || !spill_nnz_off) return BWS_EINVAL;
  if (tiles > 0 && (!spill_cnt || !spill_nnz || !spill_table || !row_end || !col || !val))
    return BWS_EINVAL;
  if (!ws_ok(workspace, workspace_bytes,
               bws_group_spilled_workspace_size(A->rows, A->cols, tile_rows)))
    return BWS_EWORKSPACE;
  const int64_t wpr = words_for(A->cols);
  Carve c(workspace);
  int32_t* woff = c.take<int32_t>((size_t)(A->rows * wpr));
  void* scan_ws = c.take<char>(scan_workspace_bytes(tiles));
  BWS_TRY(launch_exclusive_scan_i32(spill_cnt, tiles, spill_row_off, scan_ws, stream));
  BWS_TRY(launch_exclusive_scan_i32(spill_nnz, tiles, spill_nnz_off, scan_ws, stream));
  if (tiles == 0) return BWS_OK;
  if (A->values) {
    BWS_TRY(launch_word_offsets(A->words, A->rows, A->cols, A->ldw, woff, stream));
    hipLaunchKernelGGL(k_group_spilled_csr<true>, dim3(grid_for(tiles, 4)), dim3(kBlock), 0,
                       stream, A->words, A->cols, wpr, A->ldw, A->row_ptr, A->values, woff,
                       (int)tile_rows, tiles, spill_cnt, spill_table, spill_row_off,
                       spill_nnz_off, row_end, col, val);
  } else {
    hipLaunchKernelGGL(k_group_spilled_csr<false>, dim3(grid_for(tiles, 4)), dim3(kBlock), 0,
                       stream, A->words, A->cols, wpr, A->ldw, A->row_ptr, A->values, woff,
                       (int)tile_rows, tiles, spill_cnt, spill_table, spill_row_off,
                       spill_nnz_off, row_end, col, val);
  }
  return launch_check();
}

bws_status bws_group_scatter(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                             const int32_t* group_id, const float* group_vals,
                             int32_t tile_rows, int32_t max_groups, float* D, int64_t ld,
                             bws_stream stream) {
  if (!dims_ok(rows, cols) || !ldw_ok(cols, ldw) || !ld_ok(cols, ld)) return BWS_EINVAL;
  if (tile_rows < 1 || tile_rows > 65536 || max_groups < 1 || max_groups > 32) return BWS_EINVAL;
  if (rows == 0 || cols == 0) return BWS_OK;
  if (!W || !group_id || !group_vals || !D) return BWS_EINVAL;
  const int64_t wpr = words_for(cols);
  hipLaunchKernelGGL(k_group_scatter<false>, dim3(grid_for(rows * wpr, kBlock)), dim3(kBlock), 0,
                     stream, W, rows, cols, wpr, ldw, group_id, const_cast<float*>(group_vals),
                     (int)tile_rows, (int)max_groups, D, ld);
  return launch_check();
}

bws_status bws_group_gather(const float* D, int64_t ld, const uint32_t* W, int64_t rows,
                            int64_t cols, int64_t ldw, const int32_t* group_id,
                            int32_t tile_rows, int32_t max_groups, float* group_vals,
                            bws_stream stream) {
  if (!dims_ok(rows, cols) || !ldw_ok(cols, ldw) || !ld_ok(cols, ld)) return BWS_EINVAL;
  if (tile_rows < 1 || tile_rows > 65536 || max_groups < 1 || max_groups > 32) return BWS_EINVAL;
  if (rows == 0 || cols == 0) return BWS_OK;
  if (!W || !group_id || !group_vals || !D) return BWS_EINVAL;
  const int64_t wpr = words_for(cols);
  hipLaunchKernelGGL(k_group_scatter<true>, dim3(grid_for(rows * wpr, kBlock)), dim3(kBlock), 0,
                     stream, W, rows, cols, wpr, ldw, group_id, group_vals, (int)tile_rows,
                     (int)max_groups, const_cast<float*>(D), ld);
  return launch_check();
}

}  // extern "C"

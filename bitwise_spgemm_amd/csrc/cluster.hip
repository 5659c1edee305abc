// cluster.hip — the row-clustering plan of the FP32 0/1 SpMM and its C ABI
// (gfx950); the product kernel itself is the streamed one (spmm_stream.hip).
//
// The reference groups rows by bitmask so that a group's rows share work
// (generate_groups_large, source/src/main.cu:50-247; backup/exp.cu:49-88).
// Here rows with similar masks are put into the same 16-row chunk; the
// product then sums a chunk base's bits once (T = base·B, the base being the
// majority mask of a run of rows with the same leader; a chunk has up to two)
// and each row only its difference from its base:
//   C[r] = T + (row & ~base)·B - (base & ~row)·B.
// When the rows of a chunk share most of their bits, that issues far fewer
// vector adds than summing every row's own bits, and the vector adds are what
// bounds the FP32 product on gfx950 (packed FP32 is not used).  The plan's
// operand stream (bws_cluster_stream) holds the add / subtract / base masks
// per chunk and 64-row stage in the layout the streamed kernel reads.
//
// Plan (device only, no host synchronisation, no memset: capturable):
//  1. leaders: c stride-sampled candidate rows (c = 256 for up to 256 leaders
//     from the search); their pairwise Hamming distances (clamped to 65535)
//     in a c x c table, a block per candidate; then one wave runs the
//     farthest-point search over the table in LDS.  More leaders than 256:
//     the candidates themselves;
//  2. every row's nearest of those leaders (16 rows per block pass, the
//     leaders' words staged in LDS, the rows' read with scalar loads), then
//     R = min(4, L - 1) exact farthest-point rounds over all rows: each round
//     is one launch that reduces the previous pass's per-block keys to its
//     leader (no same-address atomics);
//  3. rows stably ordered by nearest leader: per 256-row block a bitonic
//     sort (shuffles within a wave, LDS across waves) gives each row its
//     rank among the block's rows of the same
//     leader and the block's count of every leader; an exclusive scan over
//     (leader, block) gives every row its slot (done by each block of the
//     placing kernel itself up to 8192 pairs);
//  4. a block per chunk: its two longest runs of equal leader, each run's
//     majority mask (bit-sliced counters) as the chunk's bases, then a wave
//     per slot compares popcount(row ^ base_j) with popcount(row).
#include <algorithm>

#include "mfma_common.hpp"

namespace bws {

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s);

namespace {

constexpr int kChunk = 16;     // slots per chunk (rows per wave of the streamed product)
constexpr int32_t kMaxLeaders = 65536;
constexpr int kSearchMax = 256;  // candidates of the farthest-point search
constexpr int kRankBlock = 256;  // rows per block of the stable ranking
constexpr int kPadKey = kMaxLeaders;  // sorts after every leader

int32_t leaders_for(int64_t rows, int32_t max_clusters) {
  if (rows <= 0) return 0;
  int64_t L = max_clusters > 0 ? max_clusters : rows / 64;
  L = std::max<int64_t>(L, 1);
  L = std::min<int64_t>(L, rows);
  return (int32_t)std::min<int64_t>(L, kMaxLeaders);
}

// Leaders taken by exact farthest-point rounds over all rows (the last ones).
constexpr int32_t kExactRounds = 4;
int32_t exact_rounds(int32_t L) { return std::max(0, std::min(kExactRounds, L - 1)); }

// Candidate rows of the search for the first L - exact_rounds(L) leaders
// (candidate j = row j * rows / c).
int64_t candidates_for(int64_t rows, int32_t L) {
  if (rows <= 0) return 0;
  const int32_t Lc = L - exact_rounds(L);
  if (Lc > kSearchMax) return Lc;
  return std::min<int64_t>(rows, kSearchMax);
}

__host__ __device__ __forceinline__ int64_t cand_row(int64_t j, int64_t rows, int64_t c) {
  return j * rows / c;
}

// Rows per block pass of the nearest-leader kernels (4 per wave), and the
// cap on their grid (the exact rounds reduce one key per block).
constexpr int kPassRows = 16;
constexpr int64_t kMaxPassBlocks = 1024;
int64_t pass_blocks(int64_t rows) { return std::min<int64_t>(ceil_div(rows, kPassRows), kMaxPassBlocks); }

typedef const __attribute__((address_space(4))) uint32_t* sgpr_src;

// ---------------------------------------------------------------------------
// wave and block reductions
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}
struct OpMax {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};
struct OpMin {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a < b ? a : b; }
};
struct OpAdd {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
// 64-lane reduction, result uniform: xor 1 and xor 2 (quad permutes), then the
// half-row and row mirrors leave every lane of a 16-lane row with the row's
// result; the four rows are combined from lanes 0, 16, 32 and 48.  Every lane
// of the wave must be active.
template <class Op>
__device__ __forceinline__ uint32_t wave_reduce(uint32_t v, Op op) {
  v = op(v, dpp<0xB1>(v));
  v = op(v, dpp<0x4E>(v));
  v = op(v, dpp<0x141>(v));
  v = op(v, dpp<0x140>(v));
  const uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
  const uint32_t b = (uint32_t)__builtin_amdgcn_readlane((int)v, 16);
  const uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 32);
  const uint32_t d = (uint32_t)__builtin_amdgcn_readlane((int)v, 48);
  return op(op(a, b), op(c, d));
}
__device__ __forceinline__ uint64_t max_u64(uint64_t a, uint64_t b) { return a > b ? a : b; }
template <int CTRL>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
  return ((uint64_t)dpp<CTRL>(hi32(v)) << 32) | dpp<CTRL>(lo32(v));
}
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int l) {
  return ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi32(v), l) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)lo32(v), l);
}
__device__ __forceinline__ uint64_t wave_max_u64(uint64_t v) {
  v = max_u64(v, dpp64<0xB1>(v));
  v = max_u64(v, dpp64<0x4E>(v));
  v = max_u64(v, dpp64<0x141>(v));
  v = max_u64(v, dpp64<0x140>(v));
  return max_u64(max_u64(lane_u64(v, 0), lane_u64(v, 16)), max_u64(lane_u64(v, 32), lane_u64(v, 48)));
}
// Max over a 256-thread block (every thread reaches it), returned to all.
__device__ __forceinline__ uint64_t block_max_u64(uint64_t v, uint64_t* sh) {
  v = wave_max_u64(v);
  __syncthreads();  // sh may still be read from an earlier call
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  return max_u64(max_u64(sh[0], sh[1]), max_u64(sh[2], sh[3]));
}

// ---------------------------------------------------------------------------
// 1. leaders
// ---------------------------------------------------------------------------
constexpr int kDistWords = 128;  // words of a candidate row per pass

// D[i * c + j] = min(Hamming(candidate i, candidate j), 65535): a block per
// candidate i, its words staged in LDS 128 at a time; a thread per candidate
// j loads its own 128 words at once (32 16-byte loads when VEC: ldw % 4 == 0
// and W 16-byte aligned) and counts against the staged words.
template <bool VEC>
__global__ void __launch_bounds__(256)
k_cand_dist(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw, int64_t c,
            uint16_t* __restrict__ D) {
  __shared__ __align__(16) uint32_t si[kDistWords];
  const int64_t i = blockIdx.x;
  const int j = threadIdx.x;
  const int64_t wpr = words_for(K);
  const uint32_t* ri = W + cand_row(i, M, c) * ldw;
  const uint32_t* rj = W + cand_row(j < c ? j : 0, M, c) * ldw;
  uint32_t d = 0;
  for (int64_t w0 = 0; w0 < wpr; w0 += kDistWords) {
    const int n = (int)std::min<int64_t>(kDistWords, wpr - w0);
    __syncthreads();
    if (j < kDistWords) si[j] = j < n ? ri[w0 + j] & valid_bits(w0 + j, K) : 0u;
    __syncthreads();
    if (j < c) {
      if (VEC && n == kDistWords) {
        uint4 x[kDistWords / 4];
#pragma unroll
        for (int s = 0; s < kDistWords / 4; ++s) x[s] = reinterpret_cast<const uint4*>(rj + w0)[s];
#pragma unroll
        for (int s = 0; s < kDistWords / 4; ++s) {
          const uint4 y = reinterpret_cast<const uint4*>(si)[s];
          d += __popc(y.x ^ x[s].x) + __popc(y.y ^ x[s].y) + __popc(y.z ^ x[s].z) +
               __popc(y.w ^ x[s].w);
        }
      } else {
#pragma unroll 8
        for (int e = 0; e < n; ++e) d += __popc(si[e] ^ rj[w0 + e]);
      }
    }
  }
  if (j < c) {
    if (wpr > 0) d -= __popc(rj[wpr - 1] & ~valid_bits(wpr - 1, K));  // row j's bits past K
    D[i * c + j] = (uint16_t)(d < 65535u ? d : 65535u);
  }
}

// Farthest-point search over c <= 256 candidates: the block stages the table
// in LDS, then one wave runs the search (lane l holds candidates l, l + 64,
// l + 128, l + 192): leader 0 is candidate 0; each candidate keeps its
// distance to the nearest leader so far; the next leader is the candidate
// with the largest (lowest index on ties).  Leaders' rows are written after
// the search.
__global__ void __launch_bounds__(kSearchMax)
k_leader_search(const uint16_t* __restrict__ D, int64_t c, int64_t M, int32_t L,
                int32_t* __restrict__ lead_row) {
  // the table and one spare 16-byte slot (the target of loads past its end)
  __shared__ __align__(16) uint16_t sD[kSearchMax * kSearchMax + 8];
  __shared__ int32_t sidx[kSearchMax];
  const int64_t n = c * c, n8 = n / 8;
  // D is 256-byte aligned: 16-byte loads, eight in flight per thread (loads
  // and stores unconditional, so that none waits for another)
  const uint4* src = reinterpret_cast<const uint4*>(D);
  uint4* dst = reinterpret_cast<uint4*>(sD);
  constexpr int64_t kSpare = (int64_t)kSearchMax * kSearchMax / 8;
  for (int64_t e0 = threadIdx.x; e0 < n8; e0 += 8 * kSearchMax) {
    uint4 v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t e = e0 + u * kSearchMax;
      v[u] = src[e < n8 ? e : 0];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t e = e0 + u * kSearchMax;
      dst[e < n8 ? e : kSpare] = v[u];
    }
  }
  for (int64_t e = n8 * 8 + threadIdx.x; e < n; e += kSearchMax) sD[e] = D[e];
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int lane = threadIdx.x;
  const int cc = (int)c;
  uint32_t near[4] = {0u, 0u, 0u, 0u};
  int lead = 0;
  for (int32_t it = 0; it < L; ++it) {
    if (lane == 0) sidx[it] = lead;
    const uint16_t* row = sD + lead * cc;
    uint32_t key = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int j = lane + 64 * q;
      const uint32_t d = j < cc ? row[j] : 0u;
      near[q] = (it == 0 || d < near[q]) ? d : near[q];
      const uint32_t k = j < cc ? (near[q] << 16) | (uint32_t)(0xFFFF - j) : 0u;
      key = k > key ? k : key;
    }
    lead = 0xFFFF - (int)(wave_reduce(key, OpMax()) & 0xFFFF);
  }
  for (int it = lane; it < L; it += 64) lead_row[it] = (int32_t)cand_row(sidx[it], M, c);
}

// More leaders than the search takes: leader i is candidate i.
__global__ void __launch_bounds__(256)
k_stride_leaders(int64_t M, int32_t L, int32_t* __restrict__ lead_row) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < L) lead_row[i] = (int32_t)cand_row(i, M, L);
}

// ---------------------------------------------------------------------------
// 2. nearest leaders
// ---------------------------------------------------------------------------
constexpr int kLeadWords = 128;  // words per LDS pass
constexpr int kLeadPitch = 65;   // padded LDS row (one word of 64 leaders)

// Nearest of the first Lc leaders for every row (full Hamming distance,
// earliest leader on ties).  A block takes 16 rows at a time, 4 per wave;
// lane i of a wave is leader g + i of a group of 64.  Per pass of 128 words
// the block stages the group's leader words (transposed) and its rows' words
// in LDS, bits past K cleared.  The row's distance goes to dist and the
// block's max of (distance, ~row) over its rows to part[block] (reduced by
// the first exact round).
__global__ void __launch_bounds__(256)
k_assign(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
         const int32_t* __restrict__ lead_row, int32_t Lc, int32_t* __restrict__ owner,
         int32_t* __restrict__ dist, unsigned long long* __restrict__ part) {
  __shared__ uint32_t sl[kLeadWords * kLeadPitch];
  __shared__ __align__(16) uint32_t srw[kPassRows][kLeadWords];
  __shared__ int32_t slr[64];
  __shared__ uint64_t sh[4];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t wpr = words_for(K);
  uint64_t far_key = 0ull;
  for (int64_t rb = blockIdx.x; rb * kPassRows < M; rb += gridDim.x) {
    uint32_t best_d[4] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
    uint32_t best_i[4] = {0u, 0u, 0u, 0u};
    for (int32_t g = 0; g < Lc; g += 64) {
      __syncthreads();  // slr is rewritten
      if (t < 64) slr[t] = lead_row[g + t < Lc ? g + t : g];
      uint32_t d[4] = {0u, 0u, 0u, 0u};
      for (int64_t w0 = 0; w0 < wpr; w0 += kLeadWords) {
        const int n = (int)std::min<int64_t>(kLeadWords, wpr - w0);
        __syncthreads();  // slr written; the previous pass's readers are done
        {
          // unconditional loads from clamped addresses (every one in flight
          // at once), words past n or rows past M zeroed after
          const int k = t & 127, i0 = t >> 7;
          const int kk = k < n ? k : 0;
          const uint32_t vb = k < n ? valid_bits(w0 + k, K) : 0u;
          uint32_t v[32];
#pragma unroll
          for (int s = 0; s < 32; ++s) v[s] = W[(int64_t)slr[i0 + 2 * s] * ldw + w0 + kk];
#pragma unroll
          for (int s = 0; s < 32; ++s) sl[k * kLeadPitch + i0 + 2 * s] = v[s] & vb;
          uint32_t x[8];
#pragma unroll
          for (int s = 0; s < 8; ++s) {
            const int64_t r = rb * kPassRows + i0 + 2 * s;
            x[s] = W[(r < M ? r : M - 1) * ldw + w0 + kk];
          }
#pragma unroll
          for (int s = 0; s < 8; ++s)
            srw[i0 + 2 * s][k] = rb * kPassRows + i0 + 2 * s < M ? x[s] & vb : 0u;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t* rw = srw[wave * 4 + q];
          uint32_t acc = d[q];
          for (int k = 0; k < n; k += 4) {
            const uint4 y = *reinterpret_cast<const uint4*>(rw + k);
            acc += __popc(sl[k * kLeadPitch + lane] ^ y.x) +
                   __popc(sl[(k + 1) * kLeadPitch + lane] ^ y.y) +
                   __popc(sl[(k + 2) * kLeadPitch + lane] ^ y.z) +
                   __popc(sl[(k + 3) * kLeadPitch + lane] ^ y.w);
          }
          d[q] = acc;
        }
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t v = g + lane < Lc ? d[q] : 0xFFFFFFFFu;
        const uint32_t m = wave_reduce(v, OpMin());
        const uint32_t f = (uint32_t)__builtin_ctzll(__ballot(v == m));
        if (m < best_d[q]) {
          best_d[q] = m;
          best_i[q] = (uint32_t)g + f;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t r = rb * kPassRows + wave * 4 + q;
      if (r < M) {
        if (lane == 0) {
          owner[r] = (int32_t)best_i[q];
          dist[r] = (int32_t)best_d[q];
        }
        far_key = max_u64(far_key, ((uint64_t)best_d[q] << 32) | (uint32_t)~(uint32_t)r);
      }
    }
  }
  const uint64_t bm = block_max_u64(far_key, sh);
  if (t == 0) part[blockIdx.x] = bm;
}

// Exact round `it`: leader Lc + it is the row of the largest of the previous
// pass's block keys prev[0..np) (the row farthest from its nearest leader,
// lowest row on ties); every row takes its distance to it (strictly nearer:
// the earlier leader keeps ties) and the block's max of (distance, ~row) over
// its rows goes to next[block].  16 rows per block pass, 4 per wave, a lane
// per word; the first pass's first 128 words of each row are loaded before
// the leader is known.
__global__ void __launch_bounds__(256)
k_exact_round(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw, int32_t Lc,
              int32_t it, const unsigned long long* __restrict__ prev, int64_t np,
              unsigned long long* __restrict__ next, int32_t* __restrict__ owner,
              int32_t* __restrict__ dist) {
  __shared__ uint64_t sh[4];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int64_t wpr = words_for(K);
  const int64_t rb0 = blockIdx.x;
  uint32_t pre[4][2];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t r = std::min<int64_t>(rb0 * kPassRows + wave * 4 + q, M - 1);
#pragma unroll
    for (int u = 0; u < 2; ++u) {  // clamped load, then zeroed past the row
      const int64_t w = lane + 64 * u;
      const uint32_t x = W[r * ldw + (w < wpr ? w : 0)];
      pre[q][u] = w < wpr ? x & valid_bits(w, K) : 0u;
    }
  }
  uint64_t k0 = 0ull;
  for (int64_t e = t; e < np; e += 256) k0 = max_u64(k0, prev[e]);
  const int64_t lead = (int64_t)(uint32_t)~(uint32_t)block_max_u64(k0, sh);
  const uint32_t* lw = W + lead * ldw;
  uint64_t far_key = 0ull;
  for (int64_t rb = rb0; rb * kPassRows < M; rb += gridDim.x) {
    const int64_t r0 = rb * kPassRows + wave * 4;
    uint32_t d[4] = {0u, 0u, 0u, 0u};
    for (int64_t wb = 0; wb < wpr; wb += 128) {
      // clamped loads, all in flight at once; words past the row count zero
      const bool first = rb == rb0 && wb == 0;
      int64_t wc[2];
      uint32_t vb[2], l[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int64_t w = wb + lane + 64 * u;
        wc[u] = w < wpr ? w : 0;
        vb[u] = w < wpr ? valid_bits(w, K) : 0u;
        l[u] = lw[wc[u]];
      }
      if (first) {
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int q = 0; q < 4; ++q) d[q] += __popc(pre[q][u] ^ (l[u] & vb[u]));
      } else {
        uint32_t x[4][2];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t r = r0 + q < M ? r0 + q : M - 1;
#pragma unroll
          for (int u = 0; u < 2; ++u) x[q][u] = W[r * ldw + wc[u]];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
          for (int q = 0; q < 4; ++q) d[q] += __popc((x[q][u] ^ l[u]) & vb[u]);
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint32_t dq = wave_reduce(d[q], OpAdd());
      const int64_t r = r0 + q;
      if (r < M) {
        const int32_t pd = dist[r];
        const bool nearer = (int32_t)dq < pd;
        if (lane == 0 && nearer) {
          dist[r] = (int32_t)dq;
          owner[r] = Lc + it;
        }
        const uint32_t dd = nearer ? dq : (uint32_t)pd;
        far_key = max_u64(far_key, ((uint64_t)dd << 32) | (uint32_t)~(uint32_t)r);
      }
    }
  }
  const uint64_t bm = block_max_u64(far_key, sh);
  if (t == 0) next[blockIdx.x] = bm;
}

// ---------------------------------------------------------------------------
// 3. stable order by leader
// ---------------------------------------------------------------------------
// Stable ranking of a 256-row block by leader: bitonic sort of (leader << 8 |
// row in block), a key per thread (partners within a wave exchanged by
// shuffles, across waves through LDS); a row's rank is its sorted position
// minus the start of its leader's run (a max-scan of run starts); the
// block's count of every leader l < L (zero when it has none) goes to
// hist[l * nb + block], by binary search over the sorted keys.
__global__ void __launch_bounds__(kRankBlock)
k_block_rank(const int32_t* __restrict__ owner, int64_t M, int64_t nb, int32_t L,
             int32_t* __restrict__ lrank, int32_t* __restrict__ hist) {
  static_assert(kRankBlock == 256, "keys hold the row in block in 8 bits");
  __shared__ uint32_t sk[kRankBlock];
  __shared__ int32_t wmax[kRankBlock / 64];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int64_t b = blockIdx.x;
  const int64_t r = b * kRankBlock + t;
  uint32_t key = ((r < M ? (uint32_t)owner[r] : (uint32_t)kPadKey) << 8) | (uint32_t)t;
  for (int k = 2; k <= kRankBlock; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint32_t other;
      if (j >= 64) {
        __syncthreads();  // the previous exchange is read
        sk[t] = key;
        __syncthreads();
        other = sk[t ^ j];
      } else {
        other = (uint32_t)__shfl_xor((int)key, j, 64);
      }
      // ascending runs: the lower position of a pair keeps the smaller key
      const bool keep_min = ((t & j) == 0) == ((t & k) == 0);
      key = keep_min ? (key < other ? key : other) : (key > other ? key : other);
    }
  }
  __syncthreads();
  sk[t] = key;
  __syncthreads();
  // start of each position's run: inclusive max-scan of run starts
  const bool head = t == 0 || (sk[t - 1] >> 8) != (key >> 8);
  int32_t v = head ? t : 0;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v = u > v ? u : v;
  }
  if (lane == 63) wmax[wave] = v;
  __syncthreads();
  for (int w = 0; w < wave; ++w) v = wmax[w] > v ? wmax[w] : v;
  const int64_t row = b * kRankBlock + (int64_t)(key & 255u);
  if (row < M) lrank[row] = t - v;
  // first position whose leader is >= x
  auto lower = [&](uint32_t x) {
    int lo = 0, hi = kRankBlock;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if ((sk[mid] >> 8) < x) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  for (int64_t l = t; l < L; l += kRankBlock)
    hist[l * nb + b] = lower((uint32_t)l + 1) - lower((uint32_t)l);
}

// Slot of every row: off[leader * nb + block] + rank in block.
__global__ void __launch_bounds__(256)
k_place(const int32_t* __restrict__ owner, const int32_t* __restrict__ lrank,
        const int32_t* __restrict__ off, int64_t M, int64_t nb, int32_t* __restrict__ perm,
        int32_t* __restrict__ keys) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= M) return;
  const int32_t o = owner[r];
  const int64_t pos = (int64_t)off[(int64_t)o * nb + r / kRankBlock] + lrank[r];
  perm[pos] = (int32_t)r;
  keys[pos] = o;
}

// k_place for up to kPlaceScan (leader, block) pairs: every block scans the
// counts itself in LDS (a thread per run of 32 counts, then a scan of the
// 256 run totals) instead of a separate scan launch.
constexpr int kPlaceScan = 8192;
__global__ void __launch_bounds__(256)
k_place_scan(const int32_t* __restrict__ owner, const int32_t* __restrict__ lrank,
             const int32_t* __restrict__ hist, int64_t pairs, int64_t M, int64_t nb,
             int32_t* __restrict__ perm, int32_t* __restrict__ keys) {
  __shared__ int32_t so[kPlaceScan];
  __shared__ int32_t tot[256];
  const int t = threadIdx.x;
  for (int64_t e = t; e < pairs; e += 256) so[e] = hist[e];
  __syncthreads();
  const int per = (int)ceil_div(pairs, 256);
  const int64_t a = (int64_t)t * per, z = std::min<int64_t>(a + per, pairs);
  int32_t s = 0;
  for (int64_t e = a; e < z; ++e) s += so[e];
  tot[t] = s;
  __syncthreads();
  // inclusive scan of the 256 totals (Hillis-Steele)
  for (int o = 1; o < 256; o <<= 1) {
    const int32_t v = t >= o ? tot[t - o] : 0;
    __syncthreads();
    tot[t] += v;
    __syncthreads();
  }
  int32_t run = t > 0 ? tot[t - 1] : 0;
  for (int64_t e = a; e < z; ++e) {
    const int32_t v = so[e];
    so[e] = run;
    run += v;
  }
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * 256 + t;
  if (r >= M) return;
  const int32_t o = owner[r];
  const int64_t pos = (int64_t)so[(int64_t)o * nb + r / kRankBlock] + lrank[r];
  perm[pos] = (int32_t)r;
  keys[pos] = o;
}

// ---------------------------------------------------------------------------
// 4, 5. chunk bases and slot flags: a block per chunk of 16 slots
// ---------------------------------------------------------------------------
// The chunk's two longest runs of equal leader (the earlier on equal
// lengths); each base's word w is the majority (count * 2 > n + 1) of its
// run's rows, counted with 5-bit bit-sliced counters; then a wave per slot
// compares popcount(row ^ base_j) with popcount(row) and flags the base with
// the fewest bits when that is fewer than the row's own.
__global__ void __launch_bounds__(256)
k_chunk_plan(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const int32_t* __restrict__ perm, const int32_t* __restrict__ keys,
             uint32_t* __restrict__ CM, int32_t* __restrict__ slot_row) {
  __shared__ int4 srun;
  __shared__ int32_t cp[kChunk], ck[kChunk];
  const int64_t c = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63;
  const int n = (int)std::min<int64_t>(kChunk, M - kChunk * c);
  if (t < n) {
    cp[t] = perm[kChunk * c + t];
    ck[t] = keys[kChunk * c + t];
  }
  __syncthreads();
  if (t == 0) {
    int s0 = 0, n0 = 0, s1 = 0, n1 = 0;
    for (int i = 0; i < n;) {
      int j = i + 1;
      while (j < n && ck[j] == ck[i]) ++j;
      const int len = j - i;
      if (len > n0) {
        s1 = s0, n1 = n0;
        s0 = i, n0 = len;
      } else if (len > n1) {
        s1 = i, n1 = len;
      }
      i = j;
    }
    srun = make_int4(s0, n0, s1, n1);
  }
  __syncthreads();
  const int4 rn = srun;
  const int64_t wpr = words_for(K);
  uint32_t* cm = CM + 2 * c * ldw;
  for (int64_t e = t; e < 2 * ldw; e += 256) {
    const int b = e >= ldw ? 1 : 0;
    const int64_t w = e - b * ldw;
    const int start = b ? rn.z : rn.x, len = b ? rn.w : rn.y;
    uint32_t out = 0u;
    if (w < wpr && len > 0) {
      const uint32_t vb = valid_bits(w, K);
      uint32_t xs[kChunk];  // clamped loads all in flight, then zeroed past the run
#pragma unroll
      for (int i = 0; i < kChunk; ++i) xs[i] = W[(int64_t)cp[start + (i < len ? i : 0)] * ldw + w];
#pragma unroll
      for (int i = 0; i < kChunk; ++i) xs[i] = i < len ? xs[i] & vb : 0u;
      uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0;  // bit-sliced count, up to 16
#pragma unroll
      for (int i = 0; i < kChunk; ++i) {
        uint32_t x = xs[i];
        uint32_t y = c0 & x; c0 ^= x; x = y;
        y = c1 & x; c1 ^= x; x = y;
        y = c2 & x; c2 ^= x; x = y;
        y = c3 & x; c3 ^= x; x = y;
        c4 ^= x;
      }
      // count >= T with T = (len + 1) / 2 + 1, compared from the top bit
      const uint32_t T = (uint32_t)(len + 1) / 2 + 1;
      const uint32_t cs[5] = {c0, c1, c2, c3, c4};
      uint32_t gt = 0u, eq = 0xFFFFFFFFu;
#pragma unroll
      for (int q = 4; q >= 0; --q) {
        if ((T >> q) & 1u) {
          eq &= cs[q];
        } else {
          gt |= eq & cs[q];
          eq &= ~cs[q];
        }
      }
      out = gt | eq;
    }
    cm[e] = out;
  }
  __syncthreads();  // the bases are read back by every wave
  const uint32_t* m0 = cm;
  const uint32_t* m1 = cm + ldw;
  for (int j = t >> 6; j < n; j += 4) {
    const int64_t s = kChunk * c + j;
    const int32_t r = cp[j];
    const uint32_t* a = W + (int64_t)r * ldw;
    uint32_t pa = 0, p0 = 0, p1 = 0;
    for (int64_t w = lane; w < wpr; w += 64) {
      const uint32_t x = a[w] & valid_bits(w, K);
      pa += __popc(x);
      p0 += __popc(x ^ m0[w]);
      p1 += __popc(x ^ m1[w]);
    }
    pa = wave_reduce(pa, OpAdd());
    p0 = wave_reduce(p0, OpAdd());
    p1 = wave_reduce(p1, OpAdd());
    int32_t flag = 0;
    uint32_t best = pa;
    if (p0 < best) flag = BWS_SLOT_DELTA, best = p0;
    if (p1 < best) flag = BWS_SLOT_DELTA | BWS_SLOT_BASE1;
    if (lane == 0) slot_row[s] = r | flag;
  }
}

struct ClusterWs {
  uint16_t* D;
  int32_t* lead_row;
  int32_t* owner;
  int32_t* dist;
  unsigned long long* part;
  int32_t* lrank;
  int32_t* hist;
  int32_t* off;
  void* scan_ws;
  int32_t* keys;
  int32_t* perm;
  int64_t c, nb, G;
  size_t total;
};

ClusterWs carve_cluster(void* base, int64_t M, int L) {
  ClusterWs w{};
  w.c = candidates_for(M, L);
  w.nb = ceil_div(M, kRankBlock);
  w.G = pass_blocks(M);
  const int32_t Lc = L - exact_rounds(L);
  const bool search = Lc <= kSearchMax;
  const int64_t pairs = (int64_t)L * w.nb;
  const bool big = pairs > kPlaceScan;
  Carve cv(base);
  w.D = cv.take<uint16_t>(search ? (size_t)(w.c * w.c) : 0);
  w.lead_row = cv.take<int32_t>((size_t)L);
  w.owner = cv.take<int32_t>((size_t)M);
  w.dist = cv.take<int32_t>((size_t)M);
  w.part = cv.take<unsigned long long>((size_t)((exact_rounds(L) + 1) * w.G));
  w.lrank = cv.take<int32_t>((size_t)M);
  w.hist = cv.take<int32_t>((size_t)pairs);
  w.off = cv.take<int32_t>(big ? (size_t)pairs + 1 : 0);
  w.scan_ws = cv.take<uint8_t>(big ? scan_workspace_bytes(pairs) : 0);
  w.keys = cv.take<int32_t>((size_t)M);
  w.perm = cv.take<int32_t>((size_t)M);
  w.total = cv.used;
  return w;
}

// (leader, block) pairs of the stable ranking: the plan's capacity.
constexpr int64_t kMaxRankPairs = (int64_t)1 << 27;

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_cluster_workspace_size(int64_t rows, int64_t cols, int32_t max_clusters) {
  if (rows <= 0 || rows >= BWS_SLOT_BASE1 || cols < 0 || max_clusters < 0) return 0;
  const int L = leaders_for(rows, max_clusters);
  if ((int64_t)L * ceil_div(rows, kRankBlock) > kMaxRankPairs) return 0;
  return carve_cluster(nullptr, rows, L).total + 256;
}

bws_status bws_cluster_rows(const bws_bitmap* A, int32_t max_clusters, int32_t* slot_row,
                            uint32_t* chunk_mask, void* workspace, size_t workspace_bytes,
                            bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (A->values) return BWS_EUNSUPPORTED;
  if (max_clusters < 0) return BWS_EINVAL;
  if (A->rows >= BWS_SLOT_BASE1) return BWS_ECAPACITY;
  if (A->rows == 0) return BWS_OK;
  if (!slot_row || !chunk_mask) return BWS_EINVAL;
  const int64_t M = A->rows, K = A->cols;
  const int L = leaders_for(M, max_clusters);
  if ((int64_t)L * ceil_div(M, kRankBlock) > kMaxRankPairs) return BWS_ECAPACITY;
  if (!ws_ok(workspace, workspace_bytes, bws_cluster_workspace_size(M, K, max_clusters)))
    return BWS_EWORKSPACE;
  hipStream_t s = stream;
  const ClusterWs w = carve_cluster(workspace, M, L);
  auto grid = [](int64_t n, int per) { return dim3((unsigned)std::max<int64_t>(1, ceil_div(n, per))); };
  // 1. leaders: the first Lc by the candidate search (or the candidates)
  const int32_t R = exact_rounds(L), Lc = L - R;
  if (Lc <= kSearchMax) {
    const bool vec = A->ldw % 4 == 0 && reinterpret_cast<uintptr_t>(A->words) % 16 == 0;
    if (vec)
      hipLaunchKernelGGL(k_cand_dist<true>, dim3((unsigned)w.c), dim3(256), 0, s, A->words, M, K,
                         A->ldw, w.c, w.D);
    else
      hipLaunchKernelGGL(k_cand_dist<false>, dim3((unsigned)w.c), dim3(256), 0, s, A->words, M, K,
                         A->ldw, w.c, w.D);
    BWS_TRY(launch_check());
    hipLaunchKernelGGL(k_leader_search, dim3(1), dim3(kSearchMax), 0, s, w.D, w.c, M, Lc,
                       w.lead_row);
  } else {
    hipLaunchKernelGGL(k_stride_leaders, grid(Lc, 256), dim3(256), 0, s, M, Lc, w.lead_row);
  }
  BWS_TRY(launch_check());
  // 2. nearest of those leaders, then the exact rounds over all rows; each
  //    pass leaves one key per block for the next round's leader
  hipLaunchKernelGGL(k_assign, dim3((unsigned)w.G), dim3(256), 0, s, A->words, M, K, A->ldw,
                     w.lead_row, Lc, w.owner, w.dist, w.part);
  BWS_TRY(launch_check());
  for (int32_t it = 0; it < R; ++it) {
    hipLaunchKernelGGL(k_exact_round, dim3((unsigned)w.G), dim3(256), 0, s, A->words, M, K, A->ldw,
                       Lc, it, w.part + it * w.G, w.G, w.part + (it + 1) * w.G, w.owner, w.dist);
    BWS_TRY(launch_check());
  }
  // 3. stable order by leader
  const int64_t pairs = (int64_t)L * w.nb;
  hipLaunchKernelGGL(k_block_rank, dim3((unsigned)w.nb), dim3(kRankBlock), 0, s, w.owner, M, w.nb,
                     (int32_t)L, w.lrank, w.hist);
  BWS_TRY(launch_check());
  if (pairs <= kPlaceScan) {
    hipLaunchKernelGGL(k_place_scan, grid(M, 256), dim3(256), 0, s, w.owner, w.lrank, w.hist,
                       pairs, M, w.nb, w.perm, w.keys);
  } else {
    BWS_TRY(launch_exclusive_scan_i32(w.hist, pairs, w.off, w.scan_ws, s));
    hipLaunchKernelGGL(k_place, grid(M, 256), dim3(256), 0, s, w.owner, w.lrank, w.off, M, w.nb,
                       w.perm, w.keys);
  }
  BWS_TRY(launch_check());
  // 4, 5. chunk bases and slot flags
  hipLaunchKernelGGL(k_chunk_plan, dim3((unsigned)ceil_div(M, kChunk)), dim3(256), 0, s, A->words,
                     M, K, A->ldw, w.perm, w.keys, chunk_mask, slot_row);
  return launch_check();
}

size_t bws_cluster_stream_size(int64_t rows, int64_t cols) {
  if (rows < 0 || cols < 0) return 0;
  return cluster_stream_bytes(rows, cols);
}

bws_status bws_cluster_stream(const bws_bitmap* A, const int32_t* slot_row,
                              const uint32_t* chunk_mask, void* plan_stream, size_t stream_bytes,
                              bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (A->values) return BWS_EUNSUPPORTED;
  if (A->rows >= BWS_SLOT_BASE1) return BWS_ECAPACITY;
  if (A->rows == 0 || A->cols == 0) return BWS_OK;
  if (!slot_row || !chunk_mask) return BWS_EINVAL;
  if (!ws_ok(plan_stream, stream_bytes, cluster_stream_bytes(A->rows, A->cols)))
    return BWS_EWORKSPACE;
  return launch_cluster_stream(A, slot_row, chunk_mask, static_cast<uint32_t*>(plan_stream),
                               stream);
}

bws_status bws_spmm_clustered(const bws_bitmap* A, const int32_t* slot_row,
                              const void* plan_stream, const float* B, int64_t ldb, int64_t N,
                              float* C, int64_t ldc, float alpha, float beta, bws_stream stream) {
  BWS_TRY(check_bitmap(A, false));
  if (A->values) return BWS_EUNSUPPORTED;
  if (N < 0 || !ld_ok(N, ldc) || (A->cols > 0 && !ld_ok(N, ldb))) return BWS_EINVAL;
  if (A->rows >= BWS_SLOT_BASE1) return BWS_ECAPACITY;
  if (A->rows == 0 || N == 0) return BWS_OK;
  if (!C || (A->cols > 0 && !B)) return BWS_EINVAL;
  hipStream_t s = stream;
  if (A->cols == 0)  // C = beta * C
    return spmm_dispatch(A, B, BWS_DTYPE_F32, ldb, N, C, ldc, alpha, beta, BWS_PREC_FP32,
                         nullptr, s);
  if (!slot_row || !plan_stream) return BWS_EINVAL;
  if (N % 4 != 0 || ldb % 4 != 0 || ldc % 4 != 0 || reinterpret_cast<uintptr_t>(B) % 16 != 0 ||
      reinterpret_cast<uintptr_t>(C) % 16 != 0 || reinterpret_cast<uintptr_t>(plan_stream) % 256)
    return BWS_EUNSUPPORTED;
  return launch_spmm_stream(static_cast<const uint32_t*>(plan_stream), slot_row, A->rows, A->cols,
                            B, ldb, N, C, ldc, alpha, beta, s);
}

}  // extern "C"

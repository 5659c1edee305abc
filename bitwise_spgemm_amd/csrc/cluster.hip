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
//  1. leaders: c stride-sampled candidate rows (c = 256 for up to 256 leaders);
//     their pairwise Hamming distances (clamped to 65535) in a c x c table,
//     then one workgroup runs the farthest-point search over the table in
//     LDS.  More than 256 leaders: the candidates themselves;
//  2. every row's nearest leader: a wave per row, a lane per leader;
//  3. rows stably ordered by nearest leader: per 1024-row block a bitonic
//     sort in LDS gives each row its rank among the block's rows of the same
//     leader and the block's count per leader; one exclusive scan over
//     (leader, block) gives every row its slot;
//  4. chunk runs: a thread per chunk finds its two longest runs of equal
//     leader; chunk masks: a thread per (chunk, base, word) counts the run's
//     bits;
//  5. slot flags: a wave per slot compares popcount(row ^ base_j) with
//     popcount(row).
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
constexpr int kRankBlock = 1024; // rows per block of the stable ranking
constexpr int kPadKey = kMaxLeaders;  // sorts after every leader

int32_t leaders_for(int64_t rows, int32_t max_clusters) {
  if (rows <= 0) return 0;
  int64_t L = max_clusters > 0 ? max_clusters : rows / 64;
  L = std::max<int64_t>(L, 1);
  L = std::min<int64_t>(L, rows);
  return (int32_t)std::min<int64_t>(L, kMaxLeaders);
}

// Leaders taken by exact farthest-point rounds over all rows (the last ones).
constexpr int32_t kExactRounds = 8;
int32_t exact_rounds(int32_t L) { return std::max(0, std::min(kExactRounds, L - 1)); }

// Candidate rows of the search for the first L - exact_rounds(L) leaders
// (candidate j = row j * rows / c).
int64_t candidates_for(int64_t rows, int32_t L) {
  if (rows <= 0) return 0;
  const int32_t Lc = L - exact_rounds(L);
  if (Lc > kSearchMax) return Lc;
  return std::min<int64_t>(rows, kSearchMax);
}

__device__ __forceinline__ int64_t cand_row(int64_t j, int64_t rows, int64_t c) {
  return j * rows / c;
}

// Transposed words of candidate (or leader) rows: T[w * n + j] = word w of
// row rowsel(j), bits past K cleared.  A thread per (w, j).
__global__ void __launch_bounds__(256)
k_rows_t(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
         const int32_t* __restrict__ rowsel, int64_t n, uint32_t* __restrict__ T) {
  const int64_t wpr = words_for(K);
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= wpr * n) return;
  const int64_t w = t / n, j = t - w * n;
  const int64_t r = rowsel ? (int64_t)rowsel[j] : cand_row(j, M, n);
  T[t] = W[r * ldw + w] & valid_bits(w, K);
}

// D[i * c + j] = min(Hamming(candidate i, candidate j), 65535): a wave per
// (i, 64 candidates j), a lane per j.
__global__ void __launch_bounds__(256)
k_cand_dist(const uint32_t* __restrict__ T, int64_t wpr, int64_t c, uint16_t* __restrict__ D) {
  const int lane = threadIdx.x & 63;
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t jb = ceil_div(c, 64);
  if (g >= c * jb) return;
  const int64_t i = g / jb, j = (g - i * jb) * 64 + lane;
  if (j >= c) return;
  int64_t d = 0;
  for (int64_t w = 0; w < wpr; ++w) d += __popc(T[w * c + i] ^ T[w * c + j]);
  D[i * c + j] = (uint16_t)(d < 65535 ? d : 65535);
}

// Farthest-point search over c <= 256 candidates, one workgroup of 256
// threads (thread j = candidate j), the distance table in LDS: leader 0 is
// candidate 0; each candidate keeps its distance to the nearest leader so
// far; the next leader is the candidate with the largest (lowest j on ties).
__global__ void __launch_bounds__(kSearchMax)
k_leader_search(const uint16_t* __restrict__ D, int64_t c, int64_t M, int32_t L,
                int32_t* __restrict__ lead_row) {
  __shared__ uint16_t sD[kSearchMax * kSearchMax];
  __shared__ uint32_t best[kSearchMax / 64];
  const int j = threadIdx.x, lane = j & 63, wave = j >> 6;
  for (int64_t e = j; e < c * c; e += kSearchMax) sD[e] = D[e];
  __syncthreads();
  uint32_t near = 0;
  int lead = 0;
  for (int32_t it = 0; it < L; ++it) {
    if (j == 0) lead_row[it] = (int32_t)cand_row(lead, M, c);
    if (j < c) {
      const uint32_t d = sD[lead * c + j];
      near = (it == 0 || d < near) ? d : near;
    }
    // argmax of near, lowest candidate on ties: max of (near << 16 | ~j)
    uint32_t key = j < c ? (near << 16) | (uint32_t)(0xFFFF - j) : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t v = (uint32_t)__shfl_xor((int)key, o, 64);
      key = v > key ? v : key;
    }
    if (lane == 0) best[wave] = key;
    __syncthreads();
    uint32_t m = best[0];
#pragma unroll
    for (int q = 1; q < kSearchMax / 64; ++q) m = best[q] > m ? best[q] : m;
    lead = 0xFFFF - (int)(m & 0xFFFF);
    __syncthreads();  // best[] is rewritten next iteration
  }
}

// More leaders than the search takes: leader i is candidate i.
__global__ void __launch_bounds__(256)
k_stride_leaders(int64_t M, int32_t L, int32_t* __restrict__ lead_row) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < L) lead_row[i] = (int32_t)cand_row(i, M, L);
}

// Block max of a 64-bit key folded into *slot (one atomicMax per block).
__device__ __forceinline__ void block_max_to(unsigned long long key, unsigned long long* slot) {
  __shared__ unsigned long long best[4];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long v = (unsigned long long)__shfl_xor((long long)key, o, 64);
    key = v > key ? v : key;
  }
  if (lane == 0) best[wave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = best[0];
    for (int i = 1; i < 4; ++i) m = best[i] > m ? best[i] : m;
    if (m) atomicMax(slot, m);
  }
}

// Nearest of the first Lc leaders for every row (full Hamming distance,
// earliest leader on ties): a wave per row, a lane per leader of each group of
// 64; LT holds the leaders' words transposed (LT[w * Lc + i]).  The row's
// distance goes to dist, and the block's max of (distance, ~row) to *far —
// the first exact round's leader.
__global__ void __launch_bounds__(256)
k_assign(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
         const uint32_t* __restrict__ LT, int32_t Lc, int32_t* __restrict__ owner,
         int32_t* __restrict__ dist, unsigned long long* __restrict__ far) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t wpr = words_for(K);
  unsigned long long far_key = 0ull;
  if (r < M) {
    uint64_t bestkey = ~0ull;
    for (int32_t g = 0; g < Lc; g += 64) {
      const int32_t i = g + lane;
      int64_t d = 0;
      if (i < Lc)
        for (int64_t w = 0; w < wpr; ++w)
          d += __popc((W[r * ldw + w] & valid_bits(w, K)) ^ LT[w * Lc + i]);
      uint64_t key = i < Lc ? ((uint64_t)d << 32) | (uint32_t)i : ~0ull;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        const uint64_t v = (uint64_t)__shfl_xor((long long)key, o, 64);
        key = v < key ? v : key;
      }
      bestkey = key < bestkey ? key : bestkey;
    }
    const uint32_t d = (uint32_t)(bestkey >> 32);
    if (lane == 0) {
      owner[r] = (int32_t)(uint32_t)bestkey;
      dist[r] = (int32_t)d;
    }
    far_key = ((unsigned long long)d << 32) | (uint32_t)~(uint32_t)r;
  }
  block_max_to(far_key, far);
}

// Exact round `it`: leader Lc + it is the row recorded in far[it] (the row
// farthest from its nearest leader, lowest row on ties); every row takes its
// distance to it (strictly nearer: the earlier leader keeps ties) and the
// block's max of (distance, ~row) goes to far[it + 1].
__global__ void __launch_bounds__(256)
k_exact_round(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw, int32_t Lc,
              int32_t it, unsigned long long* __restrict__ far, int32_t* __restrict__ owner,
              int32_t* __restrict__ dist) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t wpr = words_for(K);
  const int64_t lead = (int64_t)(uint32_t)~(uint32_t)far[it];
  unsigned long long far_key = 0ull;
  if (r < M) {
    const int32_t prev = dist[r];
    int32_t d = 0;
    for (int64_t w = lane; w < wpr; w += 64)
      d += __popc((W[r * ldw + w] ^ W[lead * ldw + w]) & valid_bits(w, K));
    d = wave_sum(d);
    const bool nearer = d < prev;
    if (lane == 0 && nearer) {
      dist[r] = d;
      owner[r] = Lc + it;
    }
    const uint32_t dd = (uint32_t)(nearer ? d : prev);
    far_key = ((unsigned long long)dd << 32) | (uint32_t)~(uint32_t)r;
  }
  block_max_to(far_key, far + it + 1);
}

__global__ void __launch_bounds__(256)
k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v) {
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)gridDim.x * 256)
    p[e] = v;
}

// Stable ranking of a 1024-row block by leader: bitonic sort of (leader << 10
// | row in block) in LDS; a row's rank is its sorted position minus the start
// of its leader's run, and each run's length is the block's count of that
// leader: hist[leader * nb + block].  hist is zero elsewhere (filled before).
__global__ void __launch_bounds__(kRankBlock)
k_block_rank(const int32_t* __restrict__ owner, int64_t M, int64_t nb, int32_t* __restrict__ lrank,
             int32_t* __restrict__ hist) {
  __shared__ uint32_t sk[kRankBlock];
  __shared__ int32_t ss[kRankBlock];
  const int t = threadIdx.x;
  const int64_t b = blockIdx.x;
  const int64_t r = b * kRankBlock + t;
  const uint32_t key = r < M ? (uint32_t)owner[r] : (uint32_t)kPadKey;
  sk[t] = (key << 10) | (uint32_t)t;
  __syncthreads();
  for (int k = 2; k <= kRankBlock; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      const int p = t ^ j;
      if (p > t) {
        const uint32_t x = sk[t], y = sk[p];
        const bool up = (t & k) == 0;
        if ((x > y) == up) {
          sk[t] = y;
          sk[p] = x;
        }
      }
      __syncthreads();
    }
  }
  // start of each position's run: inclusive max-scan of run starts
  const uint32_t e = sk[t];
  const bool head = t == 0 || (sk[t - 1] >> 10) != (e >> 10);
  ss[t] = head ? t : 0;
  __syncthreads();
  for (int o = 1; o < kRankBlock; o <<= 1) {
    const int32_t v = t >= o ? ss[t - o] : 0;
    __syncthreads();
    ss[t] = v > ss[t] ? v : ss[t];
    __syncthreads();
  }
  const uint32_t k_e = e >> 10;
  const int src = (int)(e & 1023u);
  const int64_t row = b * kRankBlock + src;
  if (row < M) lrank[row] = t - ss[t];
  const bool tail = t == kRankBlock - 1 || (sk[t + 1] >> 10) != k_e;
  if (tail && k_e != (uint32_t)kPadKey) hist[(int64_t)k_e * nb + b] = t - ss[t] + 1;
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

// A thread per chunk: the two longest runs of equal leader among the chunk's
// slots (sorted keys), the earlier run first on equal lengths.
__global__ void __launch_bounds__(256)
k_chunk_runs(const int* __restrict__ keys, int64_t M, int4* __restrict__ runs) {
  const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= ceil_div(M, kChunk)) return;
  const int n = (int)std::min<int64_t>(kChunk, M - kChunk * c);
  const int* k = keys + kChunk * c;
  int s0 = 0, n0 = 0, s1 = 0, n1 = 0;
  for (int i = 0; i < n;) {
    int j = i + 1;
    while (j < n && k[j] == k[i]) ++j;
    const int len = j - i;
    if (len > n0) {
      s1 = s0, n1 = n0;
      s0 = i, n0 = len;
    } else if (len > n1) {
      s1 = i, n1 = len;
    }
    i = j;
  }
  runs[c] = make_int4(s0, n0, s1, n1);
}

// A thread per (chunk, base, word): majority bits of the base's run.
__global__ void __launch_bounds__(256)
k_chunk_mask(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const int* __restrict__ perm, const int4* __restrict__ runs,
             uint32_t* __restrict__ CM) {
  const int64_t wpr = words_for(K);
  const int64_t chunks = ceil_div(M, kChunk);
  const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= 2 * chunks * ldw) return;
  const int64_t cb = t / ldw, w = t - cb * ldw;
  const int64_t c = cb >> 1;
  const int4 rn = runs[c];
  const int start = (cb & 1) ? rn.z : rn.x, n = (cb & 1) ? rn.w : rn.y;
  if (w >= wpr || n == 0) {
    CM[t] = 0u;
    return;
  }
  const uint32_t vb = valid_bits(w, K);
  uint32_t cnt[32];
#pragma unroll
  for (int b = 0; b < 32; ++b) cnt[b] = 0;
  for (int i = 0; i < n; ++i) {
    const uint32_t x = W[(int64_t)perm[kChunk * c + start + i] * ldw + w] & vb;
#pragma unroll
    for (int b = 0; b < 32; ++b) cnt[b] += (x >> b) & 1u;
  }
  uint32_t out = 0;
#pragma unroll
  for (int b = 0; b < 32; ++b) out |= (2 * cnt[b] > (uint32_t)n + 1 ? 1u : 0u) << b;
  CM[t] = out;
}

// A wave per slot: the row, flagged with the base whose difference from the
// row has the fewest bits when that is fewer than the row's own.
__global__ void __launch_bounds__(256)
k_slot_flags(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const int* __restrict__ perm, const uint32_t* __restrict__ CM,
             int32_t* __restrict__ slot_row) {
  const int lane = threadIdx.x & 63;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= M) return;
  const int64_t wpr = words_for(K);
  const int r = perm[s];
  const uint32_t* a = W + (int64_t)r * ldw;
  const uint32_t* m0 = CM + (s / kChunk) * 2 * ldw;
  const uint32_t* m1 = m0 + ldw;
  int pa = 0, p0 = 0, p1 = 0;
  for (int64_t w = lane; w < wpr; w += 64) {
    const uint32_t x = a[w] & valid_bits(w, K);
    pa += __popc(x);
    p0 += __popc(x ^ m0[w]);
    p1 += __popc(x ^ m1[w]);
  }
  pa = wave_sum(pa);
  p0 = wave_sum(p0);
  p1 = wave_sum(p1);
  int32_t flag = 0;
  int best = pa;
  if (p0 < best) flag = BWS_SLOT_DELTA, best = p0;
  if (p1 < best) flag = BWS_SLOT_DELTA | BWS_SLOT_BASE1;
  if (lane == 0) slot_row[s] = r | flag;
}

struct ClusterWs {
  uint32_t* candT;
  uint16_t* D;
  int32_t* lead_row;
  uint32_t* leadT;
  int32_t* owner;
  int32_t* dist;
  unsigned long long* far;
  int32_t* lrank;
  int32_t* hist;
  int32_t* off;
  void* scan_ws;
  int32_t* keys_out;
  int32_t* perm;
  int4* runs;
  int64_t c, nb;
  size_t total;
};

ClusterWs carve_cluster(void* base, int64_t M, int64_t K, int L) {
  ClusterWs w{};
  const int64_t wpr = words_for(K);
  w.c = candidates_for(M, L);
  w.nb = ceil_div(M, kRankBlock);
  const int32_t Lc = L - exact_rounds(L);
  const bool search = Lc <= kSearchMax;
  Carve cv(base);
  w.candT = cv.take<uint32_t>(search ? (size_t)(wpr * w.c) : 0);
  w.D = cv.take<uint16_t>(search ? (size_t)(w.c * w.c) : 0);
  w.lead_row = cv.take<int32_t>((size_t)L);
  w.leadT = cv.take<uint32_t>((size_t)(wpr * Lc));
  w.owner = cv.take<int32_t>((size_t)M);
  w.dist = cv.take<int32_t>((size_t)M);
  w.far = cv.take<unsigned long long>((size_t)kExactRounds + 1);
  w.lrank = cv.take<int32_t>((size_t)M);
  w.hist = cv.take<int32_t>((size_t)(L * w.nb));
  w.off = cv.take<int32_t>((size_t)(L * w.nb) + 1);
  w.scan_ws = cv.take<uint8_t>(scan_workspace_bytes(L * w.nb));
  w.keys_out = cv.take<int32_t>((size_t)M);
  w.perm = cv.take<int32_t>((size_t)M);
  w.runs = cv.take<int4>((size_t)ceil_div(M, kChunk));
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
  return carve_cluster(nullptr, rows, cols, L).total + 256;
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
  const ClusterWs w = carve_cluster(workspace, M, K, L);
  const int64_t wpr = words_for(K);
  auto grid = [](int64_t n, int per) { return dim3((unsigned)std::max<int64_t>(1, ceil_div(n, per))); };
  // 1. leaders: the first Lc by the candidate search (or the candidates)
  const int32_t R = exact_rounds(L), Lc = L - R;
  if (Lc <= kSearchMax) {
    if (wpr > 0) {
      hipLaunchKernelGGL(k_rows_t, grid(wpr * w.c, 256), dim3(256), 0, s, A->words, M, K, A->ldw,
                         nullptr, w.c, w.candT);
      BWS_TRY(launch_check());
    }
    hipLaunchKernelGGL(k_cand_dist, grid(w.c * ceil_div(w.c, 64), 4), dim3(256), 0, s, w.candT,
                       wpr, w.c, w.D);
    BWS_TRY(launch_check());
    hipLaunchKernelGGL(k_leader_search, dim3(1), dim3(kSearchMax), 0, s, w.D, w.c, M, Lc,
                       w.lead_row);
  } else {
    hipLaunchKernelGGL(k_stride_leaders, grid(Lc, 256), dim3(256), 0, s, M, Lc, w.lead_row);
  }
  BWS_TRY(launch_check());
  // 2. nearest of those leaders, then the exact rounds over all rows
  if (wpr > 0) {
    hipLaunchKernelGGL(k_rows_t, grid(wpr * Lc, 256), dim3(256), 0, s, A->words, M, K, A->ldw,
                       w.lead_row, (int64_t)Lc, w.leadT);
    BWS_TRY(launch_check());
  }
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, s, reinterpret_cast<int32_t*>(w.far),
                     (int64_t)2 * (kExactRounds + 1), 0);
  BWS_TRY(launch_check());
  hipLaunchKernelGGL(k_assign, grid(M, 4), dim3(256), 0, s, A->words, M, K, A->ldw, w.leadT, Lc,
                     w.owner, w.dist, w.far);
  BWS_TRY(launch_check());
  for (int32_t it = 0; it < R; ++it) {
    hipLaunchKernelGGL(k_exact_round, grid(M, 4), dim3(256), 0, s, A->words, M, K, A->ldw, Lc, it,
                       w.far, w.owner, w.dist);
    BWS_TRY(launch_check());
  }
  // 3. stable order by leader
  const int64_t pairs = (int64_t)L * w.nb;
  hipLaunchKernelGGL(k_fill_i32, grid(std::min<int64_t>(pairs, 1 << 24), 256), dim3(256), 0, s,
                     w.hist, pairs, 0);
  BWS_TRY(launch_check());
  hipLaunchKernelGGL(k_block_rank, dim3((unsigned)w.nb), dim3(kRankBlock), 0, s, w.owner, M, w.nb,
                     w.lrank, w.hist);
  BWS_TRY(launch_check());
  BWS_TRY(launch_exclusive_scan_i32(w.hist, pairs, w.off, w.scan_ws, s));
  hipLaunchKernelGGL(k_place, grid(M, 256), dim3(256), 0, s, w.owner, w.lrank, w.off, M, w.nb,
                     w.perm, w.keys_out);
  BWS_TRY(launch_check());
  // 4, 5. chunk runs, bases, slot flags
  const int64_t chunks = ceil_div(M, kChunk);
  hipLaunchKernelGGL(k_chunk_runs, grid(chunks, 256), dim3(256), 0, s, w.keys_out, M, w.runs);
  BWS_TRY(launch_check());
  const int64_t threads = 2 * chunks * A->ldw;
  hipLaunchKernelGGL(k_chunk_mask, grid(threads, 256), dim3(256), 0, s, A->words, M, K, A->ldw,
                     w.perm, w.runs, chunk_mask);
  BWS_TRY(launch_check());
  hipLaunchKernelGGL(k_slot_flags, grid(M, 4), dim3(256), 0, s, A->words, M, K, A->ldw, w.perm,
                     chunk_mask, slot_row);
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

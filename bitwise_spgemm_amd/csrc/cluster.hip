// cluster.hip — row clustering and the clustered FP32 0/1 SpMM (gfx950).
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
// bounds the FP32 product on gfx950 (packed FP32 is not used).
//
// Plan (device only, no host synchronisation):
//  1. farthest-point leaders: one launch per leader; a wave per row takes the
//     Hamming distance to the new leader and keeps the nearest; one 64-bit
//     atomicMax per block of (distance, ~row) picks the next leader;
//  2. rows stably sorted by nearest leader (hipCUB radix sort);
//  3. chunk runs: a thread per chunk finds its two longest runs of equal
//     leader; chunk masks: a thread per (chunk, base, word) counts the run's
//     bits;
//  4. slot flags: a wave per slot compares popcount(row ^ base_j) with
//     popcount(row).
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "mfma_common.hpp"

namespace bws {

bws_status spmm_dispatch(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                         int64_t N, float* C, int64_t ldc, float alpha, float beta,
                         int32_t precision, const bws_plan* plan, hipStream_t s);

namespace {

constexpr int kDistWaves = 8;  // rows per block of the distance kernel
constexpr int kChunk = 16;     // slots per chunk = rows per wave of the product
constexpr int kCWaves = 16;    // waves (chunks) per workgroup of the product
constexpr int kCStage = 64;    // K rows per LDS stage of the product
constexpr int kCBN = 256;      // columns per workgroup (4 per lane)
constexpr int32_t kMaxLeaders = 65536;

int32_t leaders_for(int64_t rows, int32_t max_clusters) {
  if (rows <= 0) return 0;
  int64_t L = max_clusters > 0 ? max_clusters : rows / 64;
  L = std::max<int64_t>(L, 1);
  L = std::min<int64_t>(L, rows);
  return (int32_t)std::min<int64_t>(L, kMaxLeaders);
}

int key_bits(int L) {
  int b = 1;
  while ((1 << b) < L) ++b;
  return b;
}

__global__ void __launch_bounds__(256)
k_cluster_init(unsigned long long* __restrict__ slots, int L, int* __restrict__ iota, int64_t M) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < L) slots[i] = 0ull;
  if (i < M) iota[i] = (int)i;
}

// Iteration `iter`: distance of every row to leader `iter` (row 0, or the
// argmax recorded by iteration iter-1), nearest-leader update (strict: the
// earlier leader keeps ties), and the block's max of (distance, ~row) folded
// into slots[iter] — the next leader is the farthest row, lowest on ties.
__global__ void __launch_bounds__(kDistWaves * 64)
k_cluster_iter(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
               unsigned long long* __restrict__ slots, int* __restrict__ dist,
               int* __restrict__ owner, int iter) {
  __shared__ unsigned long long best[kDistWaves];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t r = (int64_t)blockIdx.x * kDistWaves + wave;
  const int64_t wpr = words_for(K);
  // the row's current nearest distance does not depend on the new leader:
  // its load is issued first, beside the leader's
  const int prev = (r < M && iter > 0) ? dist[r] : 0;
  const int64_t lead = iter == 0 ? 0 : (int64_t)(~(uint32_t)slots[iter - 1]);
  unsigned long long key = 0ull;
  if (r < M) {
    int d = 0;
    for (int64_t w = lane; w < wpr; w += 64)
      d += __popc((W[r * ldw + w] ^ W[lead * ldw + w]) & valid_bits(w, K));
    d = wave_sum(d);
    const bool nearer = iter == 0 || d < prev;
    if (lane == 0 && nearer) {
      dist[r] = d;
      owner[r] = iter;
    }
    const int dd = nearer ? d : prev;
    key = ((unsigned long long)(uint32_t)dd << 32) | (uint32_t)(~(uint32_t)r);
  }
  if (lane == 0) best[wave] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = best[0];
    for (int i = 1; i < kDistWaves; ++i) m = best[i] > m ? best[i] : m;
    if (m) atomicMax(&slots[iter], m);
  }
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
  int* dist;
  int* owner;
  int* iota;
  int* keys_out;
  int* perm;
  int4* runs;
  unsigned long long* slots;
  void* sort_tmp;
  size_t sort_bytes;
  size_t total;
};

ClusterWs carve_cluster(void* base, int64_t M, int L) {
  ClusterWs c{};
  size_t sort_bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (const int*)nullptr, (int*)nullptr,
                                           (const int*)nullptr, (int*)nullptr, (int)M, 0,
                                           key_bits(L));
  Carve cv(base);
  c.dist = cv.take<int>((size_t)M);
  c.owner = cv.take<int>((size_t)M);
  c.iota = cv.take<int>((size_t)M);
  c.keys_out = cv.take<int>((size_t)M);
  c.perm = cv.take<int>((size_t)M);
  c.runs = cv.take<int4>((size_t)ceil_div(M, kChunk));
  c.slots = cv.take<unsigned long long>((size_t)L);
  c.sort_tmp = cv.take<uint8_t>(sort_bytes + 256);
  c.sort_bytes = sort_bytes + 256;
  c.total = cv.used;
  return c;
}

// ---------------------------------------------------------------------------
// The clustered product.  Workgroup = 16 waves x 16 slots = 256 slots x 256
// columns; B is staged 64 K rows at a time into LDS with
// global_load_lds_dwordx4 (double-buffered, 2 x 64 KiB), as in the plain FP32
// kernel.  In each wave lane l < 16 holds slot l's bitmap words (word 2s in
// lanes 0..15, word 2s+1 in lanes 32..47) and lanes 16 / 17 (48 / 49) the
// chunk's two bases'.  Per stage every lane forms its add mask (row & ~base
// for a delta row, the row otherwise; lanes 16, 17: the base itself when some
// row of the chunk uses it) and its subtract mask (base & ~row for a delta
// row), bit reversed once in VALU; the rows then walk their bits with scalar
// bit scans, each bit adding (or subtracting) one LDS row of 4 floats per
// lane.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void take_bit(uint64_t& m, int& k) {
  asm volatile("s_ff1_i32_b64 %0, %1" : "=s"(k) : "s"(m));
  asm volatile("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(k));
}

template <bool NEG>
__device__ __forceinline__ void acc4(float4& a, const float4& b) {
  if (NEG) {
    a.x -= b.x;
    a.y -= b.y;
    a.z -= b.z;
    a.w -= b.w;
  } else {
    a.x += b.x;
    a.y += b.y;
    a.z += b.z;
    a.w += b.w;
  }
}

// Adds (NEG: subtracts) LDS rows k of the n set bits of m (bit k <-> stage
// row k) in ascending k, up to four reads in flight.
template <bool NEG>
__device__ __forceinline__ void walk_bits(float4& acc, uint64_t m, int n, float4 (*cur)[64],
                                          int lane) {
  while (n >= 4) {
    int k0, k1, k2, k3;
    take_bit(m, k0);
    take_bit(m, k1);
    take_bit(m, k2);
    take_bit(m, k3);
    const float4 b0 = cur[k0][lane], b1 = cur[k1][lane], b2 = cur[k2][lane], b3 = cur[k3][lane];
    acc4<NEG>(acc, b0);
    acc4<NEG>(acc, b1);
    acc4<NEG>(acc, b2);
    acc4<NEG>(acc, b3);
    n -= 4;
  }
  if (n >= 2) {
    int k0, k1;
    take_bit(m, k0);
    take_bit(m, k1);
    const float4 b0 = cur[k0][lane], b1 = cur[k1][lane];
    acc4<NEG>(acc, b0);
    acc4<NEG>(acc, b1);
    n -= 2;
  }
  if (n) {
    int k0;
    take_bit(m, k0);
    acc4<NEG>(acc, cur[k0][lane]);
  }
}

__device__ __forceinline__ uint64_t readlane_u64(uint32_t hi, uint32_t lo, int l) {
  return (((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, l)) << 32) |
         (uint32_t)__builtin_amdgcn_readlane((int)lo, l);
}

__global__ void __launch_bounds__(kCWaves * 64)
k_spmm_cluster(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
               const int32_t* __restrict__ slot_row, const uint32_t* __restrict__ chunk_mask,
               const float* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
               int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
               int64_t per_xcd) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  __shared__ float4 sB0[kCStage][64];
  __shared__ float4 sB1[kCStage][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t n0 = cb * kCBN;
  const int64_t chunk = rb * kCWaves + wave;
  const int64_t slot0 = chunk * kChunk;
  const int64_t nst = (K + kCStage - 1) / kCStage;
  const int64_t wpr = (K + 31) / 32;
  const int rl = lane & 31;

  // slot rows: lane l < 16 (and l + 32) holds slot l's row and flags
  int row = -1;
  bool delta = false, base1 = false;
  if (rl < kChunk && slot0 + rl < M) {
    const int32_t v = slot_row[slot0 + rl];
    row = v & (BWS_SLOT_BASE1 - 1);
    delta = (v & BWS_SLOT_DELTA) != 0;
    base1 = (v & BWS_SLOT_BASE1) != 0;
  }
  // slots summed as differences from base 0 / base 1
  const uint64_t dmask0 = __ballot(delta && !base1 && lane < kChunk);
  const uint64_t dmask1 = __ballot(delta && base1 && lane < kChunk);
  const uint32_t* src = nullptr;
  if (rl < kChunk) {
    if (row >= 0) src = W + (int64_t)row * ldw;
  } else if (rl == kChunk && dmask0 != 0ull) {
    src = chunk_mask + 2 * chunk * ldw;
  } else if (rl == kChunk + 1 && dmask1 != 0ull) {
    src = chunk_mask + (2 * chunk + 1) * ldw;
  }
  const int64_t ncol = (n0 + 4 * lane + 4 <= N) ? n0 + 4 * lane : N - 4;  // clamped column
  const float* bcol = B + ncol;
  const uint32_t lds0 = lds_u32(&sB0[0][0]), lds1 = lds_u32(&sB1[0][0]);

  float4 acc[kChunk];
#pragma unroll
  for (int r = 0; r < kChunk; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);
  float4 accT0 = make_float4(0.f, 0.f, 0.f, 0.f), accT1 = accT0;

  auto issue = [&](int64_t st, uint32_t base) {
#pragma unroll
    for (int kr = 0; kr < kCStage / kCWaves; ++kr) {
      const int brow = wave + kCWaves * kr;
      int64_t k = st * kCStage + brow;
      k = k < K ? k : K - 1;
      dma16(bcol + k * ldb, base + (uint32_t)brow * 1024u);
    }
  };
  auto load_words = [&](int64_t st) -> uint32_t {
    const int64_t w = 2 * st + (lane >> 5);
    return (src && st < nst && w < wpr) ? src[w] : 0u;
  };
  auto stage_mask = [&](int64_t st) -> uint64_t {
    const int64_t kvalid = K - st * kCStage;
    return kvalid >= 64 ? ~0ull : (~0ull << (64 - kvalid));
  };

  issue(0, lds0);
  uint32_t wcur = load_words(0);
  uint32_t wnext = load_words(1);

  auto stage = [&](int64_t st, float4 (*cur)[64], uint32_t nxt_base) {
    wait_vmcnt<0>();  // this wave's loads for stage st (and its words) landed
    // re-define the already-landed word registers so that the compiler does
    // not wait on the prefetch issued below
    asm volatile("" : "+v"(wcur), "+v"(wnext));
    __syncthreads();  // everyone's; the other buffer is no longer being read
    if (st + 1 < nst) issue(st + 1, nxt_base);
    const uint32_t wn2 = load_words(st + 2);
    const uint64_t smask = stage_mask(st);

    const uint32_t wo = (uint32_t)__shfl_xor((int)wcur, 32, 64);
    const uint64_t mm = ((((uint64_t)wcur) << 32) | wo) & smask;
    const uint64_t bm0 = readlane_u64((uint32_t)(mm >> 32), (uint32_t)mm, kChunk);
    const uint64_t bm1 = readlane_u64((uint32_t)(mm >> 32), (uint32_t)mm, kChunk + 1);
    uint64_t am = mm, sm = 0ull;
    if (delta && lane < kChunk) {
      const uint64_t bm = base1 ? bm1 : bm0;
      am = mm & ~bm;
      sm = bm & ~mm;
    }
    // bit-reversed halves: bit k of (hi:lo) <-> stage row k
    const uint32_t a_lo = __brev((uint32_t)(am >> 32)), a_hi = __brev((uint32_t)am);
    const int a_n = __popc(a_lo) + __popc(a_hi);
    const uint32_t s_lo = __brev((uint32_t)(sm >> 32)), s_hi = __brev((uint32_t)sm);
    const int s_n = __popc(s_lo) + __popc(s_hi);
    const uint64_t srows = __ballot(s_n != 0);
#pragma unroll
    for (int r = 0; r < kChunk; ++r) {
      walk_bits<false>(acc[r], readlane_u64(a_hi, a_lo, r), __builtin_amdgcn_readlane(a_n, r),
                       cur, lane);
      if ((srows >> r) & 1ull)
        walk_bits<true>(acc[r], readlane_u64(s_hi, s_lo, r), __builtin_amdgcn_readlane(s_n, r),
                        cur, lane);
    }
    if (dmask0)
      walk_bits<false>(accT0, readlane_u64(a_hi, a_lo, kChunk),
                       __builtin_amdgcn_readlane(a_n, kChunk), cur, lane);
    if (dmask1)
      walk_bits<false>(accT1, readlane_u64(a_hi, a_lo, kChunk + 1),
                       __builtin_amdgcn_readlane(a_n, kChunk + 1), cur, lane);
    wcur = wnext;
    wnext = wn2;
  };

  for (int64_t st = 0; st < nst; st += 2) {
    stage(st, sB0, lds1);
    if (st + 1 < nst) stage(st + 1, sB1, lds0);
  }

  const int64_t n = n0 + 4 * lane;
  if (n >= N) return;
#pragma unroll
  for (int r = 0; r < kChunk; ++r) {
    const int rr = __builtin_amdgcn_readlane(row, r);
    if (rr < 0) continue;
    float4 v = acc[r];
    if ((dmask0 >> r) & 1ull) acc4<false>(v, accT0);
    if ((dmask1 >> r) & 1ull) acc4<false>(v, accT1);
    float4* c = reinterpret_cast<float4*>(C + (int64_t)rr * ldc + n);
    v.x *= alpha;
    v.y *= alpha;
    v.z *= alpha;
    v.w *= alpha;
    if (beta != 0.0f) {
      const float4 o = *c;
      v.x = fmaf(beta, o.x, v.x);
      v.y = fmaf(beta, o.y, v.y);
      v.z = fmaf(beta, o.z, v.z);
      v.w = fmaf(beta, o.w, v.w);
    }
    *c = v;
  }
}

}  // namespace
}  // namespace bws

using namespace bws;

extern "C" {

size_t bws_cluster_workspace_size(int64_t rows, int32_t max_clusters) {
  if (rows <= 0 || rows >= BWS_SLOT_BASE1 || max_clusters < 0) return 0;
  return carve_cluster(nullptr, rows, leaders_for(rows, max_clusters)).total + 256;
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
  const int64_t M = A->rows;
  const int L = leaders_for(M, max_clusters);
  if (!ws_ok(workspace, workspace_bytes, bws_cluster_workspace_size(M, max_clusters)))
    return BWS_EWORKSPACE;
  hipStream_t s = stream;
  ClusterWs c = carve_cluster(workspace, M, L);
  const int64_t n_init = std::max<int64_t>(M, L);
  hipLaunchKernelGGL(k_cluster_init, dim3((unsigned)ceil_div(n_init, 256)), dim3(256), 0, s,
                     c.slots, L, c.iota, M);
  BWS_TRY(launch_check());
  const unsigned blocks = (unsigned)ceil_div(M, kDistWaves);
  for (int it = 0; it < L; ++it) {
    hipLaunchKernelGGL(k_cluster_iter, dim3(blocks), dim3(kDistWaves * 64), 0, s, A->words, M,
                       A->cols, A->ldw, c.slots, c.dist, c.owner, it);
    BWS_TRY(launch_check());
  }
  size_t sb = c.sort_bytes;
  if (hipcub::DeviceRadixSort::SortPairs(c.sort_tmp, sb, c.owner, c.keys_out, c.iota, c.perm,
                                         (int)M, 0, key_bits(L), s) != hipSuccess)
    return BWS_EHIP;
  BWS_TRY(launch_check());
  const int64_t chunks = ceil_div(M, kChunk);
  hipLaunchKernelGGL(k_chunk_runs, dim3((unsigned)ceil_div(chunks, 256)), dim3(256), 0, s,
                     c.keys_out, M, c.runs);
  BWS_TRY(launch_check());
  const int64_t threads = 2 * chunks * A->ldw;
  hipLaunchKernelGGL(k_chunk_mask, dim3((unsigned)ceil_div(threads, 256)), dim3(256), 0, s,
                     A->words, M, A->cols, A->ldw, c.perm, c.runs, chunk_mask);
  BWS_TRY(launch_check());
  hipLaunchKernelGGL(k_slot_flags, dim3((unsigned)ceil_div(M, 4)), dim3(256), 0, s, A->words, M,
                     A->cols, A->ldw, c.perm, chunk_mask, slot_row);
  return launch_check();
}

bws_status bws_spmm_clustered(const bws_bitmap* A, const int32_t* slot_row,
                              const uint32_t* chunk_mask, const float* B, int64_t ldb, int64_t N,
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
  if (!slot_row || !chunk_mask) return BWS_EINVAL;
  if (N % 4 != 0 || ldb % 4 != 0 || ldc % 4 != 0 || reinterpret_cast<uintptr_t>(B) % 16 != 0 ||
      reinterpret_cast<uintptr_t>(C) % 16 != 0)
    return BWS_EUNSUPPORTED;
  const int64_t row_blocks = ceil_div(ceil_div(A->rows, kChunk), kCWaves);
  const int64_t col_blocks = ceil_div(N, kCBN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_spmm_cluster, dim3((unsigned)(8 * per_xcd)), dim3(kCWaves * 64), 0, s,
                     A->words, A->rows, A->cols, A->ldw, slot_row, chunk_mask, B, ldb, N, C, ldc,
                     alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

}  // extern "C"

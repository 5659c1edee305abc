// spmm_stream.hip — FP32 SpMM of a 0/1 bitmap with the row masks streamed
// through the scalar unit and B staged through a ring of LDS buffers (gfx950).
//
// C = alpha * A·B + beta * C: each set bit (r, k) adds row k of B to row r of C
// ("replace the multiply with an add": the FP32 add path of
// backup/layout_trans.cu:158-171 and the add-only loop of backup/sgemm.cu:428-436).
//
// Workgroup: 16 waves, 256 rows x 256 columns of C; a wave owns 16 rows (its
// "chunk"), a lane 4 columns of each (16 float4 accumulators).
//
// B: K is walked in stages of 64 rows.  A stage's 64 x 256 panel of B goes
// HBM -> LDS by global_load_lds_dwordx4 (1 KiB per wave instruction, no
// registers) into one of two ring slots while the other is summed; one
// s_barrier per stage makes every wave's copies visible and frees the slot
// the next copy overwrites.  Each slot starts with a row of zeros.
//
// Masks: the bitmap is read as a mask stream — per chunk and stage, the
// chunk's 16 row masks are contiguous 64-bit pairs, bit reversed (bit j of a
// stage mask <-> K row 64*stage + j).  A wave's masks of the next stage are
// loaded into SGPRs (s_load) at the end of a stage, so the wait for them falls
// into the barrier.  Set bits are taken with s_ff1 / s_bitset0.
//
// Per stage a wave first takes the lowest set bit of every row without a
// branch: s_ff1 of an empty mask is -1, which addresses the slot's zero row,
// so the 16 reads (four in flight) and their adds run straight through.  The
// bits left over (rows with two or more in the stage, about a third of the
// rows at 5% density) are then walked row by row.
//
// Plain product: bws_mask_stream writes the stream from the bitmap (a pre-pass
// per launch, 4 B read and 4 B written per word).  Row-clustered product: the
// stream of the row-clustering plan (bws_cluster_rows) holds, per chunk and
// stage, the slots' add masks (row & ~base for a delta row, the row
// otherwise), the chunk's two base masks (walked as two more rows into their
// own accumulators, added to their rows at the store), a flag word (bit j:
// slot j has bits to subtract in this stage), the subtract masks (base & ~row)
// of the first two flagged slots (loaded with the stage's masks) and every
// slot's subtract mask (read on the rare stages with more than two).
#include "mfma_common.hpp"

namespace bws {
namespace {

constexpr int kQW = 16;            // waves per workgroup
constexpr int kQR = 16;            // rows (slots) per wave
constexpr int kQRows = kQW * kQR;  // rows per workgroup
constexpr int kQBN = 256;          // columns per workgroup (4 per lane)

// Words of one chunk-stage record of the stream (MW = mask words per row).
template <bool CL, int MW>
struct Rec {
  static constexpr int words = CL ? 48 * MW : 16 * MW;
  static constexpr int base = 16 * MW;      // CL: base 0, base 1
  static constexpr int flags = 18 * MW;     // CL: bit j = slot j subtracts
  static constexpr int inl = 19 * MW;       // CL: subtract masks of the first two flagged slots
  static constexpr int sub = 32 * MW;       // CL: subtract masks of every slot
};

typedef const __attribute__((address_space(4))) uint32_t* sgpr_src;

// A stage mask of MW words in SGPRs.  take(m): index of the lowest set bit,
// which is cleared; -1 for an empty mask (s_ff1 of zero), which s_bitset0
// leaves empty (it clears bit 31 or 63).  count(m): set bits.
template <int MW>
struct Mask;
template <>
struct Mask<1> {
  using T = uint32_t;
  static __device__ __forceinline__ T load(sgpr_src p, int i) { return p[i]; }
  static __device__ __forceinline__ int take(T& m) {
    int k;
    asm volatile("s_ff1_i32_b32 %0, %1\n\ts_bitset0_b32 %1, %0" : "=&s"(k), "+s"(m));
    return k;
  }
  static __device__ __forceinline__ int count(T m) {
    int n;
    asm("s_bcnt1_i32_b32 %0, %1" : "=s"(n) : "s"(m));
    return n;
  }
};
template <>
struct Mask<2> {
  using T = uint64_t;
  static __device__ __forceinline__ T load(sgpr_src p, int i) {
    return ((uint64_t)p[2 * i + 1] << 32) | p[2 * i];
  }
  static __device__ __forceinline__ int take(T& m) {
    int k;
    asm volatile("s_ff1_i32_b64 %0, %1\n\ts_bitset0_b64 %1, %0" : "=&s"(k), "+s"(m));
    return k;
  }
  static __device__ __forceinline__ int count(T m) {
    int n;
    asm("s_bcnt1_i32_b64 %0, %1" : "=s"(n) : "s"(m));
    return n;
  }
};

__device__ __forceinline__ f32x4 lds_row(uint32_t vbase, int k) {
  return *reinterpret_cast<const __attribute__((address_space(3))) f32x4*>(
      (uintptr_t)(vbase + ((uint32_t)k << 10)));
}

template <bool NEG, class V>
__device__ __forceinline__ void acc_add(float4& a, const V& b) {
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

// Adds (NEG: subtracts) the LDS rows of the set bits of m to acc.
// The scalar unit issues every instruction of these walks for the 16 waves of
// a CU, so the common cases are kept short: an empty mask costs a compare and
// a branch, a one-bit mask the bit scan and one more compare and branch.  The
// row's bits are summed into t first (ascending k, up to four LDS reads in
// flight) and t is added to acc once: every path then ends in the same add, so
// the accumulators keep their registers (no copies at the joins).
template <int MW, bool NEG>
__device__ __forceinline__ void walk(float4& acc, typename Mask<MW>::T m, uint32_t vbase) {
  if (m != 0) {
    f32x4 t = lds_row(vbase, Mask<MW>::take(m));
    if (m != 0) {
      const uint32_t n = (uint32_t)Mask<MW>::count(m);
      if (n & 1u) t += lds_row(vbase, Mask<MW>::take(m));
      if (n & 2u) {
        const int k0 = Mask<MW>::take(m), k1 = Mask<MW>::take(m);
        const f32x4 b0 = lds_row(vbase, k0), b1 = lds_row(vbase, k1);
        t += b0;
        t += b1;
      }
      for (uint32_t q = n >> 2; q != 0u; --q) {
        const int k0 = Mask<MW>::take(m), k1 = Mask<MW>::take(m), k2 = Mask<MW>::take(m),
                  k3 = Mask<MW>::take(m);
        const f32x4 b0 = lds_row(vbase, k0), b1 = lds_row(vbase, k1), b2 = lds_row(vbase, k2),
                    b3 = lds_row(vbase, k3);
        t += b0;
        t += b1;
        t += b2;
        t += b3;
      }
    }
    acc_add<NEG>(acc, t);
  }
}

__device__ __forceinline__ void s_barrier() { asm volatile("s_barrier" ::: "memory"); }

// Keeps a at this point of the program (its adds are not moved into later
// branches, where every path would need its own copy of the registers).
__device__ __forceinline__ void pin(float4& a) {
  asm volatile("" : "+v"(a.x), "+v"(a.y), "+v"(a.z), "+v"(a.w));
}

// acc[r] += the LDS row of the lowest set bit of m[r], that bit cleared, for
// every r < NR without a branch (an empty mask reads the zero row at k = -1).
// Four rows' reads are in flight while the previous four rows are added.
template <int MW, int NR>
__device__ __forceinline__ void first_bits(float4 (&acc)[NR], typename Mask<MW>::T (&m)[NR],
                                           uint32_t vbase) {
  constexpr int G = (NR + 3) / 4;
  f32x4 b[4], c[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i < NR) b[i] = lds_row(vbase, Mask<MW>::take(m[i]));
#pragma unroll
  for (int g = 1; g < G; ++g) {
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (4 * g + i < NR) c[i] = lds_row(vbase, Mask<MW>::take(m[4 * g + i]));
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      acc_add<false>(acc[4 * (g - 1) + i], b[i]);
      pin(acc[4 * (g - 1) + i]);
      b[i] = c[i];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (4 * (G - 1) + i < NR) {
      acc_add<false>(acc[4 * (G - 1) + i], b[i]);
      pin(acc[4 * (G - 1) + i]);
    }
}

template <bool CL>
__global__ void __launch_bounds__(kQW * 64)
k_spmm_stream(const uint32_t* __restrict__ S, const int32_t* __restrict__ slot_row, int64_t M,
              int64_t nst, const float* __restrict__ B, int64_t ldb, int64_t K, int64_t N,
              float* __restrict__ C, int64_t ldc, float alpha, float beta, int64_t row_blocks,
              int64_t col_blocks, int64_t per_xcd) {
  constexpr int KS = 64, MW = 2;
  constexpr int RPS = KS / kQW;               // copies per wave per stage (one B row each)
  constexpr uint32_t SLOT = (KS + 1) * 1024;  // bytes per ring slot: the zero row, then KS rows
  constexpr int NR = CL ? kQR + 2 : kQR;      // CL: the two bases are rows 16 and 17
  using R = Rec<CL, MW>;
  using MT = typename Mask<MW>::T;
  __shared__ float4 sB[2][KS + 1][64];

  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t n0 = cb * kQBN;
  const int64_t chunk = rb * kQW + wave;
  const int64_t ncol = (n0 + 4 * lane + 4 <= N) ? n0 + 4 * lane : N - 4;  // clamped column
  const float* bcol = B + ncol;
  const uint32_t lds0 = lds_u32(&sB[0][0][0]);
  const sgpr_src rec = (sgpr_src)(S + chunk * nst * R::words);
  if (wave < 2) sB[wave][0][lane] = make_float4(0.f, 0.f, 0.f, 0.f);
  __syncthreads();

  // copy stage st's B rows of this wave into ring slot `slot`: B rows
  // st*KS + wave + 16j, j < RPS (rows past K read row K-1, never summed)
  const float* bw = bcol + (int64_t)wave * ldb;
  const int64_t row_step = (int64_t)kQW * ldb;  // elements between a wave's rows
  auto issue = [&](int64_t st, int slot) {
    const uint32_t m0 = lds0 + (uint32_t)slot * SLOT + 1024u + (uint32_t)wave * 1024u;
    if ((st + 1) * KS <= K) {
      const float* p = bw + st * KS * ldb;
#pragma unroll
      for (int j = 0; j < RPS; ++j) dma16(p + j * row_step, m0 + (uint32_t)(j * kQW) * 1024u);
    } else {
#pragma unroll
      for (int j = 0; j < RPS; ++j) {
        int64_t k = st * KS + wave + kQW * j;
        k = k < K ? k : K - 1;
        dma16(bcol + k * ldb, m0 + (uint32_t)(j * kQW) * 1024u);
      }
    }
  };

  float4 acc[NR];
#pragma unroll
  for (int r = 0; r < NR; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);

  MT m[NR];
  MT s0 = 0, s1 = 0;  // CL: subtract masks of the first two flagged slots
  uint32_t fl = 0;    // CL: bit j = slot j subtracts in this stage
  auto load = [&](int64_t st) {
    const sgpr_src p = rec + st * R::words;
#pragma unroll
    for (int r = 0; r < kQR; ++r) m[r] = Mask<MW>::load(p, r);
    if (CL) {
      m[NR - 2] = Mask<MW>::load(p + R::base, 0);
      m[NR - 1] = Mask<MW>::load(p + R::base, 1);
      fl = p[R::flags];
      s0 = Mask<MW>::load(p + R::inl, 0);
      s1 = Mask<MW>::load(p + R::inl, 1);
    }
  };
  issue(0, 0);
  load(0);
  int slot = 0;
  for (int64_t st = 0; st < nst; ++st) {
    wait_vmcnt<0>();  // this wave's copies of stage st have landed
    s_barrier();      // every wave's have; every wave is done with the other slot
    if (st + 1 < nst) issue(st + 1, slot ^ 1);
    const uint32_t vbase = lds0 + (uint32_t)slot * SLOT + 1024u + (uint32_t)lane * 16u;
    first_bits<MW, NR>(acc, m, vbase);
#pragma unroll
    for (int r = 0; r < NR; ++r) walk<MW, false>(acc[r], m[r], vbase);
    if (CL && fl) {
      const sgpr_src p = rec + st * R::words + R::sub;
      int i = 0;
#pragma unroll
      for (int r = 0; r < kQR; ++r)
        if ((fl >> r) & 1u) {
          MT sm = i == 0 ? s0 : s1;
          if (i >= 2) sm = Mask<MW>::load(p, r);
          walk<MW, true>(acc[r], sm, vbase);
          ++i;
        }
    }
    if (st + 1 < nst) load(st + 1);
    slot ^= 1;
  }

  // slot rows (CL) and which slots add a base, taken before any lane leaves
  int32_t srow = -1;
  if (CL) {
    const int64_t s0r = chunk * kQR;
    if (lane < kQR && s0r + lane < M) srow = slot_row[s0r + lane];
  }
  const bool d = lane < kQR && srow >= 0 && (srow & BWS_SLOT_DELTA);
  const uint32_t use0 = CL ? (uint32_t)__ballot(d && !(srow & BWS_SLOT_BASE1)) : 0u;
  const uint32_t use1 = CL ? (uint32_t)__ballot(d && (srow & BWS_SLOT_BASE1)) : 0u;
  const int64_t n = n0 + 4 * lane;
  if (n >= N) return;
#pragma unroll
  for (int r = 0; r < kQR; ++r) {
    int64_t row;
    float4 v = acc[r];
    if (CL) {
      const int32_t sr = __builtin_amdgcn_readlane(srow, r);
      if (sr < 0) continue;
      row = sr & (BWS_SLOT_BASE1 - 1);
      if ((use0 >> r) & 1u) acc_add<false>(v, acc[NR - 2]);
      if ((use1 >> r) & 1u) acc_add<false>(v, acc[NR - 1]);
    } else {
      row = chunk * kQR + r;
      if (row >= M) break;
    }
    float4* c = reinterpret_cast<float4*>(C + row * ldc + n);
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

// Subset-sum variant for dense masks (the bit-wise subset sums of
// backup/bit_wise_gemm.cu:35-120, restated for LDS): K in sub-stages of 16
// rows, 4 groups of 4.  Per sub-stage an LDS table holds, per group, the sums
// of the group's B rows over all 16 subsets (entry e = sum of rows 4g + j for
// the bits j of e, entry 0 a zero row), built from the sub-stage's B rows, so
// a row's nibble of the group costs one LDS read and four adds whatever its
// bits.  B rows come HBM -> LDS by global_load_lds_dwordx4 into a ring of RB
// sub-stages (RB - 1 ahead); the masks are the 32-row mask stream (a record
// per two sub-stages, the next one loaded while the current is used).
constexpr int kSubRows = 16;
constexpr int kSubRing = 4;
__constant__ uint8_t kComposite[11] = {3, 5, 6, 7, 9, 10, 11, 12, 13, 14, 15};

__global__ void __launch_bounds__(kQW * 64)
k_spmm_subset(const uint32_t* __restrict__ S, int64_t M, int64_t nst, const float* __restrict__ B,
              int64_t ldb, int64_t K, int64_t N, float* __restrict__ C, int64_t ldc, float alpha,
              float beta, int64_t row_blocks, int64_t col_blocks, int64_t per_xcd) {
  using R = Rec<false, 1>;
  constexpr int RB = kSubRing;
  __shared__ f32x4 raw[RB][kSubRows][64];  // 64 KiB
  __shared__ f32x4 tab[4][16][64];         // [group][subset][lane]: 64 KiB

  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t n0 = cb * kQBN;
  const int64_t chunk = rb * kQW + wave;
  const int64_t ncol = (n0 + 4 * lane + 4 <= N) ? n0 + 4 * lane : N - 4;  // clamped column
  const float* bcol = B + ncol;
  const sgpr_src rec = (sgpr_src)(S + chunk * nst * R::words);
  const int64_t nsub = ceil_div(K, kSubRows);

  if (wave < 4) tab[wave][0][lane] = f32x4{0.f, 0.f, 0.f, 0.f};
  // wave w copies row 16 s + w of sub-stage s into raw[s % RB][w]
  const uint32_t dst0 = lds_u32(&raw[0][wave][0]);
  auto issue = [&](int64_t sub) {
    int64_t k = sub * kSubRows + wave;
    k = k < K ? k : K - 1;  // rows past K meet zero mask bits
    dma16(bcol + k * ldb, dst0 + (uint32_t)(sub % RB) * (kSubRows * 1024u));
  };

  float4 acc[kQR];
#pragma unroll
  for (int r = 0; r < kQR; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);
  uint32_t mc[kQR], mn[kQR];
#pragma unroll
  for (int r = 0; r < kQR; ++r) mc[r] = rec[r];
#pragma unroll
  for (int j = 0; j < RB - 1; ++j)
    if (j < nsub) issue(j);
  for (int64_t sub = 0; sub < nsub; ++sub) {
    // this wave's copy of sub-stage sub has landed (younger ones may be in flight)
    const int64_t younger = nsub - 1 - sub;
    if (younger >= RB - 2) wait_vmcnt<RB - 2>();
    else if (younger == 1) wait_vmcnt<1>();
    else wait_vmcnt<0>();
    s_barrier();  // every copy of sub landed; the table of sub - 1 is no longer read
    if (sub + RB - 1 < nsub) issue(sub + RB - 1);
    if ((sub & 1) == 0 && (sub >> 1) + 1 < nst) {  // the next record's masks
      const sgpr_src p = rec + ((sub >> 1) + 1) * R::words;
#pragma unroll
      for (int r = 0; r < kQR; ++r) mn[r] = p[r];
    }
    // the table of this sub-stage: singles copied, composites summed
    const int slot = (int)(sub % RB);
    for (int t = wave; t < 60; t += kQW) {
      const int g = t / 15, e = t - 15 * g;  // e: 0..3 singles (1 << e), 4..14 composites
      f32x4 v;
      int ent;
      if (e < 4) {
        ent = 1 << e;
        v = raw[slot][4 * g + e][lane];
      } else {
        ent = kComposite[e - 4];
        int rest = ent;
        v = raw[slot][4 * g + __builtin_ctz(rest)][lane];
        rest &= rest - 1;
        while (rest) {
          v += raw[slot][4 * g + __builtin_ctz(rest)][lane];
          rest &= rest - 1;
        }
      }
      tab[g][ent][lane] = v;
    }
    asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");  // table complete
    const int sh = (int)(sub & 1) * kSubRows;
#pragma unroll
    for (int r = 0; r < kQR; ++r) {
      const uint32_t m16 = (mc[r] >> sh) & 0xFFFFu;
      if (m16 != 0u) {
#pragma unroll
        for (int g = 0; g < 4; ++g) acc_add<false>(acc[r], tab[g][(m16 >> (4 * g)) & 15u][lane]);
      }
    }
    if (sub & 1) {
#pragma unroll
      for (int r = 0; r < kQR; ++r) mc[r] = mn[r];
    }
  }

  const int64_t n = n0 + 4 * lane;
  if (n >= N) return;
#pragma unroll
  for (int r = 0; r < kQR; ++r) {
    const int64_t row = chunk * kQR + r;
    if (row >= M) break;
    float4* c = reinterpret_cast<float4*>(C + row * ldc + n);
    float4 v = acc[r];
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

// Bits of word w (LSB-first after reversal) that are columns < cols.
__device__ __forceinline__ uint32_t valid_lsb(int64_t w, int64_t cols) {
  const int64_t rem = cols - 32 * w;
  return rem >= 32 ? 0xFFFFFFFFu : (rem <= 0 ? 0u : ((1u << rem) - 1u));
}

// Mask stream of a plain bitmap: a thread per (row, group of 8 stages).  Rows
// beyond M and bits beyond K are written as zero.
template <int MW>
__global__ void __launch_bounds__(256)
k_mask_stream(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw, int64_t nst,
              uint32_t* __restrict__ S) {
  const int t = threadIdx.x;
  const int64_t r = (int64_t)blockIdx.y * kQRows + t;
  const int64_t chunk = r / kQR;
  const int in = t % kQR;
  const int64_t wpr = words_for(K);
  for (int j = 0; j < 8; ++j) {
    const int64_t st = (int64_t)blockIdx.x * 8 + j;
    if (st >= nst) break;
    uint32_t* out = S + (chunk * nst + st) * (16 * MW) + in * MW;
#pragma unroll
    for (int h = 0; h < MW; ++h) {
      const int64_t w = st * MW + h;
      uint32_t x = 0;
      if (r < M && w < wpr) x = __brev(W[r * ldw + w]) & valid_lsb(w, K);
      out[h] = x;
    }
  }
}

// Mask stream of a row-clustering plan: 16 threads per (chunk, stage)
// record, thread j for slot j (its add and subtract masks); the record's
// bases, flag word (bit j: slot j subtracts in this stage), the first two
// flagged slots' subtract masks and zero words are written by its threads
// together, the chunk's base use and the flags gathered with wave ballots
// (four records per wave).
template <int MW>
__global__ void __launch_bounds__(256)
k_cluster_stream(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const int32_t* __restrict__ slot_row, const uint32_t* __restrict__ CM,
                 int64_t chunks, int64_t nst, uint32_t* __restrict__ S) {
  using R = Rec<true, MW>;
  const int64_t rec = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 4;
  const int j = threadIdx.x & 15;
  const int sh = threadIdx.x & 48;  // first lane of the record within the wave
  const bool live = rec < chunks * nst;
  const int64_t c = live ? rec / nst : 0, st = live ? rec - c * nst : 0;
  const int64_t wpr = words_for(K);
  const bool real = live && c * kQR < M;  // chunks past the last row block's end are all zero
  const int64_t s = c * kQR + j;
  const int32_t sr = real && s < M ? slot_row[s] : -1;
  const bool delta = sr >= 0 && (sr & BWS_SLOT_DELTA);
  const bool one = (sr & BWS_SLOT_BASE1) != 0;
  // which bases some slot of the chunk uses
  const bool used0 = ((__ballot(delta && !one) >> sh) & 0xFFFFull) != 0;
  const bool used1 = ((__ballot(delta && one) >> sh) & 0xFFFFull) != 0;
  uint32_t* out = S + rec * R::words;
  uint32_t any_sub = 0;
  uint32_t subs[MW];
#pragma unroll
  for (int h = 0; h < MW; ++h) {
    const int64_t w = st * MW + h;
    const bool ok = real && w < wpr;
    const uint32_t vb = ok ? valid_lsb(w, K) : 0u;
    const uint32_t b0 = ok ? __brev(CM[(2 * c) * ldw + w]) & vb : 0u;
    const uint32_t b1 = ok ? __brev(CM[(2 * c + 1) * ldw + w]) & vb : 0u;
    uint32_t add = 0, sub = 0;
    if (sr >= 0 && ok) {
      const uint32_t x = __brev(W[(int64_t)(sr & (BWS_SLOT_BASE1 - 1)) * ldw + w]) & vb;
      if (delta) {
        const uint32_t b = one ? b1 : b0;
        add = x & ~b;
        sub = b & ~x;
      } else {
        add = x;
      }
    }
    if (live) {
      out[j * MW + h] = add;
      out[R::sub + j * MW + h] = sub;
      if (j == 0) {
        out[R::base + h] = used0 ? b0 : 0u;
        out[R::base + MW + h] = used1 ? b1 : 0u;
      }
    }
    subs[h] = sub;
    any_sub |= sub;
  }
  const uint32_t flags = (uint32_t)((__ballot(any_sub != 0u) >> sh) & 0xFFFFull);
  if (live) {
    if (j == 0) out[R::flags] = flags;
    // the first two flagged slots' subtract masks again, in slot order, after the flags
    const int rank = __popc(flags & ((1u << j) - 1u));
    if (any_sub != 0u && rank < 2) {
#pragma unroll
      for (int h = 0; h < MW; ++h) out[R::inl + rank * MW + h] = subs[h];
    }
    if (j == 0)
      for (int q = __popc(flags); q < 2; ++q)
#pragma unroll
        for (int h = 0; h < MW; ++h) out[R::inl + q * MW + h] = 0u;
    for (int i = R::flags + 1 + j; i < R::sub; i += 16)
      if (i < R::inl || i >= R::inl + 2 * MW) out[i] = 0u;
  }
}

template <int KS>
size_t mask_stream_bytes_t(int64_t M, int64_t K) {
  return (size_t)ceil_div(M, kQRows) * kQW * ceil_div(K, KS) * Rec<false, KS / 32>::words * 4;
}
// The cluster stream covers whole row blocks (16 chunks each), as the product
// reads them; chunks past the rows are zero.
template <int KS>
size_t cluster_stream_bytes_t(int64_t M, int64_t K) {
  return (size_t)ceil_div(M, kQRows) * kQW * ceil_div(K, KS) * Rec<true, KS / 32>::words * 4;
}

template <int KS>
bws_status launch_mask_stream_t(const bws_bitmap* A, uint32_t* S, hipStream_t s) {
  const int64_t nst = ceil_div(A->cols, KS);
  const int64_t rbs = ceil_div(A->rows, kQRows);
  if (rbs > 65535 || ceil_div(nst, 8) > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_mask_stream<KS / 32>, dim3((unsigned)ceil_div(nst, 8), (unsigned)rbs),
                     dim3(256), 0, s, A->words, A->rows, A->cols, A->ldw, nst, S);
  return launch_check();
}

template <int KS>
bws_status launch_cluster_stream_t(const bws_bitmap* A, const int32_t* slot_row,
                                   const uint32_t* chunk_mask, uint32_t* S, hipStream_t s) {
  const int64_t nst = ceil_div(A->cols, KS);
  const int64_t chunks = ceil_div(A->rows, kQRows) * kQW;
  const int64_t n = chunks * nst * 16;  // 16 threads per (chunk, stage) record
  if (ceil_div(n, 256) > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_cluster_stream<KS / 32>, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, slot_row, chunk_mask, chunks, nst, S);
  return launch_check();
}

// The streamed FP32 product over a stream of launch_mask_stream_t<64> (slot_row
// NULL) or launch_cluster_stream_t<64> (slot_row = the plan's).  N, ldb and
// ldc multiples of 4 and B, C 16-byte aligned (checked by the callers).
bws_status launch_spmm_stream_t(const uint32_t* S, const int32_t* slot_row, int64_t M, int64_t K,
                                const float* B, int64_t ldb, int64_t N, float* C, int64_t ldc,
                                float alpha, float beta, hipStream_t s) {
  const int64_t nst = ceil_div(K, 64);
  const int64_t row_blocks = ceil_div(M, kQRows);
  const int64_t col_blocks = ceil_div(N, kQBN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  if (slot_row)
    hipLaunchKernelGGL(k_spmm_stream<true>, dim3((unsigned)(8 * per_xcd)), dim3(kQW * 64), 0, s,
                       S, slot_row, M, nst, B, ldb, K, N, C, ldc, alpha, beta, row_blocks,
                       col_blocks, per_xcd);
  else
    hipLaunchKernelGGL(k_spmm_stream<false>, dim3((unsigned)(8 * per_xcd)), dim3(kQW * 64), 0, s,
                       S, slot_row, M, nst, B, ldb, K, N, C, ldc, alpha, beta, row_blocks,
                       col_blocks, per_xcd);
  return launch_check();
}

// The subset-sum product over the 32-row mask stream of launch_mask_stream_t<32>.
bws_status launch_spmm_subset_t(const uint32_t* S, int64_t M, int64_t K, const float* B,
                                int64_t ldb, int64_t N, float* C, int64_t ldc, float alpha,
                                float beta, hipStream_t s) {
  const int64_t nst = ceil_div(K, 32);
  const int64_t row_blocks = ceil_div(M, kQRows);
  const int64_t col_blocks = ceil_div(N, kQBN);
  const int64_t per_xcd = ceil_div(row_blocks * col_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_spmm_subset, dim3((unsigned)(8 * per_xcd)), dim3(kQW * 64), 0, s, S, M,
                     nst, B, ldb, K, N, C, ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

}  // namespace

#ifndef BWS_STREAM_NO_PRODUCT
// K rows per stage of the streamed product.
constexpr int kStreamKS = 64;

size_t mask_stream_bytes(int64_t M, int64_t K) { return mask_stream_bytes_t<kStreamKS>(M, K); }
size_t cluster_stream_bytes(int64_t M, int64_t K) {
  return cluster_stream_bytes_t<kStreamKS>(M, K);
}
bws_status launch_mask_stream(const bws_bitmap* A, uint32_t* S, hipStream_t s) {
  return launch_mask_stream_t<kStreamKS>(A, S, s);
}
bws_status launch_cluster_stream(const bws_bitmap* A, const int32_t* slot_row,
                                 const uint32_t* chunk_mask, uint32_t* S, hipStream_t s) {
  return launch_cluster_stream_t<kStreamKS>(A, slot_row, chunk_mask, S, s);
}
bws_status launch_spmm_stream(const uint32_t* S, const int32_t* slot_row, int64_t M, int64_t K,
                              const float* B, int64_t ldb, int64_t N, float* C, int64_t ldc,
                              float alpha, float beta, hipStream_t s) {
  return launch_spmm_stream_t(S, slot_row, M, K, B, ldb, N, C, ldc, alpha, beta, s);
}
size_t subset_stream_bytes(int64_t M, int64_t K) { return mask_stream_bytes_t<32>(M, K); }
bws_status launch_spmm_subset(const bws_bitmap* A, uint32_t* S, const float* B, int64_t ldb,
                              int64_t N, float* C, int64_t ldc, float alpha, float beta,
                              hipStream_t s) {
  BWS_TRY(launch_mask_stream_t<32>(A, S, s));
  return launch_spmm_subset_t(S, A->rows, A->cols, B, ldb, N, C, ldc, alpha, beta, s);
}
#endif

}  // namespace bws

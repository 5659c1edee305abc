// spmm_valu.hip — FP32 vector-add tile GEMM: C = alpha * A·B + beta * C with a
// bitmap A (0/1 or value-carrying) and a dense B, for gfx950.
//
// Tile: a workgroup of 8 waves owns 8*RPW rows x 256 columns of C.  K is
// walked one 32-column word at a time: the matching 32 x 256 panel of B is
// staged in LDS (double-buffered, coalesced 16-byte global loads), and every
// wave walks the set bits of its rows' words with wave-uniform scalar bit
// scans (s_flbit / s_andn2), so there is no divergence: each set bit (r, k)
// adds LDS row k (4 floats per lane) into row r's accumulators ("replace the
// multiply with an add", backup/sgemm.cu:428-436, backup/layout_trans.cu:158-171).
// Workgroups are mapped XCD-aware: consecutive tiles of one 256-column panel
// go to the same XCD, so that panel of B is fetched from HBM about once per
// XCD and re-read from that XCD's L2.
#include "bws_internal.hpp"

namespace bws {
namespace {

constexpr int kWaves = 8;
constexpr int kThreads = kWaves * 64;
constexpr int kBN = 256;  // columns per workgroup (4 per lane)
constexpr int kKC = 32;   // K rows per stage (one bitmap word)

// B element types: 0 = f32, 1 = bf16 (stored as uint16), 2 = f16.
template <int TB>
struct BType;
template <>
struct BType<0> {
  using T = float;
  static __device__ __forceinline__ float cvt(T x) { return x; }
};
template <>
struct BType<1> {
  using T = uint16_t;
  static __device__ __forceinline__ float cvt(T x) { return __uint_as_float((uint32_t)x << 16); }
};
template <>
struct BType<2> {
  using T = _Float16;
  static __device__ __forceinline__ float cvt(T x) { return (float)x; }
};

template <int TB, bool VEC>
__device__ __forceinline__ float4 load_b4(const typename BType<TB>::T* __restrict__ B, int64_t k,
                                          int64_t n, int64_t ldb, int64_t K, int64_t N) {
  using T = typename BType<TB>::T;
  float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
  if (k >= K) return v;
  const T* p = B + k * ldb + n;
  if (VEC && n + 3 < N) {
    if (TB == 0) {
      v = *reinterpret_cast<const float4*>(p);
    } else {
      const uint2 u = *reinterpret_cast<const uint2*>(p);
      T e[4];
      __builtin_memcpy(e, &u, 8);
      v = make_float4(BType<TB>::cvt(e[0]), BType<TB>::cvt(e[1]), BType<TB>::cvt(e[2]),
                      BType<TB>::cvt(e[3]));
    }
  } else {
    if (n + 0 < N) v.x = BType<TB>::cvt(p[0]);
    if (n + 1 < N) v.y = BType<TB>::cvt(p[1]);
    if (n + 2 < N) v.z = BType<TB>::cvt(p[2]);
    if (n + 3 < N) v.w = BType<TB>::cvt(p[3]);
  }
  return v;
}

__device__ __forceinline__ float readlane_f(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int64_t readlane_i64(int64_t v, int l) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)((uint64_t)v >> 32), l);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ void fma4(float4& acc, float a, const float4& b) {
  acc.x = fmaf(a, b.x, acc.x);
  acc.y = fmaf(a, b.y, acc.y);
  acc.z = fmaf(a, b.z, acc.z);
  acc.w = fmaf(a, b.w, acc.w);
}
__device__ __forceinline__ void add4(float4& acc, const float4& b) {
  acc.x += b.x;
  acc.y += b.y;
  acc.z += b.z;
  acc.w += b.w;
}

// Tile index -> (row block, column block), XCD-aware: workgroup g runs on
// XCD g % 8; XCD x receives the contiguous tile range [x*per_xcd, (x+1)*per_xcd)
// of the column-block-major tile order.
__device__ __forceinline__ int64_t xcd_tile(int64_t g, int64_t per_xcd) {
  return (g & 7) * per_xcd + (g >> 3);
}

template <int TB, int RPW, bool VALS, bool VEC>
__global__ void __launch_bounds__(kThreads)
k_spmm_rows(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
            const int64_t* __restrict__ row_ptr, const float* __restrict__ values,
            const typename BType<TB>::T* __restrict__ B, int64_t ldb, int64_t N,
            float* __restrict__ C, int64_t ldc, float alpha, float beta, int64_t row_blocks,
            int64_t col_blocks, int64_t per_xcd) {
  __shared__ float4 sB[2][kKC][64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t n0 = cb * kBN;
  const int64_t mw = rb * (kWaves * RPW) + (int64_t)wave * RPW;
  const int64_t nk = (K + kKC - 1) / kKC;
  const int krem = (int)(K % kKC);
  const uint32_t last_mask = krem ? (0xFFFFFFFFu << (32 - krem)) : 0xFFFFFFFFu;
  const bool row_lane = lane < RPW && mw + lane < M;

  float4 acc[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);

  float4 st[kKC * 64 / kThreads];
  auto load_chunk = [&](int64_t kc) {
#pragma unroll
    for (int s = 0; s < kKC * 64 / kThreads; ++s) {
      const int idx = threadIdx.x + kThreads * s;
      st[s] = load_b4<TB, VEC>(B, kc * kKC + (idx >> 6), n0 + 4 * (idx & 63), ldb, K, N);
    }
  };
  auto store_chunk = [&](int buf) {
#pragma unroll
    for (int s = 0; s < kKC * 64 / kThreads; ++s) {
      const int idx = threadIdx.x + kThreads * s;
      sB[buf][idx >> 6][idx & 63] = st[s];
    }
  };

  load_chunk(0);
  store_chunk(0);
  uint32_t wcur = row_lane ? W[(mw + lane) * ldw] : 0u;
  int64_t e_l = 0;
  if (VALS) e_l = row_lane ? row_ptr[mw + lane] : 0;
  __syncthreads();

  for (int64_t kc = 0; kc < nk; ++kc) {
    const int buf = (int)(kc & 1);
    const bool more = kc + 1 < nk;
    if (more) load_chunk(kc + 1);
    const uint32_t wnext = (more && row_lane) ? W[(mw + lane) * ldw + kc + 1] : 0u;
    const uint32_t wc = more ? wcur : (wcur & last_mask);

    float vv[VALS ? RPW : 1];
    if (VALS) {
      const int cnt_l = __popc(wc);
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        const int64_t e = readlane_i64(e_l, r);
        const int c = __builtin_amdgcn_readlane(cnt_l, r);
        vv[r] = lane < c ? values[e + lane] : 0.0f;
      }
      e_l += cnt_l;
    }

#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      uint32_t m = (uint32_t)__builtin_amdgcn_readlane((int)wc, r);
      int j = 0;
      while (m) {
        const int i = __clz((int)m);
        m &= ~wbit(i);
        const float4 b = sB[buf][i][lane];
        if (VALS) {
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], j), b);
          ++j;
        } else {
          add4(acc[r], b);
        }
      }
    }
    if (more) store_chunk(buf ^ 1);
    wcur = wnext;
    __syncthreads();
  }

  const int64_t n = n0 + 4 * lane;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int64_t row = mw + r;
    if (row >= M) break;
    float* c = C + row * ldc + n;
    float4 v = acc[r];
    v.x *= alpha;
    v.y *= alpha;
    v.z *= alpha;
    v.w *= alpha;
    if (VEC && n + 3 < N) {
      if (beta != 0.0f) {
        const float4 o = *reinterpret_cast<const float4*>(c);
        v.x = fmaf(beta, o.x, v.x);
        v.y = fmaf(beta, o.y, v.y);
        v.z = fmaf(beta, o.z, v.z);
        v.w = fmaf(beta, o.w, v.w);
      }
      *reinterpret_cast<float4*>(c) = v;
    } else {
      const float vs[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (n + e < N) c[e] = beta != 0.0f ? fmaf(beta, c[e], vs[e]) : vs[e];
    }
  }
}

// Version 2 of the FP32 row kernel, for FP32 B with ldb % 4 == 0, N % 4 == 0
// and 16-byte aligned B and C.  Differences from k_spmm_rows:
//  - a stage covers 64 K rows (two bitmap words); each row's 64-bit mask
//    drives one bit loop per stage;
//  - B is staged with global_load_lds_dwordx4 (HBM -> LDS directly, 1 KB per
//    wave instruction = one 256-column row segment), issued one stage ahead,
//    so no registers hold staged data and the loads overlap the bit loops.
//    The loads are issued from a two-instruction asm block (they write LDS
//    only, no VGPR results), and the one wait on them (vmcnt(0)) sits right
//    before each stage's barrier, so nothing waits on the prefetch mid-stage;
//  - bitmap words are prefetched two stages ahead and values one stage ahead,
//    so every global load a stage consumes was issued before its barrier;
//  - set bits are consumed in batches of 4, 2 and 1 independent LDS reads,
//    keeping several reads in flight per wave; within a row the adds stay in
//    ascending k order.
// Lanes beyond N or rows beyond K load a clamped, valid address; that data is
// never used (rows beyond K have no bits, columns beyond N are never stored).
__device__ __forceinline__ void lds_dma16(const void* g, uint32_t lds_addr) {
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off"
               :
               : "v"(g), "s"(lds_addr)
               : "memory", "m0");
}
__device__ __forceinline__ void wait_vmem() { asm volatile("s_waitcnt vmcnt(0)" ::: "memory"); }
__device__ __forceinline__ uint32_t lds_addr_of(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}

// m holds the stage's row mask bit-reversed (bit k <-> stage column k), so the
// lowest set bit is the next column: s_ff1 + shift + xor on the scalar unit.
__device__ __forceinline__ void take_bit(uint64_t& m, int& k) {
  // two scalar instructions per bit (find first set, clear it); the generic
  // form k = ctz(m), m ^= 1 << k costs a shift and an xor instead of the
  // clear, and this loop is bound by scalar issue
  asm volatile("s_ff1_i32_b64 %0, %1" : "=s"(k) : "s"(m));
  asm volatile("s_bitset0_b64 %0, %1" : "+s"(m) : "s"(k));
}

// One output tile ; the
// tile index is column-block-major within the problem.
// The grouped form's value tables (bws_group_rows): the value of set bit i of
// word w of row r is gv[(G * t + g) * 32 + i] with g = gid[r * ldw + w] >= 0
// and t = (r / T) * wpr + w; a spilled word (g < 0) takes the row's own
// value (values in row order, or 1.0).
struct GroupTab {
  const int32_t* gid;
  const float* gv;
  int T, G;
};

template <int WAVES, int RPW, int KS, bool VALS, bool GRP = false>
__device__ __forceinline__ void rows2_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K,
                                           int64_t ldw, const int64_t* __restrict__ row_ptr,
                                           const float* __restrict__ values,
                                           const float* __restrict__ B, int64_t ldb, int64_t N,
                                           float* __restrict__ C, int64_t ldc, float alpha,
                                           float beta, int64_t tile, int64_t row_blocks,
                                           GroupTab gt = GroupTab{nullptr, nullptr, 1, 1}) {
  static_assert(KS == 32 || KS == 64, "stage is one or two bitmap words");
  static_assert(VALS || !GRP, "grouped values walk like values");
  constexpr int WPS = KS / 32;  // bitmap words per stage
  __shared__ float4 sB0[KS][64];
  __shared__ float4 sB1[KS][64];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t cb = tile / row_blocks, rb = tile - cb * row_blocks;
  const int64_t n0 = cb * kBN;
  const int64_t mw = rb * (WAVES * RPW) + (int64_t)wave * RPW;
  const int64_t nst = (K + KS - 1) / KS;
  const int64_t wpr = (K + 31) / 32;
  // lane l < RPW holds word WPS*s of row mw+l, lane 32+l holds word 2s+1 (KS = 64)
  const int rl = lane & 31;
  const bool row_lane = rl < RPW && mw + rl < M;
  const int64_t wrow = row_lane ? (mw + rl) * ldw : 0;
  const int64_t ncol = (n0 + 4 * lane + 4 <= N) ? n0 + 4 * lane : N - 4;  // clamped column
  const float* bcol = B + ncol;
  const uint32_t base0 = lds_addr_of(&sB0[0][0]), base1 = lds_addr_of(&sB1[0][0]);

  float4 acc[RPW];
#pragma unroll
  for (int r = 0; r < RPW; ++r) acc[r] = make_float4(0.f, 0.f, 0.f, 0.f);

  auto issue = [&](int64_t st, uint32_t base) {
#pragma unroll
    for (int kr = 0; kr < KS / WAVES; ++kr) {
      const int row = wave + WAVES * kr;
      int64_t k = st * KS + row;
      k = k < K ? k : K - 1;
      lds_dma16(bcol + k * ldb, base + (uint32_t)row * 1024u);
    }
  };
  auto load_words = [&](int64_t st) -> uint32_t {
    const int64_t w = WPS * st + (lane >> 5);
    const bool ok = WPS == 2 || lane < 32;
    return (ok && row_lane && st < nst && w < wpr) ? W[wrow + w] : 0u;
  };
  auto stage_mask = [&](int64_t st) -> uint64_t {
    const int64_t kvalid = K - st * KS;
    return kvalid >= 64 ? ~0ull : (~0ull << (64 - kvalid));
  };
  // this lane's row mask (lane l < RPW <-> row l) for the stage words w
  auto lane_row_mask = [&](uint32_t w, uint64_t smask) -> uint64_t {
    const uint32_t w_hi = (uint32_t)__shfl((int)w, rl, 64);
    const uint32_t w_lo = WPS == 2 ? (uint32_t)__shfl((int)w, rl + 32, 64) : 0u;
    return ((((uint64_t)w_hi) << 32) | w_lo) & smask;
  };

  int64_t e_l = 0;         // VALS: lane l < RPW: value offset of row l at the current stage
  int cnt_cur = 0;         // VALS: set bits of row l in the current stage
  float vv[VALS ? RPW : 1];
  if (VALS && (!GRP || values)) e_l = row_lane ? row_ptr[mw + rl] : 0;
  // GRP: lane p's value is that of stage column p of row r (zero when the bit
  // is clear), so the walk reads it by column, not by rank; tbase[r] is the
  // group-table tile index of row r's word 0
  int64_t tbase[GRP ? RPW : 1];
#pragma unroll
  for (int r = 0; r < (GRP ? RPW : 1); ++r) {
    const int64_t row = mw + r < M ? mw + r : M - 1;
    tbase[r] = GRP ? (row / gt.T) * wpr : 0;
  }
  auto grp_values = [&](float (&dst)[VALS ? RPW : 1], uint32_t words, int64_t st, int64_t e_row) {
    const uint64_t sm = stage_mask(st);
    const int p = lane;
    const int64_t wp = WPS * st + (p >> 5);
    const int64_t wc = wp < wpr ? wp : wpr - 1;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      const uint32_t whi = (uint32_t)__builtin_amdgcn_readlane((int)words, r);
      const uint32_t wlo = WPS == 2 ? (uint32_t)__builtin_amdgcn_readlane((int)words, r + 32) : 0u;
      const uint64_t m = ((((uint64_t)whi) << 32) | wlo) & sm;
      const bool set = p < KS && st < nst && ((m >> (63 - p)) & 1ull) != 0;
      int64_t row = mw + r;
      row = row < M ? row : M - 1;
      const int32_t g = gt.gid[row * ldw + wc];  // clamped, unconditional
      const int64_t t = tbase[GRP ? r : 0] + wp;
      const int64_t er = readlane_i64(e_row, r);
      float v = 0.0f;
      if (set) {
        if (g >= 0) {
          v = gt.gv[(t * gt.G + g) * 32 + (p & 31)];
        } else if (values) {
          const int rank = p == 0 ? 0 : __popcll(m >> (64 - p));
          v = values[er + rank];
        } else {
          v = 1.0f;
        }
      }
      dst[VALS ? r : 0] = v;
    }
  };

  issue(0, base0);
  uint32_t wcur = load_words(0);
  uint32_t wnext = load_words(1);
  if (VALS) {
    cnt_cur = __popcll(lane_row_mask(wcur, stage_mask(0)));
    if (GRP) {
      grp_values(vv, wcur, 0, e_l);
    } else {
#pragma unroll
      for (int r = 0; r < RPW; ++r) {
        const int64_t e = readlane_i64(e_l, r);
        const int c = __builtin_amdgcn_readlane(cnt_cur, r);
        vv[r] = lane < c ? values[e + lane] : 0.0f;
      }
    }
  }

  auto stage = [&](int64_t st, float4 (*cur)[64], uint32_t nxt_base) {
    wait_vmem();     // this wave's loads for stage st (and its words/values) landed
    // The registers below were loaded in earlier stages and have now landed;
    // re-defining them through an empty asm statement tells the compiler so,
    // which keeps it from waiting (vmcnt(0)) on the prefetch issued next.
    asm volatile("" : "+v"(wcur), "+v"(wnext));
    if (VALS) {
#pragma unroll
      for (int r = 0; r < RPW; ++r) asm volatile("" : "+v"(vv[VALS ? r : 0]));
    }
    __syncthreads(); // everyone's; the other buffer is no longer being read
    const bool more = st + 1 < nst;
    if (more) issue(st + 1, nxt_base);
    const uint32_t wn2 = load_words(st + 2);
    const uint64_t smask = stage_mask(st);
    float vn[VALS ? RPW : 1];
    int cnt_next = 0;
    if (VALS) {
      const int64_t e_next = e_l + cnt_cur;
      cnt_next = more ? __popcll(lane_row_mask(wnext, stage_mask(st + 1))) : 0;
      if (GRP) {
        grp_values(vn, wnext, st + 1, e_next);
      } else {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
          const int64_t e = readlane_i64(e_next, r);
          const int c = __builtin_amdgcn_readlane(cnt_next, r);
          vn[r] = lane < c ? values[e + lane] : 0.0f;
        }
      }
    }

    // every row's stage mask, masked and bit-reversed once in VALU (lane r
    // holds row r), so each row below only reads its mask and its count
    const uint32_t wo = WPS == 2 ? (uint32_t)__shfl_xor((int)wcur, 32, 64) : 0u;
    const uint64_t mm = ((((uint64_t)wcur) << 32) | wo) & smask;
    const uint32_t rev_lo = __brev((uint32_t)(mm >> 32)), rev_hi = __brev((uint32_t)mm);
    const int cnt_v = __popc(rev_lo) + __popc(rev_hi);
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
      uint64_t m = (((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)rev_hi, r)) << 32) |
                   (uint32_t)__builtin_amdgcn_readlane((int)rev_lo, r);
      int n = __builtin_amdgcn_readlane(cnt_v, r);
      int j = 0;
      while (n >= 4) {
        int k0, k1, k2, k3;
        take_bit(m, k0);
        take_bit(m, k1);
        take_bit(m, k2);
        take_bit(m, k3);
        const float4 b0 = cur[k0][lane], b1 = cur[k1][lane], b2 = cur[k2][lane],
                     b3 = cur[k3][lane];
        if (VALS) {
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k0 : j), b0);
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k1 : j + 1), b1);
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k2 : j + 2), b2);
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k3 : j + 3), b3);
        } else {
          add4(acc[r], b0);
          add4(acc[r], b1);
          add4(acc[r], b2);
          add4(acc[r], b3);
        }
        n -= 4;
        j += 4;
      }
      if (n >= 2) {
        int k0, k1;
        take_bit(m, k0);
        take_bit(m, k1);
        const float4 b0 = cur[k0][lane], b1 = cur[k1][lane];
        if (VALS) {
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k0 : j), b0);
          fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k1 : j + 1), b1);
        } else {
          add4(acc[r], b0);
          add4(acc[r], b1);
        }
        n -= 2;
        j += 2;
      }
      if (n) {
        int k0;
        take_bit(m, k0);
        const float4 b0 = cur[k0][lane];
        if (VALS) fma4(acc[r], readlane_f(vv[VALS ? r : 0], GRP ? k0 : j), b0);
        else add4(acc[r], b0);
      }
    }
    if (VALS) {
      e_l += cnt_cur;
      cnt_cur = cnt_next;
#pragma unroll
      for (int r = 0; r < RPW; ++r) vv[r] = vn[r];
    }
    wcur = wnext;
    wnext = wn2;
  };

  for (int64_t st = 0; st < nst; st += 2) {
    stage(st, sB0, base1);
    if (st + 1 < nst) stage(st + 1, sB1, base0);
  }

  const int64_t n = n0 + 4 * lane;
  if (n >= N) return;
#pragma unroll
  for (int r = 0; r < RPW; ++r) {
    const int64_t row = mw + r;
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

template <int WAVES, int RPW, int KS, bool VALS>
__global__ void __launch_bounds__(WAVES * 64)
k_spmm_rows2(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
             const int64_t* __restrict__ row_ptr, const float* __restrict__ values,
             const float* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
             int64_t ldc, float alpha, float beta, int64_t row_blocks, int64_t col_blocks,
             int64_t per_xcd) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  rows2_tile<WAVES, RPW, KS, VALS>(W, M, K, ldw, row_ptr, values, B, ldb, N, C, ldc, alpha, beta,
                                   tile, row_blocks);
}

// Grouped-form evaluation (bws_spmm_groups) on the LDS-staged tile.
template <int WAVES, int RPW, int KS>
__global__ void __launch_bounds__(WAVES * 64)
k_spmm_rows2_groups(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                    const int64_t* __restrict__ row_ptr, const float* __restrict__ values,
                    GroupTab gt, const float* __restrict__ B, int64_t ldb, int64_t N,
                    float* __restrict__ C, int64_t ldc, int64_t row_blocks, int64_t col_blocks,
                    int64_t per_xcd) {
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * col_blocks) return;
  rows2_tile<WAVES, RPW, KS, true, true>(W, M, K, ldw, row_ptr, values, B, ldb, N, C, ldc, 1.0f,
                                         0.0f, tile, row_blocks, gt);
}

// Grouped launch: workgroup g takes global tile xcd_tile(g), finds its problem
// by binary search over the tile offsets and runs that problem's tile.
template <int WAVES, int RPW, int KS, bool VALS>
__global__ void __launch_bounds__(WAVES * 64)
k_spmm_rows2_batch(const BatchEntry* __restrict__ entries, const int64_t* __restrict__ tile_off,
                   int count, int64_t per_xcd) {
  const int64_t g = xcd_tile(blockIdx.x, per_xcd);
  if (g >= tile_off[count]) return;
  int lo = 0, hi = count;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (tile_off[mid] <= g) lo = mid;
    else hi = mid;
  }
  const BatchEntry& e = entries[lo];
  rows2_tile<WAVES, RPW, KS, VALS>(e.W, e.M, e.K, e.ldw, e.row_ptr, e.values,
                                   static_cast<const float*>(e.B), e.ldb, e.N, e.C, e.ldc, 1.0f,
                                   0.0f, g - tile_off[lo], e.row_blocks);
}

// Grouped evaluation: one wave per (row, 256-column block), values from the
// group tables (group_vals of the row's group in this word) or, for a spilled
// row, from A's own values.
__global__ void __launch_bounds__(256)
k_spmm_groups(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
              const int64_t* __restrict__ row_ptr, const float* __restrict__ values,
              const int32_t* __restrict__ group_id, const float* __restrict__ gv, int T, int G,
              const float* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
              int64_t ldc, int64_t col_blocks) {
  const int lane = lane_id();
  const int64_t wpr = words_for(K);
  const int64_t total = M * col_blocks;
  for (int64_t gw = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); gw < total;
       gw += (int64_t)gridDim.x * 4) {
    const int64_t r = gw / col_blocks, cb = gw - r * col_blocks;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    int64_t e = values ? row_ptr[r] : 0;
    for (int64_t w = 0; w < wpr; ++w) {
      uint32_t m = W[r * ldw + w] & valid_bits(w, K);
      if (!m) continue;
      const int32_t g = group_id[r * ldw + w];
      const int64_t t = (r / T) * wpr + w;
      while (m) {
        const int i = __clz((int)m);
        m &= ~wbit(i);
        const float a = g >= 0 ? gv[(t * G + g) * 32 + i] : (values ? values[e] : 1.0f);
        ++e;
        const float* b = B + (32 * w + i) * ldb;
        // lanes past column N read the last column and never store: the four
        // loads carry no branch, so they are in flight together
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const int64_t n = cb * 256 + q * 64 + lane;
          acc[q] = fmaf(a, b[n < N ? n : N - 1], acc[q]);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t n = cb * 256 + q * 64 + lane;
      if (n < N) C[r * ldc + n] = acc[q];
    }
  }
}

// FP32 SDDMM: dA[row_ptr[r] + j] = Σ_n dC[r, n] B[k_j, n].  One wave per row;
// the row of dC is held in registers in 1024-column slices.  With dense_tile
// (one byte per 64 x 64 tile, tile index (k / 64) * tile_rows + r / 64), the
// bits of tiles flagged 1 are skipped (the tiled kernel computed them) and
// rows with nothing left are not visited.
template <bool VEC>
__global__ void __launch_bounds__(256)
k_sddmm_f32(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
            const int64_t* __restrict__ row_ptr, const float* __restrict__ dC, int64_t lddc,
            const float* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ dA,
            const uint8_t* __restrict__ dense_tile, int64_t tile_rows) {
  const int lane = lane_id();
  const int64_t wpr = words_for(K);
  auto skip_word = [&](int64_t r, int64_t w) {
    return dense_tile && dense_tile[(w >> 1) * tile_rows + (r >> 6)] != 0;
  };
  for (int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); r < M;
       r += (int64_t)gridDim.x * 4) {
    if (dense_tile) {
      bool any = false;
      for (int64_t w = lane; w < wpr; w += 64)
        any |= !skip_word(r, w) && (W[r * ldw + w] & valid_bits(w, K)) != 0u;
      if (!__any(any)) continue;
    }
    const int64_t e0 = row_ptr[r];
    for (int64_t n0 = 0; n0 < N || n0 == 0; n0 += 1024) {
      float4 x[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        x[q] = load_b4<0, VEC>(dC + r * lddc, 0, n0 + 256 * q + 4 * lane, 0, 1, N);
      int64_t e = e0;
      uint64_t skipm = 0;  // words w0 .. w0+63 to skip, fetched a group at a time
      for (int64_t w = 0; w < wpr; ++w) {
        if (dense_tile && (w & 63) == 0) skipm = __ballot(w + lane < wpr && skip_word(r, w + lane));
        uint32_t m = W[r * ldw + w] & valid_bits(w, K);
        if ((skipm >> (w & 63)) & 1ull) {
          e += __popc(m);
          continue;
        }
        while (m) {
          const int i = __clz((int)m);
          m &= ~wbit(i);
          const float* b = B + (32 * w + i) * ldb;
          float s = 0.f;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const float4 y = load_b4<0, VEC>(b, 0, n0 + 256 * q + 4 * lane, 0, 1, N);
            s = fmaf(x[q].x, y.x, s);
            s = fmaf(x[q].y, y.y, s);
            s = fmaf(x[q].z, y.z, s);
            s = fmaf(x[q].w, y.w, s);
          }
          s = wave_sum(s);
          if (lane == 0) dA[e] = n0 == 0 ? s : dA[e] + s;
          ++e;
        }
      }
    }
  }
}

template <int WAVES, int RPW, int KS, bool VALS>
bws_status launch_rows2(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N, float* C,
                        int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, WAVES * RPW);
  const int64_t col_blocks = ceil_div(N, kBN);
  const int64_t tiles = row_blocks * col_blocks;
  const int64_t per_xcd = ceil_div(tiles, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL((k_spmm_rows2<WAVES, RPW, KS, VALS>), dim3((unsigned)(8 * per_xcd)),
                     dim3(WAVES * 64), 0, s, A->words, A->rows, A->cols, A->ldw, A->row_ptr,
                     A->values, B, ldb, N, C, ldc, alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

template <int TB, int RPW, bool VALS, bool VEC>
bws_status launch_rows(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                       int64_t ldc, float alpha, float beta, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, kWaves * RPW);
  const int64_t col_blocks = ceil_div(N, kBN);
  const int64_t tiles = row_blocks * col_blocks;
  const int64_t per_xcd = ceil_div(tiles, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL((k_spmm_rows<TB, RPW, VALS, VEC>), dim3((unsigned)(8 * per_xcd)),
                     dim3(kThreads), 0, s, A->words, A->rows, A->cols, A->ldw, A->row_ptr,
                     A->values, static_cast<const typename BType<TB>::T*>(B), ldb, N, C, ldc,
                     alpha, beta, row_blocks, col_blocks, per_xcd);
  return launch_check();
}

template <int TB, int RPW>
bws_status launch_rows_v(const bws_bitmap* A, const void* B, int64_t ldb, int64_t N, float* C,
                         int64_t ldc, float alpha, float beta, bool vec, hipStream_t s) {
  const bool vals = A->values != nullptr;
  if (vals) {
    return vec ? launch_rows<TB, RPW, true, true>(A, B, ldb, N, C, ldc, alpha, beta, s)
               : launch_rows<TB, RPW, true, false>(A, B, ldb, N, C, ldc, alpha, beta, s);
  }
  return vec ? launch_rows<TB, RPW, false, true>(A, B, ldb, N, C, ldc, alpha, beta, s)
             : launch_rows<TB, RPW, false, false>(A, B, ldb, N, C, ldc, alpha, beta, s);
}

inline int grid_for(int64_t items, int64_t per_block) {
  int64_t g = ceil_div(items, per_block);
  if (g < 1) g = 1;
  if (g > 65536) g = 65536;
  return (int)g;
}

template <int WAVES, int RPW, int KS, bool VALS>
bws_status launch_rows2_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                              int64_t total_tiles, hipStream_t s) {
  const int64_t per_xcd = ceil_div(total_tiles, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  if (total_tiles == 0) return BWS_OK;
  hipLaunchKernelGGL((k_spmm_rows2_batch<WAVES, RPW, KS, VALS>), dim3((unsigned)(8 * per_xcd)),
                     dim3(WAVES * 64), 0, s, entries, tile_off, count, per_xcd);
  return launch_check();
}

}  // namespace

// Grouped vector-add launch (FP32 B meeting the direct-staging conditions).
bws_status spmm_valu_batch(const BatchEntry* entries, const int64_t* tile_off, int count,
                           int64_t total_tiles, const ValuCfg& cfg, bool vals, hipStream_t s) {
#define BWS_VB(WV, RP, KS)                                                                    \
  if (cfg.waves == WV && cfg.rows == WV * RP && cfg.kstage == KS)                             \
    return vals ? launch_rows2_batch<WV, RP, KS, true>(entries, tile_off, count, total_tiles, s) \
                : launch_rows2_batch<WV, RP, KS, false>(entries, tile_off, count, total_tiles, s);
  BWS_VB(8, 32, 64)
  BWS_VB(16, 16, 64)
  BWS_VB(16, 16, 32)
  BWS_VB(8, 16, 64)
  BWS_VB(8, 16, 32)
  BWS_VB(8, 8, 32)
  BWS_VB(16, 8, 32)
  BWS_VB(16, 8, 64)
  BWS_VB(16, 4, 32)
#undef BWS_VB
  return BWS_EINVAL;
}

// Whether bws_spmm's FP32 direct-staging kernel applies to these operands.
bool valu_direct_ok(int32_t b_dtype, const void* B, int64_t ldb, int64_t N, const float* C,
                    int64_t ldc) {
  return b_dtype == BWS_DTYPE_F32 && ldb % 4 == 0 && ldc % 4 == 0 && N % 4 == 0 && N >= 4 &&
         reinterpret_cast<uintptr_t>(B) % 16 == 0 && reinterpret_cast<uintptr_t>(C) % 16 == 0;
}

// Vector-add path launcher shared with the dispatcher in spmm.cpp.
bws_status spmm_valu(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb, int64_t N,
                     float* C, int64_t ldc, float alpha, float beta, const ValuCfg& cfg,
                     hipStream_t s) {
  const size_t esz = b_dtype == BWS_DTYPE_F32 ? 4 : 2;
  const bool vec = (ldb % 4 == 0) && (ldc % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(B) % (4 * esz) == 0) &&
                   (reinterpret_cast<uintptr_t>(C) % 16 == 0);
  if (b_dtype == BWS_DTYPE_F32 && vec && N % 4 == 0 && N >= 4) {
    const float* Bf = static_cast<const float*>(B);
    const bool vals = A->values != nullptr;
#define BWS_V2(WV, RP, KS)                                                              \
  if (cfg.waves == WV && cfg.rows == WV * RP && cfg.kstage == KS)                       \
    return vals ? launch_rows2<WV, RP, KS, true>(A, Bf, ldb, N, C, ldc, alpha, beta, s) \
                : launch_rows2<WV, RP, KS, false>(A, Bf, ldb, N, C, ldc, alpha, beta, s);
    BWS_V2(8, 32, 64)
    BWS_V2(16, 16, 64)
    BWS_V2(16, 16, 32)
    BWS_V2(8, 16, 64)
    BWS_V2(8, 16, 32)
    BWS_V2(8, 8, 32)
    BWS_V2(16, 8, 32)
    BWS_V2(16, 8, 64)
    BWS_V2(16, 4, 32)
#undef BWS_V2
    return BWS_EINVAL;
  }
  const bool big = cfg.rows >= 256;
#define BWS_ROWS(TB)                                                                   \
  return big ? launch_rows_v<TB, 32>(A, B, ldb, N, C, ldc, alpha, beta, vec, s)        \
             : launch_rows_v<TB, 8>(A, B, ldb, N, C, ldc, alpha, beta, vec, s)
  switch (b_dtype) {
    case BWS_DTYPE_F32: BWS_ROWS(0);
    case BWS_DTYPE_BF16: BWS_ROWS(1);
    case BWS_DTYPE_F16: BWS_ROWS(2);
    default: return BWS_EUNSUPPORTED;
  }
#undef BWS_ROWS
}

namespace {

// The vector side of the MFMA density split: one wave per (row, 256 columns)
// of the row blocks flagged 0; per set bit the wave reads the 256 columns of
// B's row straight from memory (8 bytes of 16-bit values per lane, four bits'
// rows in flight) and adds them in FP32.  The rows these blocks hold are
// sparse (that is why they were not given to MFMA), so staging B's panels in
// LDS for them would cost more than the bits.
template <int TB>
__global__ void __launch_bounds__(256)
k_spmm_split_rows(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                  const uint16_t* __restrict__ B, int64_t ldb, int64_t N, float* __restrict__ C,
                  int64_t ldc, float alpha, float beta, const uint8_t* __restrict__ mfma_rows,
                  int64_t col_blocks) {
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (item >= M * col_blocks) return;
  const int64_t r = item / col_blocks, cb = item - r * col_blocks;
  if (mfma_rows[r / kSplitRows]) return;
  const int64_t n = cb * 256 + 4 * lane;
  const int64_t nc = n + 4 <= N ? n : N - 4;  // lanes past N read the last columns, never store
  const int64_t wpr = words_for(K);
  const uint16_t* bcol = B + nc;
  auto cvt2 = [](uint32_t u, float& lo, float& hi) {
    if (TB == 1) {
      lo = __uint_as_float(u << 16);
      hi = __uint_as_float(u & 0xFFFF0000u);
    } else {
      _Float16 h[2];
      __builtin_memcpy(h, &u, 4);
      lo = (float)h[0];
      hi = (float)h[1];
    }
  };
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  // the row's words 64 at a time, a word per lane; the set bits taken in
  // column order, four at a time, across words (uniform scalar work)
  for (int64_t w0 = 0; w0 < wpr; w0 += 64) {
    const int64_t wl = w0 + lane;
    const uint32_t word = wl < wpr ? W[r * ldw + wl] & valid_bits(wl, K) : 0u;
    uint64_t nz = __ballot(word != 0u);
    uint32_t cur = 0;
    int64_t cw = 0;
    while (nz != 0ull || cur != 0u) {
      int64_t k[4];
      int cnt = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (cur == 0u && nz != 0ull) {
          const int l = __builtin_ctzll(nz);
          nz &= nz - 1ull;
          cw = w0 + l;
          cur = (uint32_t)__builtin_amdgcn_readlane((int)word, l);
        }
        k[i] = 32 * cw;  // a row of B that exists (never added when i >= cnt)
        if (cur != 0u) {
          const int b = __clz((int)cur);
          cur &= ~wbit(b);
          k[i] += b;
          ++cnt;
        }
      }
      uint2 u[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) u[i] = *reinterpret_cast<const uint2*>(bcol + k[i] * ldb);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        if (i < cnt) {  // uniform: the row's bits are the wave's
          float x0, x1, x2, x3;
          cvt2(u[i].x, x0, x1);
          cvt2(u[i].y, x2, x3);
          acc.x += x0;
          acc.y += x1;
          acc.z += x2;
          acc.w += x3;
        }
      }
    }
  }
  if (n >= N) return;
  float4* c = reinterpret_cast<float4*>(C + r * ldc + n);
  float4 v = make_float4(alpha * acc.x, alpha * acc.y, alpha * acc.z, alpha * acc.w);
  if (beta != 0.0f) {
    const float4 o = *c;
    v.x = fmaf(beta, o.x, v.x);
    v.y = fmaf(beta, o.y, v.y);
    v.z = fmaf(beta, o.z, v.z);
    v.w = fmaf(beta, o.w, v.w);
  }
  *c = v;
}

}  // namespace

// The vector side of the MFMA density split: the 0/1 product with 16-bit B
// (bf16 or f16; N % 8 == 0, 16-byte aligned rows of B and C) on the row
// blocks whose flag is 0.
bws_status spmm_valu_split16(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                             int64_t N, float* C, int64_t ldc, float alpha, float beta,
                             const uint8_t* mfma_rows, hipStream_t s) {
  const int64_t col_blocks = ceil_div(N, 256);
  const int64_t items = A->rows * col_blocks;
  if (ceil_div(items, 4) > kMaxGrid) return BWS_ECAPACITY;
  const uint16_t* b = static_cast<const uint16_t*>(B);
  if (b_dtype == BWS_DTYPE_BF16)
    hipLaunchKernelGGL(k_spmm_split_rows<1>, dim3((unsigned)ceil_div(items, 4)), dim3(256), 0, s,
                       A->words, A->rows, A->cols, A->ldw, b, ldb, N, C, ldc, alpha, beta,
                       mfma_rows, col_blocks);
  else if (b_dtype == BWS_DTYPE_F16)
    hipLaunchKernelGGL(k_spmm_split_rows<2>, dim3((unsigned)ceil_div(items, 4)), dim3(256), 0, s,
                       A->words, A->rows, A->cols, A->ldw, b, ldb, N, C, ldc, alpha, beta,
                       mfma_rows, col_blocks);
  else
    return BWS_EUNSUPPORTED;
  return launch_check();
}

bws_status spmm_groups_launch(const bws_bitmap* A, const int32_t* group_id, const float* gv,
                              int32_t T, int32_t G, const float* B, int64_t ldb, int64_t N,
                              float* C, int64_t ldc, hipStream_t s) {
  // B rows staged in LDS by 16-byte copies when the rows allow it (the tile
  // kernel of the value-carrying FP32 product, values from the group tables)
  if (valu_direct_ok(BWS_DTYPE_F32, B, ldb, N, C, ldc)) {
    constexpr int WV = 16, RP = 8, KS = 64;
    const int64_t row_blocks = ceil_div(A->rows, WV * RP);
    const int64_t cbs = ceil_div(N, kBN);
    const int64_t per_xcd = ceil_div(row_blocks * cbs, 8);
    if (8 * per_xcd <= kMaxGrid) {
      hipLaunchKernelGGL((k_spmm_rows2_groups<WV, RP, KS>), dim3((unsigned)(8 * per_xcd)),
                         dim3(WV * 64), 0, s, A->words, A->rows, A->cols, A->ldw, A->row_ptr,
                         A->values, GroupTab{group_id, gv, T, G}, B, ldb, N, C, ldc, row_blocks,
                         cbs, per_xcd);
      return launch_check();
    }
  }
  const int64_t col_blocks = ceil_div(N, 256);
  hipLaunchKernelGGL(k_spmm_groups, dim3(grid_for(A->rows * col_blocks, 4)), dim3(256), 0, s,
                     A->words, A->rows, A->cols, A->ldw, A->row_ptr, A->values, group_id, gv,
                     (int)T, (int)G, B, ldb, N, C, ldc, col_blocks);
  return launch_check();
}

bws_status sddmm_f32_launch(const bws_bitmap* A, const float* dC, int64_t lddc, const float* B,
                            int64_t ldb, int64_t N, float* dA, const uint8_t* dense_tile,
                            hipStream_t s) {
  const bool vec = (ldb % 4 == 0) && (lddc % 4 == 0) &&
                   (reinterpret_cast<uintptr_t>(B) % 16 == 0) &&
                   (reinterpret_cast<uintptr_t>(dC) % 16 == 0);
  const int64_t tile_rows = ceil_div(A->rows, 64);
  if (vec)
    hipLaunchKernelGGL(k_sddmm_f32<true>, dim3(grid_for(A->rows, 4)), dim3(256), 0, s, A->words,
                       A->rows, A->cols, A->ldw, A->row_ptr, dC, lddc, B, ldb, N, dA, dense_tile,
                       tile_rows);
  else
    hipLaunchKernelGGL(k_sddmm_f32<false>, dim3(grid_for(A->rows, 4)), dim3(256), 0, s, A->words,
                       A->rows, A->cols, A->ldw, A->row_ptr, dC, lddc, B, ldb, N, dA, dense_tile,
                       tile_rows);
  return launch_check();
}

}  // namespace bws

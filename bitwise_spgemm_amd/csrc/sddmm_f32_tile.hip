// sddmm_f32_tile.hip — the FP32 SDDMM adjoint over 64 x 64 bitmap tiles:
//   dA[row_ptr[r] + rank(r, k)] = Σ_n dC[r, n] · B[k, n]   for every set bit (r, k).
//
// A workgroup takes one tile (64 rows x 64 columns of A, two bitmap words per
// row).  Both operands' rows for the tile (dC rows r0.., B rows k0..) are
// copied into LDS 64 values of N at a time by global_load_lds_dwordx4 (two
// stages in flight, 16 KiB per operand per stage), so B's rows are read from
// L2 once per tile instead of once per set bit.  A work item is two
// consecutive set bits of one row (a row with an odd count ends with a single
// bit) and belongs to a quad of lanes: lane p of the quad reads 16-byte chunks
// 4i + p (i = 0..3) of the dC row once and of both B rows, i.e. 12 LDS reads
// for two dot products instead of 16.  Staged rows are 256 bytes, one full
// span of the LDS banks, in chunk order.  At step g the quad in place j of its
// 16-lane LDS pass (j = (lane >> 2) & 3) reads chunk group i = (g + j) & 3 of
// its rows, so the four quads of a pass read the four different 64-byte
// groups, all banks once, whatever rows their items have.
//
// Each lane keeps four partial sums per bit (components x, y, z, w of its
// chunks) over N in ascending stages of 64; at the end they are added as
// (x + y) + (z + w) and the quad's four as (p0 + p1) + (p2 + p3): a fixed
// order, so results are reproducible.
// Tiles with fewer than 64 set bits are only flagged (dense_tile[tile] = 0):
// staging 2·64 rows for so few dot products costs more than the wave-per-row
// kernel, which then takes exactly those tiles' bits.
//
// Launch conditions (checked by the caller): N a positive multiple of 64,
// lddc and ldb multiples of 4, dC and B 16-byte aligned, and the per-word
// popcount offsets of A in the workspace.
#include "mfma_common.hpp"

namespace bws {
namespace {

constexpr int TT = 64;                // tile: TT rows x TT columns of A
constexpr int NS = 64;                // N values per stage
constexpr int ROWB = NS * 4;          // bytes per staged row (16 chunks)
constexpr int OPB = TT * ROWB;        // bytes per operand per stage
constexpr int NCW = OPB / 1024 / 4;   // 1 KiB copies per wave per operand per stage
constexpr int G = 2;                  // set bits of one row per work item
constexpr int SLOTS = 8;              // work items per quad of lanes per round
constexpr int SPARSE = 64;            // tiles with fewer set bits are left to the row kernel

__device__ __forceinline__ void fma4(float4& acc, const float4& a, const float4& b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.y = fmaf(a.y, b.y, acc.y);
  acc.z = fmaf(a.z, b.z, acc.z);
  acc.w = fmaf(a.w, b.w, acc.w);
}

// The tile's rounds: work item i is set bits G p .. G p + G-1 (in column
// order) of one row; SLOTS items per quad per round.
__device__ __forceinline__ void tile_rounds(const float* __restrict__ dC, int64_t lddc,
                                            const float* __restrict__ B, int64_t ldb, int64_t N,
                                            float* __restrict__ dA, int64_t M, int64_t K,
                                            int64_t r0, int64_t k0, int n_items,
                                            const int32_t* sItem, const uint64_t* sMask,
                                            const int64_t* sBase, uint8_t (*sD)[OPB],
                                            uint8_t (*sB)[OPB]) {
  constexpr int ROUND = SLOTS * 64;  // work items per round (64 quads)
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // copy sources: instruction g of a stage fills staged rows 4g .. 4g+3; lane
  // l writes chunk l % 16 of row 4g + l / 16
  const float* dsrc[NCW];
  const float* bsrc[NCW];
#pragma unroll
  for (int t = 0; t < NCW; ++t) {
    const int row = 4 * (wave * NCW + t) + lane / 16;
    const int c = lane % 16;
    int64_t rd = r0 + row, rk = k0 + row;
    if (rd > M - 1) rd = M - 1;
    if (rk > K - 1) rk = K - 1;
    dsrc[t] = dC + rd * lddc + 4 * c;
    bsrc[t] = B + rk * ldb + 4 * c;
  }
  const uint32_t sD0 = lds_u32(&sD[0][0]), sB0 = lds_u32(&sB[0][0]);
  auto issue = [&](int64_t st) {
    const uint32_t buf = (uint32_t)(st & 1) * OPB;
#pragma unroll
    for (int t = 0; t < NCW; ++t) {
      const uint32_t off = buf + (uint32_t)((wave * NCW + t) * 1024);
      dma16(dsrc[t] + st * NS, sD0 + off);
      dma16(bsrc[t] + st * NS, sB0 + off);
    }
  };
  const int64_t nst = N / NS;

  const int ph = lane & 3;                  // this lane's place in its quad
  const int quad = wave * 16 + (lane >> 2);  // 0 .. 63
  // byte offset within a staged row of the chunk this lane reads at step g
  uint32_t coff[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) coff[g] = (uint32_t)(64 * ((g + (lane >> 2)) & 3) + 16 * ph);
  const uint8_t* sDp = &sD[0][0];
  const uint8_t* sBp = &sB[0][0];
  for (int q = 0; q < n_items; q += ROUND) {
    // Item i of the tile: set bits G p .. G p + G-1 (in column order) of one
    // row.  Per item and step: the dC row's read offset and, per bit, the B
    // row's (the same chunk of each).
    uint32_t offD[SLOTS][4], offB[SLOTS][G][4];
    int64_t out[SLOTS];
    int nb[SLOTS];  // set bits of the item (0 .. G)
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int i = q + 64 * s + quad;
      int rr = 0;
      nb[s] = 0;
      out[s] = 0;
      uint64_t m = 0;
      int first = 0;
      if (i < n_items) {
        int lo = 0, hi = TT;  // sItem[lo] <= i < sItem[hi]
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (sItem[mid] <= i) lo = mid;
          else hi = mid;
        }
        rr = lo;
        m = sMask[rr];
        first = G * (i - sItem[rr]);
        const int c = __popcll(m) - first;
        nb[s] = c < G ? c : G;
        out[s] = sBase[rr] + first;
        for (int x = 0; x < first; ++x) m &= ~(0x8000000000000000ull >> __clzll(m));
      }
#pragma unroll
      for (int g = 0; g < 4; ++g) offD[s][g] = (uint32_t)(rr * ROWB) + coff[g];
#pragma unroll
      for (int b = 0; b < G; ++b) {
        // column of the item's bit b (row 0 when the item has fewer bits)
        const int kk = b < nb[s] ? __clzll(m) : 0;
        if (b < nb[s]) m &= ~(0x8000000000000000ull >> kk);
#pragma unroll
        for (int g = 0; g < 4; ++g) offB[s][b][g] = (uint32_t)(kk * ROWB) + coff[g];
      }
    }
    float4 acc[SLOTS][G];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s)
#pragma unroll
      for (int b = 0; b < G; ++b) acc[s][b] = make_float4(0.f, 0.f, 0.f, 0.f);
    // slots used by some quad of this wave (uniform)
    const int left = n_items - q - 16 * wave;
    const int used = left <= 0 ? 0 : (left + 63) / 64 < SLOTS ? (left + 63) / 64 : SLOTS;

    auto stage = [&](int64_t st, uint32_t buf) {
      wait_vmcnt<0>();  // this wave's copies of stage st landed
      __syncthreads();  // everyone's; stage st-1's buffer is free
      if (st + 1 < nst) issue(st + 1);
#pragma unroll
      for (int s = 0; s < SLOTS; ++s) {
        if (s < used) {  // some quad of this wave has slot s
#pragma unroll
          for (int g = 0; g < 4; ++g) {
            const float4 a = *reinterpret_cast<const float4*>(sDp + buf + offD[s][g]);
#pragma unroll
            for (int b = 0; b < G; ++b) {
              const float4 v = *reinterpret_cast<const float4*>(sBp + buf + offB[s][b][g]);
              fma4(acc[s][b], a, v);
            }
          }
        }
      }
    };
    issue(0);
    for (int64_t st = 0; st < nst; st += 2) {
      stage(st, 0u);
      if (st + 1 < nst) stage(st + 1, (uint32_t)OPB);
    }
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      if (s >= used) break;
#pragma unroll
      for (int b = 0; b < G; ++b) {
        const float v = (acc[s][b].x + acc[s][b].y) + (acc[s][b].z + acc[s][b].w);
        const float v2 = v + __shfl_xor(v, 1, kWave);    // (p0 + p1), (p2 + p3)
        const float v4 = v2 + __shfl_xor(v2, 2, kWave);  // (p0 + p1) + (p2 + p3)
        if (ph == 0 && b < nb[s]) dA[out[s] + b] = v4;
      }
    }
    __syncthreads();  // the buffers are reused by the next round
  }
}

__global__ void __launch_bounds__(256, 2)
k_sddmm_f32_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
                 const float* __restrict__ dC, int64_t lddc, const float* __restrict__ B,
                 int64_t ldb, int64_t N, float* __restrict__ dA, uint8_t* __restrict__ dense_tile,
                 int64_t row_blocks, int64_t k_blocks, int64_t per_xcd) {
  __shared__ __attribute__((aligned(1024))) uint8_t sD[2][OPB];
  __shared__ __attribute__((aligned(1024))) uint8_t sB[2][OPB];
  __shared__ int64_t sBase[TT];      // value index of each row's first bit in the tile
  __shared__ uint64_t sMask[TT];     // each row's 64 tile columns, MSB = first column
  __shared__ int32_t sItem[TT + 1];  // exclusive prefix of the rows' item counts
  __shared__ int32_t sTotal;         // set bits in the tile
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * k_blocks) return;
  // groups of 8 row blocks, column blocks in order within a group, row blocks
  // fastest: the ~64 tiles an XCD runs together are 8 row blocks x 8 column
  // blocks, so each stage's slice of a dC row or a B row is fetched from HBM
  // once for 8 tiles (row blocks fastest over all of them had every tile of
  // a column block read its own dC rows: dC came from HBM 64 times)
  const int64_t gsz = 8 * k_blocks, grp = tile / gsz, first = grp * 8;
  const int64_t rows_g = row_blocks - first < 8 ? row_blocks - first : 8;
  const int64_t in = tile - grp * gsz;
  const int64_t rb = first + in % rows_g, kb = in / rows_g;
  const int64_t r0 = rb * TT, k0 = kb * TT;
  const int64_t wpr = words_for(K);

  // wave 0: lane l takes row r0 + l (its two words, its bit count, its items)
  const int64_t wr = r0 + lane, w0 = k0 / 32;
  const bool row_in = wave == 0 && wr < M;
  uint64_t mask = 0;
  if (row_in) {
    const uint32_t a = w0 < wpr ? W[wr * ldw + w0] & valid_bits(w0, K) : 0u;
    const uint32_t b = w0 + 1 < wpr ? W[wr * ldw + w0 + 1] & valid_bits(w0 + 1, K) : 0u;
    mask = ((uint64_t)a << 32) | b;
  }
  const int cnt = __popcll(mask);
  if (wave == 0) {
    const int bits = wave_sum(cnt);
    const int incl = wave_incl_scan((cnt + G - 1) / G);
    sItem[lane + 1] = incl;
    if (lane == 0) sItem[0] = 0;
    sMask[lane] = mask;
    if (lane == 0) sTotal = bits;
  }
  __syncthreads();
  const int total = sTotal;
  if (tid == 0) dense_tile[kb * row_blocks + rb] = total >= SPARSE ? 1 : 0;
  if (total < SPARSE) return;
  if (row_in && cnt > 0) sBase[lane] = row_ptr[wr] + woff[wr * wpr + w0];
  const int n_items = sItem[TT];
  __syncthreads();

  tile_rounds(dC, lddc, B, ldb, N, dA, M, K, r0, k0, n_items, sItem, sMask, sBase, sD, sB);
}

}  // namespace

// Whether the tiled FP32 SDDMM takes these operands.
bool sddmm_f32_tile_ok(const float* dC, int64_t lddc, const float* B, int64_t ldb, int64_t N) {
  return N > 0 && N % NS == 0 && lddc % 4 == 0 && ldb % 4 == 0 &&
         reinterpret_cast<uintptr_t>(dC) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0;
}

int64_t sddmm_f32_tiles(int64_t rows, int64_t cols) {
  return ceil_div(rows, TT) * ceil_div(cols, TT);
}

// woff: per-word popcount offsets of A (launch_word_offsets); dense_tile: one
// byte per tile (sddmm_f32_tiles), written here for every tile, 1 where this
// kernel computed the tile's bits.  Tile t covers rows 64 (t % row_blocks) ..
// and columns 64 (t / row_blocks) .., row_blocks = ceil(rows / 64).
bws_status launch_sddmm_f32_tile(const bws_bitmap* A, const int32_t* woff, const float* dC,
                                 int64_t lddc, const float* B, int64_t ldb, int64_t N, float* dA,
                                 uint8_t* dense_tile, hipStream_t s) {
  const int64_t row_blocks = ceil_div(A->rows, TT), k_blocks = ceil_div(A->cols, TT);
  const int64_t per_xcd = ceil_div(row_blocks * k_blocks, 8);
  if (8 * per_xcd > kMaxGrid) return BWS_ECAPACITY;
  hipLaunchKernelGGL(k_sddmm_f32_tile, dim3((unsigned)(8 * per_xcd)), dim3(256), 0, s, A->words,
                     A->rows, A->cols, A->ldw, A->row_ptr, woff, dC, lddc, B, ldb, N, dA,
                     dense_tile, row_blocks, k_blocks, per_xcd);
  return launch_check();
}

}  // namespace bws

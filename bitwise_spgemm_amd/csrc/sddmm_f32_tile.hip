// sddmm_f32_tile.hip — the FP32 SDDMM adjoint over 64 x 64 bitmap tiles:
//   dA[row_ptr[r] + rank(r, k)] = Σ_n dC[r, n] · B[k, n]   for every set bit (r, k).
//
// A workgroup takes one tile (64 rows x 64 columns of A, two bitmap words per
// row).  Both operands' rows for the tile (dC rows r0.., B rows k0..) are
// copied into LDS 64 values of N at a time by global_load_lds_dwordx4 (two
// stages in flight, 16 KiB per operand per stage).  Each set bit's dot
// product belongs to a quad of lanes: lane p of the quad reads 16-byte chunks
// 4i + p (i = 0..3) of both rows from LDS, so B's rows are read from L2 once
// per tile instead of once per set bit.  Chunk c of staged row r sits at slot
// 4 ((c >> 2) ^ (r & 3)) + ((c & 3) ^ ((r >> 2) & 3)): a quad fills one
// aligned group of four slots, consecutive rows (the dC rows of consecutive
// bits) fall in different groups, and the four quads of a 16-lane LDS pass
// meet in a group only when their B rows agree in bits 0-1 (~2-way conflicts,
// against ~3-way when every lane reads a whole row).
//
// Each lane keeps four partial sums (components x, y, z, w of its chunks) over
// N in ascending stages of 64; at the end they are added as (x + y) + (z + w)
// and the quad's four as (p0 + p1) + (p2 + p3): a fixed order, so results are
// reproducible.
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
constexpr int TW = 2 * TT;            // bitmap words per tile
constexpr int NS = 64;                // N values per stage
constexpr int ROWB = NS * 4;          // bytes per staged row (16 chunks)
constexpr int OPB = TT * ROWB;        // bytes per operand per stage
constexpr int NCW = OPB / 1024 / 4;   // 1 KiB copies per wave per operand per stage
constexpr int SLOTS = 8;              // set bits per quad of lanes per round
constexpr int ROUND = SLOTS * 64;     // set bits per round (64 quads)
constexpr int SPARSE = 64;            // tiles with fewer set bits are left to the row kernel

__device__ __forceinline__ uint32_t chunk_slot(int r, int c) {
  return (uint32_t)(4 * ((c >> 2) ^ (r & 3)) + ((c & 3) ^ ((r >> 2) & 3)));
}

__device__ __forceinline__ void fma4(float4& acc, const float4& a, const float4& b) {
  acc.x = fmaf(a.x, b.x, acc.x);
  acc.y = fmaf(a.y, b.y, acc.y);
  acc.z = fmaf(a.z, b.z, acc.z);
  acc.w = fmaf(a.w, b.w, acc.w);
}

__global__ void __launch_bounds__(256, 2)
k_sddmm_f32_tile(const uint32_t* __restrict__ W, int64_t M, int64_t K, int64_t ldw,
                 const int64_t* __restrict__ row_ptr, const int32_t* __restrict__ woff,
                 const float* __restrict__ dC, int64_t lddc, const float* __restrict__ B,
                 int64_t ldb, int64_t N, float* __restrict__ dA, uint8_t* __restrict__ dense_tile,
                 int64_t row_blocks, int64_t k_blocks, int64_t per_xcd) {
  __shared__ __attribute__((aligned(1024))) uint8_t sD[2][OPB];
  __shared__ __attribute__((aligned(1024))) uint8_t sB[2][OPB];
  __shared__ int64_t sBase[TW];   // value index of each tile word's first bit
  __shared__ uint32_t sW[TW];     // the tile's words, row-major (two per row)
  __shared__ int32_t sPre[TW + 1];  // exclusive prefix of their popcounts
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t tile = xcd_tile(blockIdx.x, per_xcd);
  if (tile >= row_blocks * k_blocks) return;
  // row blocks vary fastest: the tiles an XCD runs together share B's rows
  const int64_t kb = tile / row_blocks, rb = tile - kb * row_blocks;
  const int64_t r0 = rb * TT, k0 = kb * TT;
  const int64_t wpr = words_for(K);

  int cnt = 0;
  const int64_t wr = r0 + (tid >> 1), ww = k0 / 32 + (tid & 1);  // thread tid < TW: word tid
  const bool in = tid < TW && wr < M && ww < wpr;
  if (tid < TW) {
    const uint32_t word = in ? W[wr * ldw + ww] & valid_bits(ww, K) : 0u;
    cnt = __popc(word);
    sW[tid] = word;
  }
  // exclusive prefix over the 128 words: a scan in each of waves 0 and 1,
  // then wave 1 adds wave 0's total
  int incl = 0;
  if (wave < 2) incl = wave_incl_scan(cnt);
  if (tid < TW) sPre[tid + 1] = incl;
  if (tid == 0) sPre[0] = 0;
  __syncthreads();
  if (tid >= TT && tid < TW) sPre[tid + 1] = incl + sPre[TT];
  __syncthreads();
  const int total = sPre[TW];
  if (tid == 0) dense_tile[tile] = total >= SPARSE ? 1 : 0;
  if (total < SPARSE) return;
  if (tid < TW) sBase[tid] = in ? row_ptr[wr] + woff[wr * wpr + ww] : 0;
  __syncthreads();

  // set bit j of the tile (row-major, MSB-first within a word): staged rows
  // rr (of dC) and kk (of B) and its value index
  auto locate = [&](int j, int& rr, int& kk, int64_t& out) {
    int lo = 0, hi = TW;  // sPre[lo] <= j < sPre[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (sPre[mid] <= j) lo = mid;
      else hi = mid;
    }
    uint32_t m = sW[lo];
    for (int x = sPre[lo]; x < j; ++x) m &= ~wbit(__clz((int)m));  // drop earlier bits
    rr = lo >> 1;
    kk = 32 * (lo & 1) + __clz((int)m);
    out = sBase[lo] + (j - sPre[lo]);
  };

  // copy sources: instruction g of a stage fills staged rows 4g .. 4g+3; lane
  // l writes slot l % 16 of row 4g + l / 16, which holds chunk
  // chunk_slot(row, l % 16) (the mapping is its own inverse)
  const float* dsrc[NCW];
  const float* bsrc[NCW];
#pragma unroll
  for (int t = 0; t < NCW; ++t) {
    const int row = 4 * (wave * NCW + t) + lane / 16;
    const int c = (int)chunk_slot(row, lane % 16);
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
  const uint8_t* sDp = &sD[0][0];
  const uint8_t* sBp = &sB[0][0];
  for (int q = 0; q < total; q += ROUND) {
    // Per slot: the byte offsets of this lane's reads.  With x = r & 3 and
    // y = (r >> 2) & 3, chunk 4i + ph of row r sits at slot 4 (i ^ x) + (ph ^ y):
    // reading the dC row's slot groups g = 0..3 in order takes its chunks
    // i = g ^ x_d, and the matching B chunk is in group g ^ x_d ^ x_b.
    uint32_t offD[SLOTS], offB[SLOTS][4];
    int64_t out[SLOTS];
    bool ok[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) {
      const int j = q + 64 * s + quad;
      ok[s] = j < total;
      int rr = 0, kk = 0;
      out[s] = 0;
      if (ok[s]) locate(j, rr, kk, out[s]);
      offD[s] = (uint32_t)(rr * ROWB + 16 * (ph ^ ((rr >> 2) & 3)));
      const int z = (rr ^ kk) & 3;
#pragma unroll
      for (int g = 0; g < 4; ++g)
        offB[s][g] = (uint32_t)(kk * ROWB + 16 * (ph ^ ((kk >> 2) & 3)) + 64 * (g ^ z));
    }
    float4 acc[SLOTS];
#pragma unroll
    for (int s = 0; s < SLOTS; ++s) acc[s] = make_float4(0.f, 0.f, 0.f, 0.f);
    // slots used by some quad of this wave (uniform)
    const int left = total - q - 16 * wave;
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
            const float4 a = *reinterpret_cast<const float4*>(sDp + buf + offD[s] + 64 * g);
            const float4 b = *reinterpret_cast<const float4*>(sBp + buf + offB[s][g]);
            fma4(acc[s], a, b);
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
      const float v = (acc[s].x + acc[s].y) + (acc[s].z + acc[s].w);
      const float v2 = v + __shfl_xor(v, 1, kWave);     // (p0 + p1), (p2 + p3)
      const float v4 = v2 + __shfl_xor(v2, 2, kWave);   // (p0 + p1) + (p2 + p3)
      if (ph == 0 && ok[s]) dA[out[s]] = v4;
    }
    __syncthreads();  // the buffers are reused by the next round
  }
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

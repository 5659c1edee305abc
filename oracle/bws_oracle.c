/*
 * bws_oracle.c — CPU restatement of the reference's bit-wise SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); see
 * bws_oracle.h.  Plain C99, compiled with gcc by oracle/Makefile.
 *
 * Citations are to files of ZSL98/bitwise_spgemm.  Where the reference is
 * race-ordered or buggy, the canonical semantics restated here are:
 *   - grouping assigns rows in ascending order (the CAS loop of
 *     source/src/main.cu:104-181 serialised), with the spill path of
 *     main.cu:153-166 taken whenever no group fits (the reference's first-scan
 *     overflow, main.cu:122-130, writes out of bounds instead);
 *   - accumulators start at zero (the reference leaves several unzeroed).
 */
#include "bws_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define WBIT(i) (0x80000000u >> (i)) /* column i of a word <-> bit 31-i */

static inline int popc32(uint32_t x) { return __builtin_popcount(x); }
static inline int64_t words_for(int64_t cols) { return (cols + 31) / 32; }

/* number of set bits of row r strictly before column c (c inside word w) */
static int64_t rank_in_row(const uint32_t* W, int64_t ldw, int64_t r, int64_t c) {
  int64_t w = c / 32, i = c % 32, n = 0;
  for (int64_t x = 0; x < w; ++x) n += popc32(W[r * ldw + x]);
  if (i) n += popc32(W[r * ldw + w] & ~(0xFFFFFFFFu >> i));
  return n;
}

/* value of set bit (r, c): values[row_ptr[r] + rank] or 1.0 for a 0/1 mask
 * (the popcount-prefix lookup of source/src/mm.cu:272) */
static float value_at(const uint32_t* W, int64_t ldw, const int64_t* row_ptr,
                      const float* values, int64_t r, int64_t c) {
  if (!values) return 1.0f;
  return values[row_ptr[r] + rank_in_row(W, ldw, r, c)];
}

/* ------------------------------------------------------------------------ */
/* layout transforms                                                         */
/* ------------------------------------------------------------------------ */

/* dense2bitmask, source/src/main.cu:591-628 (TILE_WIDTH 32 branch, 612-625):
 * bit (31-i) of word (row, seg) set iff MatB_dense[row, 32*seg + i] != 0.
 * The reference ORs into an unzeroed buffer; here every word is written. */
void oref_pack_bitmap_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                          uint32_t* W, int64_t ldw) {
  for (int64_t r = 0; r < rows; ++r) {
    for (int64_t w = 0; w < ldw; ++w) W[r * ldw + w] = 0u;
    for (int64_t c = 0; c < cols; ++c)
      if (D[r * ld + c] != 0.0f) W[r * ldw + c / 32] |= WBIT(c % 32);
  }
}

/* 0/1 unpack: the inverse of the pack above for a mask. */
void oref_unpack_bitmap_f32(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                            float* D, int64_t ld) {
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c)
      D[r * ld + c] = (W[r * ldw + c / 32] & WBIT(c % 32)) ? 1.0f : 0.0f;
}

/* layout_trans, backup/layout_trans.cu:75-85: dst[i*MatSize + tid] =
 * (bits(src[tid]) >> i) & 1 for i < bit_width. */
void oref_unpack_bitplanes_f32(const float* src, int64_t n, int32_t bit_width, int32_t* dst) {
  for (int64_t t = 0; t < n; ++t) {
    uint32_t v;
    memcpy(&v, &src[t], 4);
    for (int32_t b = 0; b < bit_width; ++b) dst[(int64_t)b * n + t] = (int32_t)((v >> b) & 1u);
  }
}

/* uint8_layout_trans, backup/layout_trans.cu:87-97: 8 planes of a uint8. */
void oref_unpack_bitplanes_u8(const uint8_t* src, int64_t n, int32_t* dst) {
  for (int64_t t = 0; t < n; ++t)
    for (int32_t b = 0; b < 8; ++b) dst[(int64_t)b * n + t] = (src[t] >> b) & 1;
}

/* Packed int8 planes: plane b of the two's-complement byte (the int8 value
 * bits the numeric kernel tests, source/src/transform.cuh:2212), each plane
 * packed as in dense2bitmask. */
void oref_pack_bitplanes_i8(const int8_t* D, int64_t rows, int64_t cols, int64_t ld,
                            uint32_t* P, int64_t ldw) {
  for (int32_t b = 0; b < 8; ++b) {
    uint32_t* Pb = P + (int64_t)b * rows * ldw;
    for (int64_t r = 0; r < rows; ++r) {
      for (int64_t w = 0; w < ldw; ++w) Pb[r * ldw + w] = 0u;
      for (int64_t c = 0; c < cols; ++c)
        if (((uint8_t)D[r * ld + c] >> b) & 1u) Pb[r * ldw + c / 32] |= WBIT(c % 32);
    }
  }
}

/* initialize_multiplicand, source/src/main.cu:889-912: the row of plane
 * weights [1,2,4,8,16,32,64,-128]. */
void oref_bitplane_weights(float* w8) {
  for (int b = 0; b < 7; ++b) w8[b] = (float)(1 << b);
  w8[7] = -128.0f;
}

/* layout_compact intent, backup/layout_trans.cu:100-118: per plane, the
 * number of set elements of each row (dst[b*row + blockIdx.x]). */
void oref_bitplane_row_counts(const uint32_t* P, int32_t planes, int64_t rows, int64_t cols,
                              int64_t ldw, int32_t* counts) {
  int64_t wpr = words_for(cols);
  for (int32_t b = 0; b < planes; ++b)
    for (int64_t r = 0; r < rows; ++r) {
      int32_t n = 0;
      for (int64_t w = 0; w < wpr; ++w) n += popc32(P[(int64_t)b * rows * ldw + r * ldw + w]);
      counts[(int64_t)b * rows + r] = n;
    }
}

/* tile_nnz / tile_nnz_acc exclusive prefix of csr2tiledcsr,
 * source/src/main.cu:517-535, at row granularity. */
void oref_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw, int64_t* row_ptr) {
  int64_t wpr = words_for(cols), acc = 0;
  for (int64_t r = 0; r < rows; ++r) {
    row_ptr[r] = acc;
    for (int64_t w = 0; w < wpr; ++w) acc += popc32(W[r * ldw + w]);
  }
  row_ptr[rows] = acc;
}

/* value placement of csr2tiledcsr, source/src/main.cu:537-563: nonzeros of
 * each row in ascending column order. */
void oref_pack_values_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                          const uint32_t* W, int64_t ldw, const int64_t* row_ptr, float* values) {
  for (int64_t r = 0; r < rows; ++r) {
    int64_t e = row_ptr[r];
    for (int64_t c = 0; c < cols; ++c)
      if (W[r * ldw + c / 32] & WBIT(c % 32)) values[e++] = D[r * ld + c];
  }
}

void oref_transpose_bitmap(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                           uint32_t* WT, int64_t ldwt) {
  for (int64_t c = 0; c < cols; ++c)
    for (int64_t w = 0; w < ldwt; ++w) WT[c * ldwt + w] = 0u;
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c)
      if (W[r * ldw + c / 32] & WBIT(c % 32)) WT[c * ldwt + r / 32] |= WBIT(r % 32);
}

void oref_transpose_values(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                           const int64_t* row_ptr, const float* values, uint32_t* WT,
                           int64_t ldwt, int64_t* row_ptr_T, float* values_T) {
  oref_transpose_bitmap(W, rows, cols, ldw, WT, ldwt);
  oref_row_ptr(WT, cols, rows, ldwt, row_ptr_T);
  for (int64_t c = 0; c < cols; ++c) {
    int64_t e = row_ptr_T[c];
    for (int64_t r = 0; r < rows; ++r)
      if (W[r * ldw + c / 32] & WBIT(c % 32))
        values_T[e++] = values ? value_at(W, ldw, row_ptr, values, r, c) : 1.0f;
  }
}

/* ------------------------------------------------------------------------ */
/* grouping                                                                  */
/* ------------------------------------------------------------------------ */

int64_t oref_group_tiles(int64_t rows, int64_t cols, int32_t tile_rows) {
  return ((rows + tile_rows - 1) / tile_rows) * words_for(cols);
}

/* generate_groups_large, source/src/main.cu:50-247, canonical form.
 * Tile bid = blockIdx.y * gridDim.x + blockIdx.x = row_tile * wpr + word
 * (main.cu:69-72).  Per row in ascending order:
 *   mask == 0                    -> group_idx 0, no member bits  (main.cu:103-107, 191-195)
 *   first g with !(row_group[g] & mask) -> join g                  (main.cu:111-181)
 *   none                         -> spill: group_id -1, hash slot = spilled_row_cnt++,
 *                                   spilled_nnz += popcount        (main.cu:153-166)
 * Outputs as written at main.cu:199-244: group masks, group values
 * d_group_ele_row_val[(G*bid+g)*32+i] from the member row of bit i
 * (main.cu:224-236), grouped nnz += popcount(row_group[g]) (main.cu:201-204),
 * spill hash table (init 0) and reverse table (init -1) (main.cu:83-84, 208-211,
 * 241-244). */
void oref_group_rows(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                     const int64_t* row_ptr, const float* values, int32_t tile_rows,
                     int32_t max_groups, int32_t* group_id, uint32_t* group_mask,
                     float* group_vals, int32_t* member_row, int32_t* spill_cnt,
                     int32_t* spill_nnz, int32_t* spill_table, int32_t* spill_rev,
                     int64_t* grouped_nnz) {
  const int64_t wpr = words_for(cols);
  const int64_t T = tile_rows, G = max_groups;
  const int64_t row_tiles = (rows + T - 1) / T;
  int64_t total = 0;
  uint32_t gm[32];
  for (int64_t rt = 0; rt < row_tiles; ++rt) {
    for (int64_t w = 0; w < wpr; ++w) {
      const int64_t t = rt * wpr + w;
      int32_t cnt = 0, snnz = 0;
      for (int64_t g = 0; g < G; ++g) {
        gm[g] = 0u;
        for (int i = 0; i < 32; ++i) {
          group_vals[(t * G + g) * 32 + i] = 0.0f;
          if (member_row) member_row[(t * G + g) * 32 + i] = -1;
        }
      }
      if (spill_table)
        for (int64_t j = 0; j < T; ++j) spill_table[t * T + j] = 0;
      if (spill_rev)
        for (int64_t j = 0; j < T; ++j) spill_rev[t * T + j] = -1;
      for (int64_t j = 0; j < T && rt * T + j < rows; ++j) {
        const int64_t r = rt * T + j;
        const uint32_t m = W[r * ldw + w];
        if (m == 0u) {
          group_id[r * ldw + w] = 0;
          continue;
        }
        int64_t g = 0;
        while (g < G && (gm[g] & m)) ++g;
        if (g < G) {
          gm[g] |= m;
          group_id[r * ldw + w] = (int32_t)g;
          for (int i = 0; i < 32; ++i)
            if (m & WBIT(i)) {
              group_vals[(t * G + g) * 32 + i] = value_at(W, ldw, row_ptr, values, r, 32 * w + i);
              if (member_row) member_row[(t * G + g) * 32 + i] = (int32_t)j;
            }
        } else {
          group_id[r * ldw + w] = -1;
          if (spill_table) spill_table[t * T + cnt] = (int32_t)j;
          if (spill_rev) spill_rev[t * T + j] = cnt;
          ++cnt;
          snnz += popc32(m);
        }
      }
      for (int64_t g = 0; g < G; ++g) {
        group_mask[t * G + g] = gm[g];
        total += popc32(gm[g]);
      }
      spill_cnt[t] = cnt;
      spill_nnz[t] = snnz;
    }
  }
  if (grouped_nnz) *grouped_nnz = total;
}

/* generate_spilled_csr, source/src/main.cu:445-496, with the host exclusive
 * prefix of spill counts (main.cu:2093-2100).  Per tile in slot order, the
 * nonzero columns j of the spilled row (main.cu:481-490) and the running end
 * offset as tile_spilled_csrRowPtr (main.cu:491). */
void oref_group_spilled_csr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                            const int64_t* row_ptr, const float* values, int32_t tile_rows,
                            const int32_t* spill_cnt, const int32_t* spill_nnz,
                            const int32_t* spill_table, int32_t* spill_row_off,
                            int32_t* spill_nnz_off, int32_t* row_end, int32_t* col, float* val) {
  const int64_t wpr = words_for(cols);
  const int64_t tiles = oref_group_tiles(rows, cols, tile_rows);
  int32_t racc = 0, nacc = 0;
  for (int64_t t = 0; t < tiles; ++t) {
    spill_row_off[t] = racc;
    spill_nnz_off[t] = nacc;
    racc += spill_cnt[t];
    nacc += spill_nnz[t];
  }
  spill_row_off[tiles] = racc;
  spill_nnz_off[tiles] = nacc;
  for (int64_t t = 0; t < tiles; ++t) {
    const int64_t rt = t / wpr, w = t % wpr;
    int32_t e = 0;
    for (int32_t s = 0; s < spill_cnt[t]; ++s) {
      const int64_t r = rt * tile_rows + spill_table[t * tile_rows + s];
      const uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i)
        if (m & WBIT(i)) {
          col[spill_nnz_off[t] + e] = i;
          val[spill_nnz_off[t] + e] = value_at(W, ldw, row_ptr, values, r, 32 * w + i);
          ++e;
        }
      row_end[spill_row_off[t] + s] = e;
    }
  }
}

/* group2dense, source/src/main.cu:829-850: scatter group values by group id
 * and the row's mask bits (31-i). */
void oref_group_scatter(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                        const int32_t* group_id, const float* group_vals, int32_t tile_rows,
                        int32_t max_groups, float* D, int64_t ld) {
  const int64_t wpr = words_for(cols);
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t w = 0; w < wpr; ++w) {
      const int32_t g = group_id[r * ldw + w];
      if (g < 0) continue;
      const int64_t t = (r / tile_rows) * wpr + w;
      const uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i)
        if (m & WBIT(i)) D[r * ld + 32 * w + i] = group_vals[(t * max_groups + g) * 32 + i];
    }
}

/* dense2group_from_idx, source/src/main.cu:853-886: the inverse gather. */
void oref_group_gather(const float* D, int64_t ld, const uint32_t* W, int64_t rows,
                       int64_t cols, int64_t ldw, const int32_t* group_id, int32_t tile_rows,
                       int32_t max_groups, float* group_vals) {
  const int64_t wpr = words_for(cols);
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t w = 0; w < wpr; ++w) {
      const int32_t g = group_id[r * ldw + w];
      if (g < 0) continue;
      const int64_t t = (r / tile_rows) * wpr + w;
      const uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i)
        if (m & WBIT(i)) group_vals[(t * max_groups + g) * 32 + i] = D[r * ld + 32 * w + i];
    }
}

/* pre_spgemm symbolic phase, source/src/transform.cuh:1028-1061:
 * bit_indicator(row, seg) = OR of B's row masks over row's A nonzeros. */
void oref_symbolic_or(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa,
                      const uint32_t* WB, int64_t N, int64_t ldwb, uint32_t* WC, int64_t ldwc) {
  const int64_t wpn = words_for(N);
  for (int64_t r = 0; r < M; ++r) {
    for (int64_t s = 0; s < ldwc; ++s) WC[r * ldwc + s] = 0u;
    for (int64_t k = 0; k < K; ++k)
      if (WA[r * ldwa + k / 32] & WBIT(k % 32))
        for (int64_t s = 0; s < wpn; ++s) WC[r * ldwc + s] |= WB[k * ldwb + s];
  }
}

/* ------------------------------------------------------------------------ */
/* row clustering                                                            */
/* ------------------------------------------------------------------------ */

/* Bits of word w that are columns < cols (MSB-first). */
static uint32_t valid_word(int64_t w, int64_t cols) {
  const int64_t rem = cols - 32 * w;
  return rem >= 32 ? 0xFFFFFFFFu : (rem <= 0 ? 0u : (0xFFFFFFFFu << (32 - rem)));
}

/* Leaders used for `rows` rows: max_clusters, or rows / 64 when it is 0;
 * at least 1 and at most rows and 65536. */
int32_t oref_cluster_leaders(int64_t rows, int32_t max_clusters) {
  if (rows <= 0) return 0;
  int64_t L = max_clusters > 0 ? max_clusters : rows / 64;
  if (L < 1) L = 1;
  if (L > rows) L = rows;
  if (L > 65536) L = 65536;
  return (int32_t)L;
}

/* Leaders taken by the exact farthest-point rounds over all rows: the last
 * min(4, L - 1) of L. */
int32_t oref_cluster_exact_rounds(int32_t L) { return L - 1 < 4 ? (L > 0 ? L - 1 : 0) : 4; }

/* Candidate rows of the leader search for the first Lc = L - exact rounds
 * leaders: Lc <= 256 searches min(rows, 256) candidates; more take the
 * candidates themselves (c = Lc).  Candidate j is row floor(j * rows / c). */
int64_t oref_cluster_candidates(int64_t rows, int32_t L) {
  if (rows <= 0) return 0;
  const int32_t Lc = L - oref_cluster_exact_rounds(L);
  if (Lc > 256) return Lc;
  return rows < 256 ? rows : 256;
}

/* The row-clustering plan, serially.  There is no reference function to
 * follow: the reference groups rows by bitmask (generate_groups_large,
 * source/src/main.cu:50-247; the row-to-group intent of backup/exp.cu:49-88)
 * so that a group's rows share work; this plan groups similar rows so that a
 * chunk's shared bits are summed once.
 *   1. Leaders.  R = min(4, L - 1) exact rounds, Lc = L - R.  c =
 *      oref_cluster_candidates(rows, L) candidate rows, candidate j = row
 *      floor(j * rows / c).  Leaders 0 .. Lc-1: with Lc <= 256, farthest
 *      point over the candidates with the distance min(popcount(x ^ y),
 *      65535): leader 0 is candidate 0; every candidate keeps its distance to
 *      the nearest leader chosen so far; the next leader is the candidate with
 *      the largest such distance (the lowest candidate on ties); with Lc > 256,
 *      leader i is candidate i.  Leaders Lc .. L-1: farthest point over all
 *      rows with the full Hamming distance: the row whose distance to its
 *      nearest leader so far is largest (the lowest row on ties).
 *   2. Every row's nearest leader by full Hamming distance (the earliest
 *      leader on ties); rows ordered by nearest leader, ascending row within
 *      a leader.
 *   3. Chunk c = slots 16c .. 16c+15 (fewer in the last).  Its rows form runs
 *      of equal leader; base 0 is made from the longest run, base 1 from the
 *      second longest (earlier run first on equal lengths; no base 1 when the
 *      chunk has one run): bit k of base j is set iff 2 * (rows of run j with
 *      bit k) > (rows in run j) + 1.  chunk_mask[(2c + j) * ldw + w].
 *   4. slot_row[s] = row, plus (1 << 30) when the row is summed as a
 *      difference from a base: the base j minimising popcount(row ^ base_j),
 *      if that is below popcount(row) (base 0 on ties); (1 << 29) marks
 *      base 1. */
static int64_t hamming_rows(const uint32_t* W, int64_t ldw, int64_t cols, int64_t a, int64_t b) {
  const int64_t wpr = words_for(cols);
  int64_t d = 0;
  for (int64_t w = 0; w < wpr; ++w)
    d += popc32((W[a * ldw + w] ^ W[b * ldw + w]) & valid_word(w, cols));
  return d;
}

void oref_cluster_rows(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                       int32_t max_clusters, int32_t* slot_row, uint32_t* chunk_mask) {
  if (rows <= 0) return;
  const int64_t wpr = words_for(cols);
  const int32_t L = oref_cluster_leaders(rows, max_clusters);
  const int64_t c = oref_cluster_candidates(rows, L);
  int64_t* lead_row = (int64_t*)malloc(sizeof(int64_t) * L);
  int32_t* owner = (int32_t*)malloc(sizeof(int32_t) * rows);
  int64_t* start = (int64_t*)calloc((size_t)L + 1, sizeof(int64_t));
  int32_t* perm = (int32_t*)malloc(sizeof(int32_t) * rows);
  const int32_t Lc = L - oref_cluster_exact_rounds(L);
  if (Lc > 256) {
    for (int32_t i = 0; i < Lc; ++i) lead_row[i] = (int64_t)i * rows / c;
  } else {
    int64_t* near = (int64_t*)malloc(sizeof(int64_t) * c);
    int64_t lead = 0;
    for (int32_t it = 0; it < Lc; ++it) {
      lead_row[it] = lead * rows / c;
      for (int64_t j = 0; j < c; ++j) {
        int64_t d = hamming_rows(W, ldw, cols, j * rows / c, lead_row[it]);
        if (d > 65535) d = 65535;
        if (it == 0 || d < near[j]) near[j] = d;
      }
      int64_t best = 0;
      for (int64_t j = 1; j < c; ++j)
        if (near[j] > near[best]) best = j;
      lead = best;
    }
    free(near);
  }
  /* nearest of the first Lc leaders, then the exact rounds */
  int64_t* dist = (int64_t*)malloc(sizeof(int64_t) * rows);
  for (int64_t r = 0; r < rows; ++r) {
    dist[r] = -1;
    for (int32_t i = 0; i < Lc; ++i) {
      const int64_t d = hamming_rows(W, ldw, cols, r, lead_row[i]);
      if (dist[r] < 0 || d < dist[r]) {
        dist[r] = d;
        owner[r] = i;
      }
    }
  }
  for (int32_t i = Lc; i < L; ++i) {
    int64_t best = 0;
    for (int64_t r = 1; r < rows; ++r)
      if (dist[r] > dist[best]) best = r;
    lead_row[i] = best;
    for (int64_t r = 0; r < rows; ++r) {
      const int64_t d = hamming_rows(W, ldw, cols, r, best);
      if (d < dist[r]) {
        dist[r] = d;
        owner[r] = i;
      }
    }
  }
  free(dist);
  for (int64_t r = 0; r < rows; ++r) start[owner[r] + 1]++;
  for (int32_t o = 0; o < L; ++o) start[o + 1] += start[o];
  for (int64_t r = 0; r < rows; ++r) perm[start[owner[r]]++] = (int32_t)r;
  const int64_t chunks = (rows + 15) / 16;
  for (int64_t c = 0; c < chunks; ++c) {
    const int64_t n = rows - 16 * c < 16 ? rows - 16 * c : 16;
    const int32_t* pr = perm + 16 * c;
    /* the two longest runs of equal leader */
    int64_t s0 = 0, n0 = 0, s1 = 0, n1 = 0;
    for (int64_t i = 0; i < n;) {
      int64_t j = i + 1;
      while (j < n && owner[pr[j]] == owner[pr[i]]) ++j;
      const int64_t len = j - i;
      if (len > n0) {
        s1 = s0, n1 = n0;
        s0 = i, n0 = len;
      } else if (len > n1) {
        s1 = i, n1 = len;
      }
      i = j;
    }
    for (int b = 0; b < 2; ++b) {
      const int64_t bs = b == 0 ? s0 : s1, bn = b == 0 ? n0 : n1;
      for (int64_t w = 0; w < ldw; ++w) {
        uint32_t m = 0u;
        if (w < wpr && bn > 0) {
          const uint32_t vb = valid_word(w, cols);
          for (int k = 0; k < 32; ++k) {
            int64_t cnt = 0;
            for (int64_t i = 0; i < bn; ++i) cnt += (W[(int64_t)pr[bs + i] * ldw + w] & vb & WBIT(k)) != 0;
            if (2 * cnt > bn + 1) m |= WBIT(k);
          }
        }
        chunk_mask[(2 * c + b) * ldw + w] = m;
      }
    }
  }
  for (int64_t s = 0; s < rows; ++s) {
    const int64_t r = perm[s], c = s / 16;
    int64_t pa = 0, p0 = 0, p1 = 0;
    for (int64_t w = 0; w < wpr; ++w) {
      const uint32_t a = W[r * ldw + w] & valid_word(w, cols);
      pa += popc32(a);
      p0 += popc32(a ^ chunk_mask[(2 * c) * ldw + w]);
      p1 += popc32(a ^ chunk_mask[(2 * c + 1) * ldw + w]);
    }
    int32_t flag = 0;
    int64_t best = pa;
    if (p0 < best) flag = 1 << 30, best = p0;
    if (p1 < best) flag = (1 << 30) | (1 << 29);
    slot_row[s] = (int32_t)r | flag;
  }
  free(lead_row);
  free(owner);
  free(start);
  free(perm);
}

/* ------------------------------------------------------------------------ */
/* numerics                                                                  */
/* ------------------------------------------------------------------------ */

/* FP64 oracle C = A·B (absolute != 0: S = |A|·|B|, the error-metric scale).
 * Semantics of backup/layout_trans.cu:145-171 (C[i,j] += B[k,j] * weight for
 * each set bit) and of the intended grouped product of
 * source/src/transform.cuh:2199-2246, accumulated exactly enough in double. */
void oref_spmm_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw, const int64_t* row_ptr,
                   const float* values, const float* B, int64_t ldb, int64_t N, double* C,
                   int64_t ldc, int32_t absolute) {
  const int64_t wpr = words_for(K);
  for (int64_t r = 0; r < M; ++r) {
    double* c = C + r * ldc;
    for (int64_t n = 0; n < N; ++n) c[n] = 0.0;
    int64_t e = values ? row_ptr[r] : 0;
    for (int64_t w = 0; w < wpr; ++w) {
      uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i) {
        if (!(m & WBIT(i))) continue;
        const int64_t k = 32 * w + i;
        double a = values ? (double)values[e++] : 1.0;
        if (absolute) a = fabs(a);
        const float* b = B + k * ldb;
        if (absolute)
          for (int64_t n = 0; n < N; ++n) c[n] += a * fabs((double)b[n]);
        else
          for (int64_t n = 0; n < N; ++n) c[n] += a * (double)b[n];
      }
    }
  }
}

/* Numeric sparse x sparse product onto an output pattern: for the j-th set
 * bit c of row r of WC, out[rpC[r] + j] = Σ_k A[r, k]·B[k, c] in FP64 (A, B
 * value-carrying bitmaps, or 0/1 when their values are NULL; absolute: Σ
 * |A|·|B|).  The numeric phase of the reference's sparse x sparse pipeline
 * (spgemm_compute_1dthread_tcore_v3, source/src/transform.cuh:2199-2246,
 * whose per-group sums land in dC_group_value, source/src/main.cu:2133) on
 * the pattern its symbolic phase builds (transform.cuh:1028-1061). */
void oref_spgemm_f64(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa, const int64_t* rpA,
                     const float* vA, const uint32_t* WB, int64_t N, int64_t ldwb,
                     const int64_t* rpB, const float* vB, const uint32_t* WC, int64_t ldwc,
                     const int64_t* rpC, double* out, int32_t absolute) {
  const int64_t wpa = words_for(K), wpn = words_for(N);
  double* row = (double*)malloc(sizeof(double) * (size_t)(N > 0 ? N : 1));
  for (int64_t r = 0; r < M; ++r) {
    for (int64_t n = 0; n < N; ++n) row[n] = 0.0;
    int64_t ea = vA ? rpA[r] : 0;
    for (int64_t x = 0; x < wpa; ++x)
      for (int i = 0; i < 32; ++i) {
        if (!(WA[r * ldwa + x] & valid_word(x, K) & WBIT(i))) continue;
        const int64_t k = 32 * x + i;
        double a = vA ? (double)vA[ea++] : 1.0;
        if (absolute) a = fabs(a);
        int64_t eb = vB ? rpB[k] : 0;
        for (int64_t w = 0; w < wpn; ++w)
          for (int j = 0; j < 32; ++j) {
            if (!(WB[k * ldwb + w] & valid_word(w, N) & WBIT(j))) continue;
            double b = vB ? (double)vB[eb++] : 1.0;
            if (absolute) b = fabs(b);
            row[32 * w + j] += a * b;
          }
      }
    int64_t ec = rpC[r];
    for (int64_t w = 0; w < wpn; ++w)
      for (int j = 0; j < 32; ++j)
        if (WC[r * ldwc + w] & valid_word(w, N) & WBIT(j)) out[ec++] = row[32 * w + j];
  }
  free(row);
}

/* FP32 in ascending k: acc = fl(acc + fl(a*b)), starting at 0 — the order of
 * the CPU check loop (backup/bit_wise_gemm.cu:17-22) restricted to A's set
 * bits (zero terms add exactly 0). */
void oref_spmm_f32_seq(const uint32_t* W, int64_t M, int64_t K, int64_t ldw,
                       const int64_t* row_ptr, const float* values, const float* B, int64_t ldb,
                       int64_t N, float* C, int64_t ldc) {
  const int64_t wpr = words_for(K);
  for (int64_t r = 0; r < M; ++r) {
    float* c = C + r * ldc;
    for (int64_t n = 0; n < N; ++n) c[n] = 0.0f;
    int64_t e = values ? row_ptr[r] : 0;
    for (int64_t w = 0; w < wpr; ++w) {
      uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i) {
        if (!(m & WBIT(i))) continue;
        const float a = values ? values[e++] : 1.0f;
        const float* b = B + (32 * w + i) * ldb;
        for (int64_t n = 0; n < N; ++n) {
          volatile float p = a * b[n]; /* keep the product rounded to FP32 */
          c[n] = c[n] + p;
        }
      }
    }
  }
}

/* Grouped evaluation (the reference's use of group values as the multiplicand
 * values, source/src/transform.cuh:2224-2239, with roles re-assigned): the
 * value of set bit (r, 32w+i) is group_vals[(G*t + g)*32 + i] for a grouped row
 * and the row's own value for a spilled row. */
void oref_spmm_groups_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw,
                          const int64_t* row_ptr, const float* values, const int32_t* group_id,
                          const float* group_vals, int32_t tile_rows, int32_t max_groups,
                          const float* B, int64_t ldb, int64_t N, double* C, int64_t ldc) {
  const int64_t wpr = words_for(K);
  for (int64_t r = 0; r < M; ++r) {
    double* c = C + r * ldc;
    for (int64_t n = 0; n < N; ++n) c[n] = 0.0;
    for (int64_t w = 0; w < wpr; ++w) {
      const uint32_t m = W[r * ldw + w];
      const int32_t g = group_id[r * ldw + w];
      const int64_t t = (r / tile_rows) * wpr + w;
      for (int i = 0; i < 32; ++i) {
        if (!(m & WBIT(i))) continue;
        const int64_t k = 32 * w + i;
        const double a = g >= 0 ? (double)group_vals[(t * max_groups + g) * 32 + i]
                                : (double)value_at(W, ldw, row_ptr, values, r, k);
        for (int64_t n = 0; n < N; ++n) c[n] += a * (double)B[k * ldb + n];
      }
    }
  }
}

/* SDDMM adjoint dA = (dC·Bᵀ)⊙mask in FP64, in the values order of A
 * (absolute != 0: Σ|dC||B|). */
void oref_sddmm_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw, const int64_t* row_ptr,
                    const float* dC, int64_t lddc, const float* B, int64_t ldb, int64_t N,
                    double* dA, int32_t absolute) {
  const int64_t wpr = words_for(K);
  for (int64_t r = 0; r < M; ++r) {
    int64_t e = row_ptr[r];
    for (int64_t w = 0; w < wpr; ++w) {
      const uint32_t m = W[r * ldw + w];
      for (int i = 0; i < 32; ++i) {
        if (!(m & WBIT(i))) continue;
        const int64_t k = 32 * w + i;
        double s = 0.0;
        for (int64_t n = 0; n < N; ++n) {
          const double x = (double)dC[r * lddc + n], y = (double)B[k * ldb + n];
          s += absolute ? fabs(x) * fabs(y) : x * y;
        }
        dA[e++] = s;
      }
    }
  }
}

/* check(), backup/bit_wise_gemm.cu:13-32: the dense i, j, p triple loop with
 * an FP32 accumulator, `sum += A[i*k + p] * B[j + p*n]`, here writing C for
 * rows [row_begin, row_end) instead of comparing.  threads > 1 splits rows
 * over OpenMP threads (each row's arithmetic is unchanged). */
void oref_check_dense_f32(const float* A, const float* B, float* C, int64_t m, int64_t n,
                          int64_t k, int64_t row_begin, int64_t row_end, int32_t threads) {
  (void)m;
  if (threads < 1) threads = 1;
#pragma omp parallel for num_threads(threads) schedule(static)
  for (int64_t i = row_begin; i < row_end; ++i) {
    for (int64_t j = 0; j < n; ++j) {
      float sum = 0.f;
      for (int64_t p = 0; p < k; ++p) sum += A[i * k + p] * B[j + p * n];
      C[i * n + j] = sum;
    }
  }
}

/* Error metrics against the FP64 oracle R with scale S = |A|·|B|:
 * out[0] max |c-r|/s, out[1] mean |c-r|/s (entries with s > 0; an entry with
 * s == 0 must be exactly 0 and counts as error 0 or +inf),
 * out[2] max |c-r|/|r| and out[3] mean over entries with |r| >= 1e-6*s,
 * out[4] number of entries excluded from out[2..3], out[5] max |c-r|,
 * out[6] max |c-r|/|r| over r != 0 (the reference check's denominator,
 * backup/bit_wise_gemm.cu:24; informational). */
void oref_err_metrics(const float* C, int64_t ldc, const double* R, const double* S,
                      int64_t M, int64_t N, double* out) {
  double mx = 0, sum = 0, mxr = 0, sumr = 0, mxa = 0, mxref = 0;
  int64_t cnt = 0, cntr = 0, excl = 0;
  for (int64_t r = 0; r < M; ++r)
    for (int64_t n = 0; n < N; ++n) {
      const double c = (double)C[r * ldc + n], ref = R[r * N + n], s = S[r * N + n];
      const double d = fabs(c - ref);
      if (!(d <= mxa)) mxa = d;
      double e;
      if (s > 0) e = d / s;
      else e = (d == 0.0) ? 0.0 : INFINITY;
      if (!(e <= mx)) mx = e;
      sum += e;
      ++cnt;
      if (fabs(ref) >= 1e-6 * s && ref != 0.0) {
        const double er = d / fabs(ref);
        if (!(er <= mxr)) mxr = er;
        sumr += er;
        ++cntr;
      } else {
        ++excl;
      }
      if (ref != 0.0) {
        const double eref = d / fabs(ref);
        if (!(eref <= mxref)) mxref = eref;
      }
    }
  out[0] = mx;
  out[1] = cnt ? sum / (double)cnt : 0.0;
  out[2] = mxr;
  out[3] = cntr ? sumr / (double)cntr : 0.0;
  out[4] = (double)excl;
  out[5] = mxa;
  out[6] = mxref;
}

/* ------------------------------------------------------------------------ */
/* 16-bit rounding, round to nearest even                                    */
/* ------------------------------------------------------------------------ */

float oref_round_bf16(float x) {
  uint32_t u;
  memcpy(&u, &x, 4);
  if ((u & 0x7FFFFFFFu) > 0x7F800000u) { /* NaN: keep it quiet */
    u |= 0x00400000u;
    u &= 0xFFFF0000u;
  } else {
    u += 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
  }
  float y;
  memcpy(&y, &u, 4);
  return y;
}

float oref_round_f16(float x) {
  if (isnan(x) || isinf(x) || x == 0.0f) return x;
  int e;
  frexp((double)x, &e); /* x = f * 2^e, f in [0.5, 1): unbiased exponent e-1 */
  int ex = e - 1;
  int qexp = ex < -14 ? -24 : ex - 10; /* half: 10 fraction bits, min normal 2^-14 */
  double q = ldexp(1.0, qexp);
  double y = rint((double)x / q) * q;
  if (fabs(y) > 65504.0) y = copysign(INFINITY, (double)x);
  return (float)y;
}

void oref_round_bf16_array(const float* x, int64_t n, float* y) {
  for (int64_t i = 0; i < n; ++i) y[i] = oref_round_bf16(x[i]);
}

void oref_round_f16_array(const float* x, int64_t n, float* y) {
  for (int64_t i = 0; i < n; ++i) y[i] = oref_round_f16(x[i]);
}

/* hi = bf16(x), lo = bf16(x - hi) (the FP32 subtraction is exact). */
void oref_split_bf16x2(const float* x, int64_t n, float* hi, float* lo) {
  for (int64_t i = 0; i < n; ++i) {
    hi[i] = oref_round_bf16(x[i]);
    lo[i] = oref_round_bf16(x[i] - hi[i]);
  }
}

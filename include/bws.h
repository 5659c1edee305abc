/*
 * bws.h — C ABI of the MI355X (gfx950) bit-wise sparse GEMM library `libbws.so`.
 *
 * C = A·B with the sparse operand A held as packed bitmasks and B dense.
 * The three entry-point families of the reference are kept as a drop-in:
 *   (1) bitmap layout transform     bws_pack_bitmap_f32, bws_unpack_bitplanes_*, ...
 *   (2) row grouping by bitmask     bws_group_rows, bws_group_spilled_csr, bws_group_scatter/gather
 *   (3) grouped tile GEMM           bws_spmm, bws_spmm_groups, bws_spmm_batched (+ adjoints)
 *
 * Conventions (ABI-visible, fixed):
 *   - A bitmap is uint32 words, row-major [rows x ldw], ldw >= ceil(cols/32).
 *     Bit (31 - i) of word w of row r is set iff element (r, 32*w + i) is nonzero
 *     (MSB-first, as dense2bitmask writes it, source/src/main.cu:619).  Bits of
 *     columns >= cols are zero.
 *   - A value-carrying bitmap adds row_ptr[rows+1] (exclusive prefix of the
 *     per-row set-bit counts) and values[nnz] stored in (row, column) order,
 *     i.e. the value of the j-th set bit of row r is values[row_ptr[r] + j]
 *     (the bitmap + popcount-prefix value layout of source/src/mm.h:83-104 and
 *     source/src/mm.cu:272, at row granularity).  values == NULL means every
 *     set bit has value 1.0 (a 0/1 mask).
 *   - Dense matrices are row-major with a leading dimension in elements.
 *   - All device pointers are caller-owned; the library never allocates device
 *     memory on any entry point (workspace is passed in, sized by the matching
 *     *_workspace_size query, 256-byte aligned as hipMalloc returns it;
 *     BWS_EWORKSPACE otherwise).  All work is enqueued on the caller's stream with
 *     no host synchronisation, so the entry points can be captured into a
 *     hipGraph, except the set-up calls bws_batch_prepare and bws_autotune
 *     (they copy or time on the host).
 *   - Every entry point returns a bws_status (0 = BWS_OK, negative = error); no
 *     entry point aborts or prints.  Argument validation happens on the host
 *     before any HIP call, so invalid calls are rejected without a GPU.
 */
#ifndef BWS_H
#define BWS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same type as hipStream_t; NULL selects the null stream. */
typedef struct ihipStream_t* bws_stream;

typedef int32_t bws_status;
enum {
  BWS_OK = 0,
  BWS_EINVAL = -1,       /* bad shape, stride, pointer or enum value            */
  BWS_ECAPACITY = -2,    /* a size exceeds a documented capacity                */
  BWS_EUNSUPPORTED = -3, /* valid but unsupported combination (dtype/precision) */
  BWS_EHIP = -4,         /* a HIP call failed; see bws_last_hip_error()         */
  BWS_EWORKSPACE = -5,   /* workspace pointer NULL or smaller than required     */
  BWS_ENODEVICE = -6,    /* no usable gfx950 device                             */
  BWS_ENOTFOUND = -7     /* no cached autotuner entry for the shape class       */
};

typedef enum { BWS_DTYPE_F32 = 0, BWS_DTYPE_BF16 = 1, BWS_DTYPE_F16 = 2 } bws_dtype;

/* Arithmetic used by the product.  FP32 is the drop-in path (FP32 vector adds /
 * FMAs, FP32 accumulation).  BF16 and FP16 round B (and A's values) to 16 bits
 * and use MFMA with FP32 accumulation.  BF16X2 splits FP32 B into hi + lo
 * bfloat16 parts (hi = bf16(b), lo = bf16(b - hi)) and accumulates both
 * products in FP32 (compensated path). */
typedef enum {
  BWS_PREC_FP32 = 0,
  BWS_PREC_BF16 = 1,
  BWS_PREC_FP16 = 2,
  BWS_PREC_BF16X2 = 3
} bws_precision;

typedef struct {
  int64_t rows;            /* M                                            */
  int64_t cols;            /* K                                            */
  int64_t ldw;             /* words per row stride, >= ceil(cols / 32)     */
  const uint32_t* words;   /* device [rows x ldw]                          */
  const int64_t* row_ptr;  /* device [rows + 1] or NULL (required if values) */
  const float* values;     /* device [nnz] or NULL (0/1 mask)              */
} bws_bitmap;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

/* Human-readable name of a status code (static storage). */
const char* bws_status_string(bws_status status);

/* The last hipError_t recorded by this thread when an entry point returned
 * BWS_EHIP (0 if none). */
int32_t bws_last_hip_error(void);

/* ABI version: (major << 16) | minor. */
int32_t bws_version(void);

/* ------------------------------------------------------------------------ */
/* (1) Bitmap layout transforms                                             */
/* ------------------------------------------------------------------------ */

/* Pack a dense FP32 matrix into the MSB-first bitmap: bit (31-i) of W[r*ldw+w]
 * is set iff D[r*ld + 32w+i] != 0.0f (IEEE compare: -0.0 packs to 0, NaN to 1).
 * Every word of every row (including padding words up to ldw) is written, so W
 * need not be zeroed first.
 * Replaces dense2bitmask (source/src/main.cu:591-628). */
bws_status bws_pack_bitmap_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                               uint32_t* W, int64_t ldw, bws_stream stream);

/* Inverse of the pack for 0/1 masks: D[r, c] = 1.0f if the bit is set, else 0.
 * Used as the dense operand of round trips and of the dense reference. */
bws_status bws_unpack_bitmap_f32(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                                 float* D, int64_t ld, bws_stream stream);

/* Unpacked bit-planes of the raw FP32 bit pattern: dst[b*n + i] = (bits(src[i]) >> b) & 1
 * for b < bit_width (1..32).  Replaces layout_trans (backup/layout_trans.cu:75-85). */
bws_status bws_unpack_bitplanes_f32(const float* src, int64_t n, int32_t bit_width,
                                    int32_t* dst, bws_stream stream);

/* Unpacked bit-planes of uint8 values: dst[b*n + i] = (src[i] >> b) & 1, b < 8.
 * Replaces uint8_layout_trans (backup/layout_trans.cu:87-97). */
bws_status bws_unpack_bitplanes_u8(const uint8_t* src, int64_t n, int32_t* dst,
                                   bws_stream stream);

/* Packed bit-planes of an int8 matrix: plane b (0..7) of the two's-complement
 * pattern of D is packed exactly like bws_pack_bitmap_f32 into
 * P[b*rows*ldw + r*ldw + w].  Σ_b w_b·plane_b reproduces D with the weights
 * [1,2,4,8,16,32,64,-128] of bws_bitplane_weights. */
bws_status bws_pack_bitplanes_i8(const int8_t* D, int64_t rows, int64_t cols, int64_t ld,
                                 uint32_t* P, int64_t ldw, bws_stream stream);

/* Two's-complement plane weights for int8 (host, no GPU): w[b] = 2^b for b<7,
 * w[7] = -128.  The reference's bit-weight multiplicand row
 * (initialize_multiplicand, source/src/main.cu:889-912). */
bws_status bws_bitplane_weights(float* w8);

/* Per-plane per-row set-bit counts of packed planes: counts[b*rows + r] =
 * popcount of row r of plane b (b < planes).  The intent of layout_compact
 * (backup/layout_trans.cu:100-118), computed on packed planes. */
bws_status bws_bitplane_row_counts(const uint32_t* P, int32_t planes, int64_t rows,
                                   int64_t cols, int64_t ldw, int32_t* counts,
                                   bws_stream stream);

/* row_ptr[r] = number of set bits in rows < r (row_ptr[rows] = nnz), with an
 * LDS-staged device scan.  Replaces the tile_nnz / tile_nnz_acc prefix of
 * csr2tiledcsr (source/src/main.cu:517-535). */
size_t bws_row_ptr_workspace_size(int64_t rows);
bws_status bws_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                       int64_t* row_ptr, void* workspace, size_t workspace_bytes,
                       bws_stream stream);

/* values[row_ptr[r] + j] = D[r, c_j] where c_j is the column of the j-th set
 * bit of row r in W.  Replaces the value placement of csr2tiledcsr
 * (source/src/main.cu:537-563). */
bws_status bws_pack_values_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                               const uint32_t* W, int64_t ldw, const int64_t* row_ptr,
                               float* values, bws_stream stream);

/* Bit-matrix transpose: WT (cols x rows bitmap, ldwt >= ceil(rows/32)) has bit
 * (c, r) set iff W has bit (r, c). */
bws_status bws_transpose_bitmap(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                                uint32_t* WT, int64_t ldwt, bws_stream stream);

/* Transpose of a value-carrying bitmap: fills AT's words, row_ptr and values
 * (in AT's (row, column) order).  AT.words/row_ptr/values must point to
 * writable device buffers of sizes cols*ldwt, cols+1 and nnz. */
size_t bws_transpose_workspace_size(int64_t rows, int64_t cols);
bws_status bws_transpose_values(const bws_bitmap* A, uint32_t* WT, int64_t ldwt,
                                int64_t* row_ptr_T, float* values_T, void* workspace,
                                size_t workspace_bytes, bws_stream stream);
/* The same transpose with the value gather left for later: perm_T[j] is the
 * index in A's values of AT's j-th value (perm_T has nnz entries; A.values
 * is ignored).  When the mask stays fixed and only its values change (a
 * trained mask), AT's values are then bws_gather_f32(values, perm_T, nnz). */
bws_status bws_transpose_perm(const bws_bitmap* A, uint32_t* WT, int64_t ldwt,
                              int64_t* row_ptr_T, int64_t* perm_T, void* workspace,
                              size_t workspace_bytes, bws_stream stream);
/* dst[i] = src[idx[i]] for i < n. */
bws_status bws_gather_f32(const float* src, const int64_t* idx, int64_t n, float* dst,
                          bws_stream stream);

/* Format converters (replace the serial tiling pass csr2tiledcsr,
 * source/src/main.cu:499-588, and cuSPARSE DenseToSparse, main.cu:630-690).
 * Into a bitmap, entries may come in any order: W (rows x ldw, every word
 * written) gets one bit per entry; row_ptr [rows+1] (optional) the popcount
 * prefix; values_out [nnz] (optional) the entries' values in the bitmap's
 * (row, column) order.  *n_invalid (device, written by the call) counts
 * entries outside the matrix and repeated positions; when it is nonzero the
 * bitmap holds the valid, distinct positions and values_out is unspecified at
 * repeated ones (values_out is never written past entry nnz - 1).  row_idx /
 * col_idx are int32; CSR row pointers int64. */
size_t bws_convert_workspace_size(int64_t rows, int64_t cols);
bws_status bws_coo_to_bitmap(const int32_t* row_idx, const int32_t* col_idx, const float* values,
                             int64_t nnz, int64_t rows, int64_t cols, uint32_t* W, int64_t ldw,
                             int64_t* row_ptr, float* values_out, unsigned long long* n_invalid,
                             void* workspace, size_t workspace_bytes, bws_stream stream);
/* CSR: nnz = the length of col_idx and values.  Each row reads only entries
 * in [0, nnz): a row pointer that is negative, decreasing, past nnz, or a
 * nonzero csr_row_ptr[0], counts one into *n_invalid per row it affects and
 * that row's range is clamped. */
bws_status bws_csr_to_bitmap(const int64_t* csr_row_ptr, const int32_t* col_idx,
                             const float* values, int64_t nnz, int64_t rows, int64_t cols,
                             uint32_t* W, int64_t ldw, int64_t* row_ptr, float* values_out,
                             unsigned long long* n_invalid, void* workspace,
                             size_t workspace_bytes, bws_stream stream);
/* Out of a bitmap: row_ptr [rows+1], col_idx [nnz] (ascending within a row,
 * may be NULL) and values_out [nnz] (A.values copied, or 1.0 for a 0/1 mask;
 * may be NULL). */
bws_status bws_bitmap_to_csr(const bws_bitmap* A, int64_t* row_ptr, int32_t* col_idx,
                             float* values_out, void* workspace, size_t workspace_bytes,
                             bws_stream stream);

/* ------------------------------------------------------------------------ */
/* (2) Row grouping by bitmask                                              */
/* ------------------------------------------------------------------------ */

/* Grouping tiles are tile_rows rows x one 32-column word.  Tile index
 * t = (r / tile_rows) * wpr + w with wpr = ceil(cols/32) (row-tile major,
 * as in source/src/main.cu:69); tiles = ceil(rows/tile_rows) * wpr.
 * Inside each tile the rows are assigned in ascending order by first fit:
 *   mask == 0              -> group_id 0, not a member;
 *   first g with GM[g]&m==0 -> group_id g, GM[g] |= m, member_row[g][i] = row
 *                              for each set bit i, group_vals[g][i] = value;
 *   no such g              -> group_id -1, spilled: slot = spill_cnt++,
 *                              spill_nnz += popcount(m), spill_table[slot] = row,
 *                              spill_rev[row] = slot.
 * This is generate_groups_large (source/src/main.cu:50-247) with its CAS race
 * resolved in ascending row order and its overflow path made the spill path of
 * source/src/main.cu:153-166.  Rows are within-tile indices.  Unused
 * spill_table slots are 0 and non-spilled spill_rev entries -1 (their
 * initial values in the reference, main.cu:83-84); uncovered group_vals are
 * 0 and uncovered member_row entries -1. */
typedef struct {
  int32_t* group_id;      /* [rows x ldw_g], ldw_g = ldw of the input bitmap      */
  uint32_t* group_mask;   /* [tiles x G]                                           */
  float* group_vals;      /* [tiles x G x 32]                                      */
  int32_t* member_row;    /* [tiles x G x 32] or NULL                              */
  int32_t* spill_cnt;     /* [tiles]                                               */
  int32_t* spill_nnz;     /* [tiles]                                               */
  int32_t* spill_table;   /* [tiles x tile_rows] or NULL                           */
  int32_t* spill_rev;     /* [tiles x tile_rows] or NULL                           */
  int64_t* grouped_nnz;   /* [1]: Σ popcount(group_mask), or NULL                  */
} bws_groups;

/* Limits: 1 <= max_groups <= 32, 1 <= tile_rows <= 65536. */
int64_t bws_group_tiles(int64_t rows, int64_t cols, int32_t tile_rows);
bws_status bws_group_rows(const bws_bitmap* A, int32_t tile_rows, int32_t max_groups,
                          const bws_groups* out, bws_stream stream);

/* Spilled rows as per-tile CSR (generate_spilled_csr, source/src/main.cu:445-496,
 * with the host prefix of main.cu:2093-2100 done on the device): tile t's
 * spilled rows in slot order occupy entries spill_row_off[t] ..
 * spill_row_off[t+1]-1; row_end[e] is the end of that row's nonzeros relative to
 * spill_nnz_off[t] (the reference's tile_spilled_csrRowPtr); col[] the column
 * within the word (0..31) and val[] the value (1.0 for a 0/1 mask).
 * spill_row_off / spill_nnz_off have tiles+1 entries. */
size_t bws_group_spilled_workspace_size(int64_t rows, int64_t cols, int32_t tile_rows);
bws_status bws_group_spilled_csr(const bws_bitmap* A, int32_t tile_rows,
                                 const int32_t* spill_cnt, const int32_t* spill_nnz,
                                 const int32_t* spill_table, int32_t* spill_row_off,
                                 int32_t* spill_nnz_off, int32_t* row_end, int32_t* col,
                                 float* val, void* workspace, size_t workspace_bytes,
                                 bws_stream stream);

/* Un-grouping (group2dense, source/src/main.cu:829-850): for every (r, w) with
 * group_id >= 0, D[r, 32w+i] = group_vals[(G*t + g)*32 + i] for each set bit i
 * of W[r, w]; other elements of D are left untouched. */
bws_status bws_group_scatter(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                             const int32_t* group_id, const float* group_vals,
                             int32_t tile_rows, int32_t max_groups, float* D, int64_t ld,
                             bws_stream stream);

/* Re-grouping (dense2group_from_idx, source/src/main.cu:853-886): the inverse
 * gather, group_vals[(G*t + g)*32 + i] = D[r, 32w+i]. */
bws_status bws_group_gather(const float* D, int64_t ld, const uint32_t* W, int64_t rows,
                            int64_t cols, int64_t ldw, const int32_t* group_id,
                            int32_t tile_rows, int32_t max_groups, float* group_vals,
                            bws_stream stream);

/* Row clustering (similarity grouping) for the FP32 0/1 product.  The
 * reference groups rows by bitmask so that a group's rows share work
 * (generate_groups_large, source/src/main.cu:50-247; the row-to-group intent
 * of backup/exp.cu:49-88); this analysis step has no reference counterpart:
 * it orders the rows so that each run of 16 slots holds similar masks, and
 * bws_spmm_clustered sums each chunk base's bits once.
 *   leaders: L = max_clusters (0: rows / 64), clamped to [1, min(rows, 65536)].
 *     Candidates: c = min(rows, 256) rows when L <= 256, else c = L; candidate
 *     j is row floor(j * rows / c).  L <= 256: farthest point over the
 *     candidates with the distance min(Hamming, 65535): leader 0 is candidate
 *     0, leader i+1 the candidate farthest from its nearest leader so far (the
 *     lowest candidate on ties).  L > 256: leader i is candidate i;
 *   order:   every row's nearest leader by Hamming distance (the earliest
 *            leader on ties); rows by nearest leader, ascending within a leader;
 *   chunks:  chunk c is slots 16c .. 16c+15 (the last may hold fewer); its
 *            slots form runs of equal leader.  Base 0 comes from the longest
 *            run, base 1 from the second longest (the earlier run on equal
 *            lengths; none when the chunk has one run): bit k of base j is set
 *            iff 2 * (rows of the run with bit k) > (rows in the run) + 1.
 *            chunk_mask[(2c + j) * A->ldw + w] holds base j;
 *   slots:   slot_row[s] = row of slot s, plus BWS_SLOT_DELTA when some base
 *            has popcount(row ^ base) < popcount(row) (the smallest wins, base
 *            0 on ties), plus BWS_SLOT_BASE1 when that base is base 1.
 * slot_row has `rows` entries, chunk_mask ceil(rows/16) x 2 x A->ldw words.
 * 0/1 masks only (A->values must be NULL); rows < 2^29 and
 * L * ceil(rows / 1024) <= 2^27 (BWS_ECAPACITY otherwise).  Deterministic, all
 * on the device with no host synchronisation and no memset: capturable into a
 * hipGraph. */
#define BWS_SLOT_DELTA (1 << 30)
#define BWS_SLOT_BASE1 (1 << 29)
size_t bws_cluster_workspace_size(int64_t rows, int64_t cols, int32_t max_clusters);
bws_status bws_cluster_rows(const bws_bitmap* A, int32_t max_clusters, int32_t* slot_row,
                            uint32_t* chunk_mask, void* workspace, size_t workspace_bytes,
                            bws_stream stream);

/* ------------------------------------------------------------------------ */
/* (3) Grouped tile GEMM                                                    */
/* ------------------------------------------------------------------------ */

/* Kernel plan.  All-zero (or a NULL plan pointer) selects the built-in
 * defaults; bws_autotune fills one by measurement.
 *   kernel        0 = auto, 1 = vector-add rows (FP32 accumulate of set bits),
 *                 2 = MFMA tiles (16-bit precisions only)
 *   rows_per_wg   vector-add path: rows of C per workgroup (64, 128 or 256);
 *                 MFMA path: rows of C per workgroup (0 = auto, 128 or 256)
 *   waves_per_wg  vector-add path: 8 or 16 (0 = 16) (MFMA: always 4, pass 0)
 *   k_stage       vector-add path: K rows staged per LDS buffer (32 or 64)
 *                 (MFMA: fixed by the tile, pass 0).  16 with kernel 1 and
 *                 rows_per_wg 0: the FP32 subset-sum product for dense 0/1
 *                 masks (per 4 K rows an LDS table of the 16 sums of their B
 *                 rows; one read and four adds per row nibble, whatever its
 *                 bits: faster than the bit walk above about 40% density),
 *                 with the workspace bws_spmm_workspace_size reports for it
 *   mfma_min_bits BF16 / FP16 on a 0/1 A (0 = off, every row block on MFMA):
 *                 128-row blocks with fewer than mfma_min_bits set bits per
 *                 64 K columns (on average) run on the vector kernel, the
 *                 others on the direct-to-LDS MFMA kernel; needs N % 8 == 0,
 *                 16-byte aligned rows of B (or of B rounded once, for FP32
 *                 B) and the workspace bws_spmm_workspace_size reports (the
 *                 block flags and room to round an FP32 B); otherwise ignored;
 *                 negative: BWS_EINVAL
 *   cols_per_wg   MFMA path: columns of C per workgroup (0 = auto, 128 or 256;
 *                 256 x 256 only with the direct-to-LDS staging of 16-bit B,
 *                 else it runs as 128 x 256; BF16X2 uses 128)
 *   mfma_staging  MFMA path: how B reaches LDS (BWS_MFMA_*; AUTO = direct-to-LDS
 *                 with 2 stages where B allows it: 16-bit B, 16-byte aligned
 *                 rows, N % 8 == 0; otherwise register staging).
 *                 BWS_MFMA_SPARSE: a 0/1 A with 16-bit B that allows the
 *                 direct copy runs on v_smfmac_f32_16x16x64 (A 4:2-compressed
 *                 per row and 4 K columns; nibbles with 3-4 bits take a second
 *                 sparse MFMA), half the MFMA instructions of the dense
 *                 kernel, results equal to it up to FP32 summation order;
 *                 other operands run as AUTO
 *   reserved      pass 0 */
enum {
  BWS_MFMA_AUTO = 0,
  BWS_MFMA_LUT = 1,   /* 64 x 64 per wave, A fragments from an LDS table      */
  BWS_MFMA_REGS = 2,  /* B loaded to registers, transposed, written to LDS    */
  BWS_MFMA_DMA3 = 3,  /* B copied straight into LDS, 3 stages in flight       */
  BWS_MFMA_DMA2 = 4,  /* B copied straight into LDS, 2 stages (the default)   */
  BWS_MFMA_SPARSE = 5 /* 0/1 A on the sparse MFMA (4:2-compressed A, see below) */
};
typedef struct {
  int32_t kernel;
  int32_t rows_per_wg;
  int32_t waves_per_wg;
  int32_t k_stage;
  int32_t mfma_min_bits;
  int32_t cols_per_wg;
  int32_t mfma_staging;
  int32_t reserved;
} bws_plan;

/* C = alpha * (A · B) + beta * C, A M x K bitmap, B K x N dense (dtype b_dtype,
 * leading dimension ldb), C M x N FP32.  beta == 0 never reads C.
 * precision/b_dtype combinations: FP32 with F32 (or exact BF16/F16 input);
 * BF16 with BF16 or F32 (rounded to nearest even); FP16 with F16 or F32;
 * BF16X2 with F32.
 * Workspace (optional; size from bws_spmm_workspace_size, 0 when unused):
 *   - FP32 with a 0/1 A and N % 4 == 0: the mask stream of A (A's words bit
 *     reversed and regrouped so that a wave's 16 row masks of a 64-row stage
 *     are contiguous), written by a pass per call; the product then reads the
 *     masks with scalar loads and walks their bits on the scalar unit (used
 *     when B and C rows are 16-byte aligned and the problem has at least 192
 *     tiles of 256 x 256; without the workspace the vector-add tile kernel
 *     runs, with the same FP32 semantics);
 *   - value-carrying A at BF16/FP16 with 16-bit B: a dense 16-bit copy of A
 *     for the MFMA kernel (without it the FP32 vector path runs);
 *   - BF16X2 with a 0/1 A and N % 8 == 0: B's hi and lo BF16 planes, split
 *     once, then multiplied by the direct-to-LDS BF16 kernel in two passes
 *     (C = alpha·A·hi + beta·C, then C += alpha·A·lo); without it B is split
 *     in registers inside one kernel.  The two forms differ by FP32 rounding
 *     only (one more rounding of the partial sum).
 *   - BF16 / FP16 with a 0/1 A and N % 8 == 0: room for B rounded once to the
 *     16-bit type, used when B is F32 (then multiplied by the direct-to-LDS
 *     kernel; without it B is rounded in registers inside the kernel).  Both
 *     forms give the same result bit for bit.
 * Semantics of the FP32 add path of backup/layout_trans.cu:145-171 and the
 * numeric kernel spgemm_compute_1dthread_tcore_v3 (source/src/transform.cuh:
 * 2062-2248) with the operand roles of a bitmap A and a dense B. */
size_t bws_spmm_workspace_size(const bws_bitmap* A, int64_t N, int32_t precision,
                               const bws_plan* plan);
bws_status bws_spmm(const bws_bitmap* A, const void* B, int32_t b_dtype, int64_t ldb,
                    int64_t N, float* C, int64_t ldc, float alpha, float beta,
                    int32_t precision, const bws_plan* plan, void* workspace,
                    size_t workspace_bytes, bws_stream stream);

/* bws_spmm at BWS_PREC_FP32 with an F32 B and a 0/1 A, allocating nothing and
 * taking no workspace: the mask stream the FP32 workspace would hold lives
 * in a device array of the library, bws_spmm_fp32_noalloc_max_bytes() bytes
 * (16 MiB) on each device it runs on, held until the process ends.  A whose
 * mask stream exceeds it (bws_spmm_workspace_size(A, N, BWS_PREC_FP32, NULL)
 * above the maximum; about rows * cols / 8 bytes, e.g. 8192 x 16384) fails
 * with BWS_ECAPACITY.  Same results as bws_spmm with that workspace.  Every
 * call on a device writes the array before reading it, and calls that
 * overlap on one device (several streams or host threads) race on it.
 * No counterpart in the reference: it exists for callers that cannot
 * allocate (code under graph capture, serving loops, harnesses that forbid
 * hipMalloc). */
size_t bws_spmm_fp32_noalloc_max_bytes(void);
bws_status bws_spmm_fp32_noalloc(const bws_bitmap* A, const float* B, int64_t ldb, int64_t N,
                                 float* C, int64_t ldc, float alpha, float beta,
                                 bws_stream stream);

/* Grouped evaluation: C = A·B where each nonzero's value is read from the
 * group tables (group_vals of its row's group) and spilled rows' values from
 * A.values (or 1.0).  Equals bws_spmm on the same A (FP32 path), since group
 * members are disjoint; this is the reference's grouped value use
 * (transform.cuh:2224-2239) with roles re-assigned. */
bws_status bws_spmm_groups(const bws_bitmap* A, const int32_t* group_id,
                           const float* group_vals, int32_t tile_rows, int32_t max_groups,
                           const float* B, int64_t ldb, int64_t N, float* C, int64_t ldc,
                           bws_stream stream);

/* The row-clustered product's operand stream, built once per plan: for every
 * chunk of 16 slots and every 64-row stage of K, the slots' add masks
 * (row & ~base for a delta row, the row itself otherwise), the chunk's two
 * base masks (zero when no slot uses them), a flag word (bit j: slot j has
 * bits to subtract in this stage), the subtract masks (base & ~row) of the
 * first two flagged slots and every slot's subtract mask, all bit reversed
 * (bit i <-> K row 64*stage + i) and laid out so that a wave's masks of a
 * stage are contiguous.  bws_cluster_stream_size gives the
 * bytes for a rows x cols bitmap (chunks padded to whole 256-row blocks).
 * Capturable (device only). */
size_t bws_cluster_stream_size(int64_t rows, int64_t cols);
bws_status bws_cluster_stream(const bws_bitmap* A, const int32_t* slot_row,
                              const uint32_t* chunk_mask, void* plan_stream, size_t stream_bytes,
                              bws_stream stream);

/* C = alpha * (A · B) + beta * C for a 0/1 bitmap A, FP32 B and C, with the
 * plan of bws_cluster_rows for this A and its stream from bws_cluster_stream:
 * a wave takes one chunk; each delta row is accumulated as (row & ~base)·B -
 * (base & ~row)·B and its base's base·B, summed once per chunk, is added at
 * the end; other rows sum their own bits.  FP32 vector adds; the summation
 * order differs from bws_spmm's, and the result stays within FP32 rounding of
 * A·B relative to |A|·|B| (the subtracted terms are bits of the base).  Needs
 * N, ldb and ldc multiples of 4 and 16-byte aligned B and C
 * (BWS_EUNSUPPORTED otherwise). */
bws_status bws_spmm_clustered(const bws_bitmap* A, const int32_t* slot_row,
                              const void* plan_stream, const float* B, int64_t ldb, int64_t N,
                              float* C, int64_t ldc, float alpha, float beta, bws_stream stream);

/* Adjoint dB = alpha * Aᵀ·dC + beta * dB given AT = transpose of A (from
 * bws_transpose_bitmap / bws_transpose_values).  Identical to bws_spmm on AT. */
bws_status bws_spmm_tA(const bws_bitmap* AT, const void* dC, int32_t dc_dtype, int64_t lddc,
                       int64_t N, float* dB, int64_t lddb, float alpha, float beta,
                       int32_t precision, const bws_plan* plan, void* workspace,
                       size_t workspace_bytes, bws_stream stream);

/* SDDMM adjoint on the bitmap pattern: for the j-th set bit (r, k) of row r,
 * dA[row_ptr[r] + j] = Σ_n dC[r, n] · B[k, n].  A.row_ptr is required; A.values
 * is ignored.  dC and B share dtype `dtype`: F32 for FP32 (vector FMAs) and
 * BF16X2; BF16/F16 or F32 (rounded on the device) for BF16/FP16, which run on
 * MFMA over 16 x 16 output blocks with FP32 accumulate and store only the set
 * bits.  The 16-bit paths need the workspace (per-word popcount offsets).
 * FP32 uses it when given (and N % 64 == 0, lddc and ldb multiples of 4, dC
 * and B 16-byte aligned): 64 x 64 tiles of A with at least 64 set bits are
 * computed with both operands staged in LDS, the other bits by a wave per
 * row; without it every row takes the wave-per-row kernel. */
size_t bws_sddmm_workspace_size(const bws_bitmap* A);
bws_status bws_sddmm(const bws_bitmap* A, const void* dC, const void* B, int32_t dtype,
                     int64_t lddc, int64_t ldb, int64_t N, float* dA_values,
                     int32_t precision, void* workspace, size_t workspace_bytes,
                     bws_stream stream);

/* Grouped launch over many bitmap problems: C_i = A_i · B_i (alpha = 1,
 * beta = 0) for every problem, with one kernel launch per tile configuration.
 * bws_batch_prepare (host, synchronous: call once per problem set and never
 * inside graph capture) validates the HOST array `problems`, picks each
 * problem's tile configuration (the plan's, or the autotuner's cached choice
 * for the problem's shape class, or the default), and writes the device
 * problem table and tile offsets into `workspace` on `stream` (after the work
 * already queued there, which may still use that memory; prepare returns
 * when the copies are done).  bws_batch_run enqueues the
 * launches on `stream` and is capturable into a hipGraph; the workspace must
 * stay untouched between the two.  Workgroups find their problem by a binary
 * search over the tile offsets, so problems of very different sizes share a
 * launch without padding. */
typedef struct {
  bws_bitmap A;      /* M x K                                     */
  const void* B;     /* K x N, dtype of the launch                */
  int64_t ldb;
  int64_t N;
  float* C;          /* M x N FP32                                */
  int64_t ldc;
  int64_t nnz_hint;  /* set bits of A if known (autotuner class), else 0 */
} bws_problem;

/* Host-side description of a prepared batch (filled by bws_batch_prepare). */
#define BWS_BATCH_MAX_LAUNCH 16
typedef struct {
  int32_t kind;          /* 1 grouped vector-add, 2 grouped MFMA, 3 per-problem launches */
  int32_t rows_per_wg;   /* workgroup tile rows (kinds 1, 2)                             */
  int32_t cols_per_wg;   /* workgroup tile columns (kind 1: 256; kind 2: 128 or 256)     */
  int32_t waves_per_wg;
  int32_t k_stage;
  int32_t values;        /* kind 1: the launch's masks carry values                     */
  int32_t vec;           /* kind 2: 1 = 16-byte B loads for every problem, 2 = and B
                            staged straight into LDS (16-bit B, N % 8 == 0)              */
  int32_t first;         /* first table entry of the launch                              */
  int32_t count;         /* entries (kinds 1, 2) or problems (kind 3)                    */
  int64_t total_tiles;   /* workgroup tiles of the launch                                */
} bws_batch_launch;

typedef struct {
  int32_t count;
  int32_t b_dtype;
  int32_t precision;
  int32_t n_launch;
  bws_batch_launch launch[BWS_BATCH_MAX_LAUNCH];
  void* workspace;                /* device problem table and tile offsets          */
  const bws_problem* problems;    /* the caller's array; kind 3 launches read it at run
                                     time, so it must outlive the batch             */
  bws_plan plan;                  /* the plan given to prepare (kernel 0 = none)    */
} bws_batch;

size_t bws_batch_workspace_size(int32_t count);
bws_status bws_batch_prepare(const bws_problem* problems, int32_t count, int32_t b_dtype,
                             int32_t precision, const bws_plan* plan, void* workspace,
                             size_t workspace_bytes, bws_batch* batch, bws_stream stream);
bws_status bws_batch_run(const bws_batch* batch, bws_stream stream);

/* Convenience: prepare + run in one call (not capturable: prepare copies). */
size_t bws_spmm_batched_workspace_size(int32_t count);
bws_status bws_spmm_batched(const bws_problem* problems, int32_t count, int32_t b_dtype,
                            int32_t precision, const bws_plan* plan, void* workspace,
                            size_t workspace_bytes, bws_stream stream);

/* Autotuner: times the candidate tile configurations of bws_spmm for this
 * problem on `stream` (HIP events; the call synchronises and overwrites C)
 * and returns the fastest in *best, also recording it in the process-wide
 * cache under the problem's shape class (M, K, N rounded up to powers of two,
 * log2 of the density rounded to an integer, precision).  bws_batch_prepare
 * consults the cache for problems with a nonzero nnz_hint when it is given no
 * plan; bws_spmm never does (its NULL plan means the built-in defaults).  nnz
 * is the bitmap's set-bit count (the caller knows it).
 * bws_autotune_save / _load write and read the cache as a text file. */
bws_status bws_autotune(const bws_bitmap* A, int64_t nnz, const void* B, int32_t b_dtype,
                        int64_t ldb, int64_t N, float* C, int64_t ldc, int32_t precision,
                        int32_t repeats, bws_plan* best, float* best_ms, bws_stream stream);
bws_status bws_autotune_lookup(int64_t M, int64_t K, int64_t N, int64_t nnz, int32_t precision,
                               bws_plan* plan);
bws_status bws_autotune_clear(void);
bws_status bws_autotune_save(const char* path);
bws_status bws_autotune_load(const char* path);

/* ------------------------------------------------------------------------ */
/* Sparse-operand phases of the reference pipeline                           */
/* ------------------------------------------------------------------------ */

/* Symbolic product of two bitmaps: WC (M x N) has bit (r, n) set iff some k
 * has bit (r, k) in WA and bit (k, n) in WB (the output mask of pre_spgemm,
 * source/src/transform.cuh:1028-1061).  Output rows are then grouped with
 * bws_group_rows on WC (the output-row grouping of transform.cuh:1065-1138). */
bws_status bws_symbolic_or(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa,
                           const uint32_t* WB, int64_t N, int64_t ldwb, uint32_t* WC,
                           int64_t ldwc, bws_stream stream);

/* Numeric sparse x sparse product onto an output pattern: for the j-th set
 * bit c of row r of C (C->row_ptr required), C_values[C->row_ptr[r] + j] =
 * alpha * Σ_k A[r, k]·B[k, c], A and B bitmaps (value-carrying, or 0/1 when
 * their values are NULL).  With C = bws_symbolic_or(A, B) this is the
 * numeric phase of the reference's sparse x sparse pipeline
 * (spgemm_compute_1dthread_tcore_v3, source/src/transform.cuh:2199-2246);
 * bws_group_rows on the result (C with these values) gives its grouped form,
 * the group values the reference stores in dC_group_value
 * (source/src/main.cu:2133).  FP32 FMA in ascending k for every entry.
 * Workspace: per-word popcount offsets of B and C. */
size_t bws_spgemm_workspace_size(int64_t M, int64_t K, int64_t N);
bws_status bws_spgemm(const bws_bitmap* A, const bws_bitmap* B, const bws_bitmap* C, float alpha,
                      float* C_values, void* workspace, size_t workspace_bytes,
                      bws_stream stream);

/* int8-valued A held as 8 packed bit-planes P (bws_pack_bitplanes_i8):
 * C = alpha * Σ_b w_b (P_b · B) + beta * C with w = [1,2,4,...,64,-128], the
 * reference's plane recombination (source/src/main.cu:889-912,
 * transform.cuh:2245).  With the workspace (bws_spmm_bitplanes_workspace_size)
 * and a NULL plan (or kernel 0) B is read once:
 *   FP32 with FP32 B: the planes become a value-carrying bitmap of the int8
 *     values (their OR as the mask), then one FP32 vector pass (FMA of each
 *     exact int8 value, FP32 accumulation: rounding bounded by |A|·|B|);
 *   BF16 / FP16 with 16-bit B (16-byte rows, N % 8 == 0): the planes are
 *     folded into a dense 16-bit copy of A (int8 values are exact in both),
 *     then one MFMA product;
 * otherwise eight weighted 0/1 products (FP32 rounding bounded by
 * Σ_b |w_b| (P_b |B|)). */
size_t bws_spmm_bitplanes_workspace_size(int64_t M, int64_t K, int32_t precision);
bws_status bws_spmm_bitplanes_i8(const uint32_t* P, int64_t M, int64_t K, int64_t ldw,
                                 const void* B, int32_t b_dtype, int64_t ldb, int64_t N, float* C,
                                 int64_t ldc, float alpha, float beta, int32_t precision,
                                 const bws_plan* plan, void* workspace, size_t workspace_bytes,
                                 bws_stream stream);

/* ------------------------------------------------------------------------ */
/* Host-side deterministic generators (no GPU; xoshiro256** seeded by        */
/* splitmix64).  Replace the reference's glibc rand() fill_random            */
/* (source/src/utils.h:14-22).                                               */
/* ------------------------------------------------------------------------ */

/* iid Bernoulli(p) bitmap, rows x cols, row-major draw order. */
bws_status bws_gen_bitmap_iid(uint64_t seed, int64_t rows, int64_t cols, double p,
                              uint32_t* W, int64_t ldw);
/* Clustered bitmap: n_proto prototypes ~ iid Bernoulli(p_proto) over cols; each
 * row copies prototype c ~ U{0..n_proto-1}, flips every bit with probability
 * flip, then the rows are shuffled by a uniform random permutation. */
bws_status bws_gen_bitmap_clustered(uint64_t seed, int64_t rows, int64_t cols,
                                    int32_t n_proto, double p_proto, double flip,
                                    uint32_t* W, int64_t ldw);
/* FP32 fills: kind 0 = N(a, b^2), 1 = U[a, b), 2 = random sign * 10^U[a, b). */
bws_status bws_gen_f32(uint64_t seed, int64_t n, int32_t kind, double a, double b,
                       float* out);

#ifdef __cplusplus
}
#endif

#endif /* BWS_H */

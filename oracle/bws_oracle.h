/*
 * bws_oracle.h — CPU restatement of the reference's bit-wise SpGEMM path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product (libbws.so) never links,
 * loads or calls it.
 *
 * Formats are those of include/bws.h: MSB-first uint32 bitmaps [rows x ldw],
 * optional row_ptr[rows+1] + values[nnz] in (row, column) order.
 * Every function cites the reference lines it restates.
 */
#ifndef BWS_ORACLE_H
#define BWS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* --- layout transforms ------------------------------------------------- */
void oref_pack_bitmap_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                          uint32_t* W, int64_t ldw);
void oref_unpack_bitmap_f32(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                            float* D, int64_t ld);
void oref_unpack_bitplanes_f32(const float* src, int64_t n, int32_t bit_width, int32_t* dst);
void oref_unpack_bitplanes_u8(const uint8_t* src, int64_t n, int32_t* dst);
void oref_pack_bitplanes_i8(const int8_t* D, int64_t rows, int64_t cols, int64_t ld,
                            uint32_t* P, int64_t ldw);
void oref_bitplane_weights(float* w8);
void oref_bitplane_row_counts(const uint32_t* P, int32_t planes, int64_t rows, int64_t cols,
                              int64_t ldw, int32_t* counts);
void oref_row_ptr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw, int64_t* row_ptr);
void oref_pack_values_f32(const float* D, int64_t rows, int64_t cols, int64_t ld,
                          const uint32_t* W, int64_t ldw, const int64_t* row_ptr, float* values);
void oref_transpose_bitmap(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                           uint32_t* WT, int64_t ldwt);
void oref_transpose_values(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                           const int64_t* row_ptr, const float* values, uint32_t* WT,
                           int64_t ldwt, int64_t* row_ptr_T, float* values_T);

/* --- grouping ---------------------------------------------------------- */
int64_t oref_group_tiles(int64_t rows, int64_t cols, int32_t tile_rows);
void oref_group_rows(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                     const int64_t* row_ptr, const float* values, int32_t tile_rows,
                     int32_t max_groups, int32_t* group_id, uint32_t* group_mask,
                     float* group_vals, int32_t* member_row, int32_t* spill_cnt,
                     int32_t* spill_nnz, int32_t* spill_table, int32_t* spill_rev,
                     int64_t* grouped_nnz);
void oref_group_spilled_csr(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                            const int64_t* row_ptr, const float* values, int32_t tile_rows,
                            const int32_t* spill_cnt, const int32_t* spill_nnz,
                            const int32_t* spill_table, int32_t* spill_row_off,
                            int32_t* spill_nnz_off, int32_t* row_end, int32_t* col, float* val);
void oref_group_scatter(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                        const int32_t* group_id, const float* group_vals, int32_t tile_rows,
                        int32_t max_groups, float* D, int64_t ld);
void oref_group_gather(const float* D, int64_t ld, const uint32_t* W, int64_t rows,
                       int64_t cols, int64_t ldw, const int32_t* group_id, int32_t tile_rows,
                       int32_t max_groups, float* group_vals);
void oref_symbolic_or(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa,
                      const uint32_t* WB, int64_t N, int64_t ldwb, uint32_t* WC, int64_t ldwc);

/* --- row clustering (an analysis step of this library; no reference
 *     counterpart beyond the intent of grouping rows by bitmask) ---------- */
int32_t oref_cluster_leaders(int64_t rows, int32_t max_clusters);
int64_t oref_cluster_candidates(int64_t rows, int32_t L);
int32_t oref_cluster_exact_rounds(int32_t L);
void oref_cluster_rows(const uint32_t* W, int64_t rows, int64_t cols, int64_t ldw,
                       int32_t max_clusters, int32_t* slot_row, uint32_t* chunk_mask);

/* --- numerics ---------------------------------------------------------- */
void oref_spgemm_f64(const uint32_t* WA, int64_t M, int64_t K, int64_t ldwa, const int64_t* rpA,
                     const float* vA, const uint32_t* WB, int64_t N, int64_t ldwb,
                     const int64_t* rpB, const float* vB, const uint32_t* WC, int64_t ldwc,
                     const int64_t* rpC, double* out, int32_t absolute);
void oref_spmm_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw, const int64_t* row_ptr,
                   const float* values, const float* B, int64_t ldb, int64_t N, double* C,
                   int64_t ldc, int32_t absolute);
void oref_spmm_f32_seq(const uint32_t* W, int64_t M, int64_t K, int64_t ldw,
                       const int64_t* row_ptr, const float* values, const float* B, int64_t ldb,
                       int64_t N, float* C, int64_t ldc);
void oref_spmm_groups_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw,
                          const int64_t* row_ptr, const float* values, const int32_t* group_id,
                          const float* group_vals, int32_t tile_rows, int32_t max_groups,
                          const float* B, int64_t ldb, int64_t N, double* C, int64_t ldc);
void oref_sddmm_f64(const uint32_t* W, int64_t M, int64_t K, int64_t ldw, const int64_t* row_ptr,
                    const float* dC, int64_t lddc, const float* B, int64_t ldb, int64_t N,
                    double* dA, int32_t absolute);
void oref_check_dense_f32(const float* A, const float* B, float* C, int64_t m, int64_t n,
                          int64_t k, int64_t row_begin, int64_t row_end, int32_t threads);
void oref_err_metrics(const float* C, int64_t ldc, const double* R, const double* S,
                      int64_t M, int64_t N, double* out);

/* --- 16-bit rounding (round to nearest even) ---------------------------- */
float oref_round_bf16(float x);
float oref_round_f16(float x);
void oref_round_bf16_array(const float* x, int64_t n, float* y);
void oref_round_f16_array(const float* x, int64_t n, float* y);
void oref_split_bf16x2(const float* x, int64_t n, float* hi, float* lo);

#ifdef __cplusplus
}
#endif

#endif /* BWS_ORACLE_H */

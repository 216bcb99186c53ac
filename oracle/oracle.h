/*
 * oracle.h — CPU restatement of AccSpMM's host algorithms, for TESTING ONLY.
 *
 * This is test infrastructure: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or the
 * timed CPU baseline.  The product (libaccspmm.so) never links or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * the AccSpMM repository root).  Pinning (what checks each function):
 *   orc_read_mtx         hand-worked MTX known-answer tests (banner, symmetric
 *                        expansion order, pattern, comments, 1-based shift)
 *   orc_coo_to_csr       stable counting sort: known answers + brute force
 *   orc_metcf_build      hand-worked METCF known answers (RW 8 and 16),
 *                        decode round trip METCF -> CSR, popcount invariants
 *   orc_adpbme_build     hand-worked group boundaries, maximality invariant
 *   orc_ibd              hand-worked IBD values incl. last-window exclusion
 *   orc_spmm_f64         brute-force dense product, scipy.sparse cross-check
 *   orc_spmm_f32_omp     CPU baseline; checked against orc_spmm_f64
 */
#ifndef ACCSPMM_ORACLE_H_
#define ACCSPMM_ORACLE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* include/mmio.hpp:171-210 (+ :24-93, :106-168).  Returns the reference's
 * codes: 0 success, -1 open failure, -2 banner error, 3 size-line failure.
 * Arrays are malloc'd; free with orc_free. */
int orc_read_mtx(const char* path, uint32_t* rows, uint32_t* cols,
                 uint32_t* nnz, uint32_t** r, uint32_t** c, float** v);
void orc_free(void* p);

/* include/format.hpp:110-142: stable counting sort by row. */
void orc_coo_to_csr(uint32_t rows, uint32_t nnz, const uint32_t* r,
                    const uint32_t* c, const float* v, uint32_t* row_ptr,
                    uint32_t* col_idx, float* val);

/* METCF result; arrays malloc'd, free with orc_metcf_free. */
typedef struct orc_metcf {
  uint32_t rw;           /* row window (8 or 16)                     */
  uint32_t num_windows;  /* W                                        */
  uint32_t num_tiles;    /* T                                        */
  uint32_t nnz;
  uint32_t* rowWindowOffset; /* W+1 */
  uint32_t* tcOffset;        /* T+1 */
  uint64_t* tcLocalBit;      /* T * rw/8 */
  uint32_t* sparseA2B;       /* 8T */
  float* data;               /* nnz */
  /* AdpBME extras (filled by orc_adpbme_build) */
  uint32_t num_groups;
  uint32_t* groupOffset;     /* G+1 */
  uint32_t* rowIndices;      /* T */
} orc_metcf;

/* include/format.hpp:215-274 (METCFBit::convertFromCSR) with
 * ROW_WINDOW = rw, COL_WINDOW = 8; column condensing per
 * include/format.hpp:179-199 and padding per :23-30. */
orc_metcf* orc_metcf_build(uint32_t rw, uint32_t rows, const uint32_t* row_ptr,
                           const uint32_t* col_idx, const float* val);
/* include/format.hpp:340-429 (AdpBME::CSR2AdpBME). */
orc_metcf* orc_adpbme_build(uint32_t rw, uint32_t rows,
                            const uint32_t* row_ptr, const uint32_t* col_idx,
                            const float* val, uint32_t t_load,
                            uint32_t t_comp, uint32_t t_write,
                            uint32_t target_num);
void orc_metcf_free(orc_metcf* m);

/* include/utils.hpp:75-85 (balanceStrategy): returns IBD; *balanced set to
 * IBD > threshold. */
uint32_t orc_ibd(const uint32_t* rowWindowOffset, uint32_t num_windows,
                 uint32_t num_tiles, uint32_t threshold, int* balanced);

/* Decodes a METCF back to CSR triples (row, col, val) in tile order. */
uint32_t orc_metcf_decode(const orc_metcf* m, uint32_t* r, uint32_t* c,
                          float* v);

/* C64[rows x K] = A * B (B rows x K row-major, ldb), FP64 accumulation. */
void orc_spmm_f64(uint32_t rows, const uint32_t* row_ptr,
                  const uint32_t* col_idx, const float* val, const float* B,
                  uint32_t ldb, uint32_t K, double* C);
/* include/ptx_tf32.hpp:19-29: the reference rounds A and B to TF32 before
 * each mma (cvt.rna.tf32.f32 in the PTX ISA: 10 explicit mantissa bits,
 * round to nearest, ties away from zero), keeping an FP32 container.
 * Infinity and NaN pass through; a carry out of the mantissa raises the
 * exponent (up to infinity). */
float orc_tf32_rna(float x);
/* C64 = tf32(A) * tf32(B) with FP64 accumulation: the input rounding of the
 * reference's TF32 kernels with their accumulation taken as exact, i.e. the
 * reference's precision class (for context beside the GPU forms' errors). */
void orc_spmm_tf32_f64(uint32_t rows, const uint32_t* row_ptr,
                       const uint32_t* col_idx, const float* val, const float* B,
                       uint32_t ldb, uint32_t K, double* C);
/* |A|*|B| (FP64), for the componentwise error metric. */
void orc_spmm_abs_f64(uint32_t rows, const uint32_t* row_ptr,
                      const uint32_t* col_idx, const float* val,
                      const float* B, uint32_t ldb, uint32_t K, double* C);
/* Error metrics: max_rel = max |C - C64| / (|A||B|) (entries with
 * (|A||B|)==0 must have C==0 exactly, else max_rel = +inf), fro_rel =
 * ||C-C64||_F / ||C64||_F.  C has row stride ldc. */
void orc_errors(uint64_t count_rows, uint32_t K, const float* C, uint32_t ldc,
                const double* C64, const double* absAB, double* max_rel,
                double* fro_rel);

/* CPU baseline: OpenMP CSR SpMM in FP32, rows scheduled dynamic(64), inner
 * loop over K vectorisable.  threads <= 0 uses all.  Returns threads used. */
int orc_spmm_f32_omp(uint32_t rows, const uint32_t* row_ptr,
                     const uint32_t* col_idx, const float* val,
                     const float* B, uint32_t ldb, uint32_t K, float* C,
                     int threads);

#ifdef __cplusplus
}
#endif
#endif

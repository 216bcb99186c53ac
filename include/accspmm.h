/*
 * accspmm.h — C ABI of the MI355X (gfx950) METCF SpMM library `libaccspmm.so`.
 *
 * C = A * B with A sparse (CSR, converted on the GPU to the METCF bitmap-tile
 * format) and B dense row-major N x K.  Every entry point returns an
 * accspmm_status and never aborts or prints.  Pointers documented as "device"
 * are HIP device pointers; `stream` is a hipStream_t passed as void* (NULL is
 * the null stream).
 *
 * The reference (AI4SClab/AccSpMM) has no library or FFI layer; its seams are
 * the `./mma` CLI, the format.hpp construction API and the TF32Compute
 * launcher.  Each declaration below names the reference interface it replaces.
 */
#ifndef ACCSPMM_H_
#define ACCSPMM_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference prints and continues (include/utils.hpp:6-13)
 * and leaves launches unchecked (include/tf32_comp.hpp:36,42). */
typedef enum accspmm_status {
  ACCSPMM_OK = 0,
  ACCSPMM_ERR_INVALID_ARG = 1,    /* null pointer, K == 0, bad option value   */
  ACCSPMM_ERR_NOT_CANONICAL = 2,  /* CSR row not strictly increasing in col   */
  ACCSPMM_ERR_HIP = 3,            /* a HIP runtime call failed                */
  ACCSPMM_ERR_OUT_OF_MEMORY = 4,  /* device or host allocation failed (or a
                                     host thread could not start); no C++
                                     exception leaves an entry point         */
  ACCSPMM_ERR_UNSUPPORTED = 5,    /* option outside the supported set         */
  ACCSPMM_ERR_BUDGET = 6,         /* serving: HBM budget exhausted            */
  ACCSPMM_ERR_NOT_FOUND = 7,      /* serving: unknown graph or request id     */
  ACCSPMM_ERR_QUEUE_FULL = 8,     /* serving: request queue full              */
  ACCSPMM_ERR_MTX = 9,            /* Matrix Market parse error                */
  ACCSPMM_ERR_NOT_READY = 10,     /* serving: request still in flight         */
  ACCSPMM_ERR_SHAPE = 11,         /* operand shapes do not agree              */
  ACCSPMM_ERR_FILE = 12,          /* file unwritable; METCF file unreadable,
                                     corrupt or of another version           */
  ACCSPMM_ERR_SCRATCH = 13        /* accspmm_spmm_noalloc: the call needs more
                                     than ACCSPMM_NOALLOC_SCRATCH_BYTES      */
} accspmm_status;

const char* accspmm_status_string(accspmm_status s);
/* Library version string, e.g. "accspmm-mi355x 0.1 gfx950". */
const char* accspmm_version(void);

/* ------------------------------------------------------------------------ */
/* Inputs                                                                   */
/* ------------------------------------------------------------------------ */

/* CSR view (include/format.hpp:96-142: row_ptr/col_idx/data, 32-bit). */
typedef struct accspmm_csr_view {
  uint32_t rows, cols, nnz;
  const uint32_t* row_ptr; /* rows + 1 */
  const uint32_t* col_idx; /* nnz      */
  const float* val;        /* nnz      */
} accspmm_csr_view;

/* Build options.  Defaults are the reference's compile-time constants:
 * ROW_WINDOW 8, COL_WINDOW 8 (include/common.hpp:46,48), TARGET_NUM 16
 * (:55), IMBALANCE_THRESHOLD 8 (:62) and the driver's CSR2AdpBME(csr,3,1,2)
 * cost model (src/mma_tf32.cu:222). */
typedef struct accspmm_build_opts {
  uint32_t row_window;          /* 8 or 16 (16 => 128-bit tiles); 0: 16 when
                                   rows average >= 128 nonzeros, or when a
                                   graph of >= 64*8192 8-row tiles has <= 92%
                                   as many 16-row ones (banded), else 8     */
  uint32_t col_window;          /* must be 8                                */
  uint32_t t_load, t_comp, t_write, target_num; /* AdpBME cost model        */
  uint32_t imbalance_threshold; /* balanceStrategy threshold                */
  uint32_t build_adpbme;        /* 1: also build AdpBME groups for export
                                   (default 0: no kernel reads them; the
                                   AdpBME drop-in class sets 1)           */
  /* 0 (default): rows kept in input order, the reference's format.  1: the
   * format is built from the rows in order of decreasing length (stable),
   * so 8-row windows of like rows share more columns: fewer tiles and fewer
   * B rows gathered.  C is still written in input row order.  No reference
   * counterpart (the reference reads an externally reordered file, named
   * ddi.my_reorder.mtx in its tree).  2: reorder when rows average >= 128
   * nonzeros (measured faster there, slower on short-row graphs). */
  uint32_t reorder_rows;
  /* Locality ordering of a square matrix, applied to rows and columns alike
   * (A' = P A P^T; reorder_rows is then ignored): 0 (default) none; 1
   * reverse Cuthill-McKee (host breadth-first search, accspmm_order_rcm):
   * banded and mesh-like graphs stored in a scattered order get their band
   * back, so neighbouring row windows gather neighbouring B rows; 2 rows
   * and columns by decreasing row length (hub rows and hub columns first).
   * B is permuted the same way inside each SpMM call (by the conversion
   * pass, or a permuted copy for the forms that gather FP32 B) and C is
   * written in input row order, so results keep their input layout.  The
   * exported arrays describe P A P^T (accspmm_metcf_row_order /
   * _col_order give P).  No reference counterpart (the reference reads an
   * externally reordered file, ddi.my_reorder.mtx in its tree).
   * ACCSPMM_ERR_UNSUPPORTED for a non-square matrix. */
  uint32_t symmetric_order;
} accspmm_build_opts;

void accspmm_build_opts_default(accspmm_build_opts* o);

/* Opaque device-resident METCF(+AdpBME) graph.  Its format arrays are
 * immutable after build.  accspmm_spmm keeps one workspace (split-window
 * partials, the BF16 form's converted B) per stream inside the handle, so
 * calls on different streams may share a handle concurrently; calls on one
 * stream are ordered by the stream.  Host threads may call concurrently (an
 * internal lock serialises the enqueue).  Destroy only after every stream
 * that used the handle is idle. */
typedef struct accspmm_metcf accspmm_metcf;

typedef struct accspmm_metcf_info {
  uint32_t rows, cols, nnz;
  uint32_t row_window;
  uint32_t num_windows;   /* rowWindowOffset.size() - 1               */
  uint32_t num_tiles;     /* tcOffset.size() - 1                      */
  uint32_t num_groups;    /* AdpBME groupOffset.size() - 1 (0 if none) */
  uint32_t ibd;           /* imbalance degree (balanceStrategy)       */
  uint32_t balanced;      /* ibd > imbalance_threshold                */
  uint32_t max_window_tiles;
  uint64_t device_bytes;  /* device memory held by the handle         */
  uint32_t reordered;     /* built with reorder_rows = 1              */
  /* values of the fullest 4-tile group (groups start at each window's
   * first tile); with 16-row windows, groups of more than 256 values
   * (tiles over half full) take a slower two-span load in the kernels */
  uint32_t max_group_values;
  uint32_t symmetric_order; /* the build's symmetric_order (0: none)      */
} accspmm_metcf_info;

/* GPU METCF builder.  Replaces METCFBit::convertFromCSR
 * (include/format.hpp:215-274), AdpBME::CSR2AdpBME (:340-429) and
 * balanceStrategy (include/utils.hpp:75-85).  `csr` holds DEVICE pointers to
 * a canonical CSR (each row strictly increasing in column); a
 * non-canonical row yields ACCSPMM_ERR_NOT_CANONICAL; more than
 * 2^32 - 1024 nonzeros (value offsets are 32-bit, as the reference's vint)
 * yields ACCSPMM_ERR_UNSUPPORTED.  Blocks until the handle is complete.
 * The handle owns its device arrays. */
accspmm_status accspmm_metcf_build(const accspmm_csr_view* csr,
                                   const accspmm_build_opts* opts,
                                   accspmm_metcf** out, void* stream);
/* Same as accspmm_metcf_build, from a HOST CSR (uploaded by the library on
 * the null stream).  Used by the format.hpp drop-in classes. */
accspmm_status accspmm_metcf_build_host(const accspmm_csr_view* host_csr,
                                        const accspmm_build_opts* opts,
                                        accspmm_metcf** out);
/* Makes a DEVICE CSR canonical in place: each row sorted by column (stable),
 * duplicate columns summed in their input order in float32, rows compacted,
 * row_ptr rewritten; *nnz_out receives the new count (col_idx/val keep their
 * length, entries past *nnz_out are unspecified).  The result equals
 * accspmm_read_mtx's host canonicalisation bit for bit.  ERR_INVALID_ARG if
 * a column is >= cols or row_ptr is not non-decreasing from 0 to nnz.
 * Blocks until done.  No reference counterpart: the reference hands its
 * builder rows as they come (include/format.hpp:110-142, 238-269), which is
 * right only for sorted, duplicate-free rows; this makes device-resident
 * graphs valid input for accspmm_metcf_build. */
accspmm_status accspmm_csr_canonicalize(uint32_t rows, uint32_t cols, uint32_t nnz,
                                        uint32_t* row_ptr, uint32_t* col_idx, float* val,
                                        uint32_t* nnz_out, void* stream);
accspmm_status accspmm_metcf_destroy(accspmm_metcf* h);
accspmm_status accspmm_metcf_get_info(const accspmm_metcf* h,
                                      accspmm_metcf_info* info);

/* Byte-exact export of the format arrays to HOST buffers laid out as the
 * public std::vector members of METCFBit (include/format.hpp:209-211,
 * :175-176) and AdpBME (:333-336).  Any pointer may be NULL (skipped).
 * Sizes: rowWindowOffset W+1, tcOffset T+1, tcLocalBit T*(row_window/8),
 * sparseA2B 8T, data nnz, groupOffset G+1, rowIndices T. */
accspmm_status accspmm_metcf_export(const accspmm_metcf* h,
                                    uint32_t* rowWindowOffset,
                                    uint32_t* tcOffset, uint64_t* tcLocalBit,
                                    uint32_t* sparseA2B, float* data,
                                    uint32_t* groupOffset,
                                    uint32_t* rowIndices);

/* Row order of the format: row i of the exported arrays is input row
 * row_order[i] (HOST buffer of `rows` entries; the identity unless built
 * with reorder_rows = 1).  No reference counterpart. */
accspmm_status accspmm_metcf_row_order(const accspmm_metcf* h, uint32_t* row_order);

/* Column order of the format: format column j is input column col_order[j]
 * (HOST buffer of `cols` entries; the identity unless built with a
 * symmetric_order, then equal to the row order).  No reference
 * counterpart. */
accspmm_status accspmm_metcf_col_order(const accspmm_metcf* h, uint32_t* col_order);

/* Binary save / load of a built graph (no reference counterpart: the
 * reference rebuilds the format from the .mtx file on every run,
 * src/mma_tf32.cu:195-226).  The file holds the build options, the arrays
 * accspmm_metcf_export returns and a checksum; load validates all of it
 * (ACCSPMM_ERR_FILE otherwise) and gives a handle equal to the saved one. */
accspmm_status accspmm_metcf_save(const accspmm_metcf* h, const char* path);
accspmm_status accspmm_metcf_load(const char* path, accspmm_metcf** out,
                                  void* stream);

/* ------------------------------------------------------------------------ */
/* SpMM                                                                     */
/* ------------------------------------------------------------------------ */

typedef enum accspmm_form {
  ACCSPMM_FORM_FP32 = 0,       /* FP32 MFMA (v_mfma_f32_16x16x4_f32)           */
  ACCSPMM_FORM_SPLIT_BF16 = 1, /* A split exactly into 3 BF16 terms, B into 2
                                  (remainder <= 2^-17 |b|), the 5 products
                                  A_i B_j (j <= 1), FP32 accumulate: each
                                  product within 2^-17 + 2^-25, <= 1e-5
                                  max-rel                                    */
  ACCSPMM_FORM_BF16 = 2        /* A,B rounded to BF16, FP32 accumulate        */
} accspmm_form;

/* Kernel tile configuration (all zero => library default / tuned). */
typedef struct accspmm_tile {
  uint32_t feat_per_lane;  /* NB: features per lane per B row: 1,2,4,8; KC=16*NB */
  uint32_t waves_per_block;/* 1..8                                           */
  uint32_t split_tiles;    /* windows above this many tiles are split (0: no)  */
  uint32_t xcd_slice;      /* XCD mapping of work: 0 round-robin, 1 pin
                              feature chunks to XCDs, 2 one contiguous range
                              of work units per XCD, 3 pin column panels to
                              XCDs (panel p on XCD p % 8, so each B row is
                              cached in one XCD's L2; needs col_panels > 1,
                              else 0)                                        */
  uint32_t pack_tiles;     /* consecutive small windows are packed into one
                              work unit of at least this many tiles (0: one
                              window per unit)                               */
  uint32_t col_panels;     /* > 1: cut long windows at this many equal column
                              ranges of B and run the pieces range by range,
                              so each range of B is gathered while it is
                              cached (0 or 1: off)                           */
  uint32_t sweep;          /* 1: resident-window sweep: one block per CU, each
                              wave owns a set of windows whose running sums
                              stay in LDS and runs their pieces panel by
                              panel (same pieces and result bits as 0; the
                              work-unit mapping fields are ignored)          */
  uint32_t b_split;        /* split-BF16 form, how B is split into bf16 hi and
                              lo: 1 a conversion pass writes a split copy of B
                              (4 bytes an element) that the kernel gathers;
                              2 the kernel gathers the FP32 B and splits it in
                              registers (no copy, the same C bits); 0 the
                              library's choice.  Multi-problem batches always
                              convert.                                     */
} accspmm_tile;

/* C[rows x K] = A * B[cols x K].  B (ldb >= K) and C (ldc >= K) are
 * caller-owned DEVICE FP32 buffers; C is written exactly (rows x K, no
 * padding rows).  Replaces TF32Compute::compute / adpBalanceCompute
 * (include/tf32_comp.hpp:52-113) and the four kernels of
 * include/mma_tf32.cuh.  The BF16 form converts B into the handle's
 * workspace for `stream` (allocated on first use, grown when a call needs
 * more, reused after), so the conversion is part of each call; split
 * windows write their partial sums there too.  Asynchronous on `stream`. */
accspmm_status accspmm_spmm(const accspmm_metcf* h, const float* B,
                            uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                            accspmm_form form, const accspmm_tile* tile,
                            void* stream);

/* The two halves of accspmm_spmm, for callers that time or overlap them:
 * accspmm_spmm_prepare_b converts B into the handle's workspace for
 * `stream` (BF16 form; a no-op for the others), accspmm_spmm_prepared runs
 * the SpMM kernel on it.  Both must be given the same B, K, form, tile and
 * stream. */
accspmm_status accspmm_spmm_prepare_b(const accspmm_metcf* h, const float* B,
                                      uint32_t ldb, uint32_t K,
                                      accspmm_form form,
                                      const accspmm_tile* tile, void* stream);
accspmm_status accspmm_spmm_prepared(const accspmm_metcf* h, const float* B,
                                     uint32_t ldb, float* C, uint32_t ldc,
                                     uint32_t K, accspmm_form form,
                                     const accspmm_tile* tile, void* stream);

/* Bytes of the per-GPU scratch array of accspmm_spmm_noalloc (64 MiB). */
#define ACCSPMM_NOALLOC_SCRATCH_BYTES (64ull << 20)

/* accspmm_spmm that allocates nothing: no hipMalloc, hipFree or other
 * allocating routine, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).  No
 * counterpart in the reference.  Same operands, tile choice, kernels and C
 * bits as accspmm_spmm; the call's device buffers (split-window partials,
 * the BF16 form's converted B and, when the handle holds no device schedule
 * for the tile yet, the work units) live in one file-scope device array of
 * ACCSPMM_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.  A
 * call that needs more returns ACCSPMM_ERR_SCRATCH (choose a tile with fewer
 * split windows and column panels); tile.sweep is refused with
 * ACCSPMM_ERR_UNSUPPORTED.  The first call for a tile on a handle without a
 * device schedule plans it on the host (one stream synchronisation, host
 * memory kept in the handle); later calls copy that plan into the array.
 * All calls on one GPU share the array and keep the last call's contents:
 * calls that overlap on one GPU race, so order them (one stream, or events
 * between streams).  Asynchronous on `stream` otherwise. */
accspmm_status accspmm_spmm_noalloc(const accspmm_metcf* h, const float* B,
                                    uint32_t ldb, float* C, uint32_t ldc,
                                    uint32_t K, accspmm_form form,
                                    const accspmm_tile* tile, void* stream);

/* Chooses the tile configuration the library would use for (h, K, form):
 * the autotuner's cached pick if one exists for this degree histogram and
 * K, else the heuristic default. */
accspmm_status accspmm_tile_select(const accspmm_metcf* h, uint32_t K,
                                   accspmm_form form, accspmm_tile* out);

/* Autotuner: times each candidate tile configuration on (h, B, C, K, form)
 * on `stream`, stores the fastest under the handle's degree-histogram key
 * and K (in memory, and in the JSON file `cache_path` if non-NULL: the
 * file's other entries are kept), and returns it.  B and C are scratch
 * operands of the right shape.  A process reads the file named by the
 * environment variable ACCSPMM_TUNE_CACHE at its first tile choice;
 * entries that are not launchable tiles are ignored. */
accspmm_status accspmm_autotune(const accspmm_metcf* h, const float* B,
                                uint32_t ldb, float* C, uint32_t ldc,
                                uint32_t K, accspmm_form form,
                                const char* cache_path, accspmm_tile* best,
                                float* best_ms, void* stream);
/* Layout autotuner: builds the DEVICE CSR in each layout (row window 8 and
 * 16; rows in input order, reordered by length, and for a square matrix
 * the reverse Cuthill-McKee symmetric order), times the SpMM of each
 * (library tile, B conversion included, median of 5 launches) on (B, C, K,
 * form) on `stream`, and returns the fastest layout as build options (the
 * other fields are the defaults) with its time.  Costs four builds (six
 * for a square matrix); no reference counterpart (the reference fixes
 * ROW_WINDOW at compile time, include/common.hpp:46). */
accspmm_status accspmm_autotune_layout(const accspmm_csr_view* csr, const float* B,
                                       uint32_t ldb, float* C, uint32_t ldc, uint32_t K,
                                       accspmm_form form, accspmm_build_opts* best,
                                       float* best_ms, void* stream);
/* Degree-histogram key of a handle (log2 buckets of tiles per window),
 * written as a NUL-terminated string of at most `cap` bytes. */
accspmm_status accspmm_tune_key(const accspmm_metcf* h, uint32_t K,
                                accspmm_form form, char* buf, size_t cap);

/* ------------------------------------------------------------------------ */
/* Batched: many (graph, B, C, K) problems in one persistent launch.        */
/* No reference counterpart (the reference runs one matrix per process).    */
/* ------------------------------------------------------------------------ */

typedef struct accspmm_problem {
  const accspmm_metcf* graph;
  const float* B; uint32_t ldb;
  float* C; uint32_t ldc;
  uint32_t K;
} accspmm_problem;

typedef struct accspmm_batch accspmm_batch;

/* Plans a batch (work-unit queue over all problems, one form); allocates its
 * device task list and workspaces.  `problems` may be re-bound later with
 * accspmm_batch_rebind as long as graphs and K stay the same.  Graphs of
 * both row-window sizes may be mixed: each size gets its own persistent
 * launch, enqueued one after the other. */
accspmm_status accspmm_batch_create(const accspmm_problem* problems,
                                    uint32_t count, accspmm_form form,
                                    accspmm_batch** out, void* stream);
accspmm_status accspmm_batch_rebind(accspmm_batch* b,
                                    const accspmm_problem* problems,
                                    uint32_t count, void* stream);
/* One persistent launch over the whole queue (plus the B conversion pass
 * for the BF16 forms).  Asynchronous on `stream`. */
accspmm_status accspmm_batch_run(accspmm_batch* b, void* stream);
/* Batch autotuner: times the batch (each run overwrites the problems' C)
 * under batch-wide choices on top of each problem's tile — the widest
 * feature chunk the persistent kernel runs (NB 4, 2, 1) and the packing of
 * short windows into work units (the tile's own, or at least 32, 128, 512
 * tiles) — and keeps the fastest; neither changes any C bit.  *best_ms
 * receives its time per run.  A captured graph of the batch is dropped (the
 * next accspmm_batch_run_graph captures again).  No reference counterpart. */
accspmm_status accspmm_batch_autotune(accspmm_batch* b, void* stream, float* best_ms);
/* Captures accspmm_batch_run into a hipGraph once and replays it. */
accspmm_status accspmm_batch_run_graph(accspmm_batch* b, void* stream);
accspmm_status accspmm_batch_destroy(accspmm_batch* b);
accspmm_status accspmm_batch_num_tasks(const accspmm_batch* b,
                                       uint64_t* tasks);

/* ------------------------------------------------------------------------ */
/* Host utilities (no GPU): Matrix Market ingest, CSR, generators.          */
/* ------------------------------------------------------------------------ */

/* Host CSR owned by the library (free with accspmm_host_csr_free). */
typedef struct accspmm_host_csr {
  uint32_t rows, cols, nnz;
  uint32_t* row_ptr;
  uint32_t* col_idx;
  float* val;
} accspmm_host_csr;

void accspmm_host_csr_free(accspmm_host_csr* c);

/* Reads a Matrix Market file with the semantics of read_from_mtx
 * (include/mmio.hpp:171-210: banner tokens, symmetric/Hermitian expansion
 * (r,c) then (c,r), pattern => 1.0) and builds the CSR by stable counting
 * sort (include/format.hpp:110-142).  If `canonicalize` is non-zero each row
 * is then sorted by column and duplicates are summed.  `mtx_code` receives
 * the reference's read_from_mtx code (0, -1 open failure, -2 banner).
 * Files with at least 1 MiB of entries are parsed line-wise on several
 * threads (ACCSPMM_HOST_THREADS, default: the usable CPUs, at most 16); a
 * file whose entries are not one per line falls back to the sequential
 * tokenizer, so the result never depends on the thread count. */
accspmm_status accspmm_read_mtx(const char* path, int canonicalize,
                                accspmm_host_csr* out, int* mtx_code);

/* Writes a Matrix Market file (coordinate real, values with 9 significant
 * digits, which read back to the same floats).  symmetric = 0: every entry,
 * row-major.  symmetric = 1: a square CSR holding a symmetric matrix is
 * written as its lower triangle in column-major order, the usual layout of
 * symmetric files (ERR_SHAPE if rows != cols).  No reference counterpart
 * (the reference only reads, include/mmio.hpp:171-210). */
accspmm_status accspmm_write_mtx(const accspmm_host_csr* csr, const char* path,
                                 int symmetric);

/* Threads that parsed the entries in this thread's last accspmm_read_mtx
 * call (1: the sequential tokenizer).  No reference counterpart: an ingest
 * diagnostic. */
int accspmm_read_mtx_threads_used(void);

/* 1 if every row is strictly increasing in column and all columns < cols. */
int accspmm_csr_is_canonical(const accspmm_host_csr* c);

/* Deterministic generators (xoshiro256** seeded by splitmix64). */
/* Symmetric Chung-Lu power-law graph (Miller-Hagberg skip sampling), a self
 * loop on every node, seeded random relabel, canonical rows.  value_mode:
 * 0 = symmetric U(0,1] per undirected edge, 1 = 1/deg(row), 2 = U[-1,1]. */
accspmm_status accspmm_gen_chung_lu(uint32_t n, double mean_degree,
                                    double exponent, uint64_t seed,
                                    int value_mode, accspmm_host_csr* out);
/* 27-point 3-D stencil on an n^3 grid (id = x + n*y + n*n*z), values
 * U[-1,1] in CSR order. */
accspmm_status accspmm_gen_stencil27(uint32_t n, uint64_t seed,
                                     accspmm_host_csr* out);
/* Dense row-major rows x cols, U[-1,1]. */
void accspmm_gen_dense(uint64_t seed, uint64_t count, float* out);

/* Gather ceiling (a measurement of the memory system, no reference
 * counterpart): `waves` waves (a multiple of 4) each read rows_per_wave (a
 * multiple of 64) rows idx[w * rows_per_wave + i] of the DEVICE region of
 * region_rows rows of row_bytes (256 or 512) bytes, in the SpMM kernels'
 * pattern (a quarter-wave per row, 16 bytes a lane, buffer loads), and
 * write one folded word per lane to out (waves * 64 words).  An index past
 * the region reads zeros (the buffer loads' range check).  Asynchronous on
 * `stream`; time it with events. */
accspmm_status accspmm_gather_probe(const void* region, uint32_t region_rows,
                                    uint32_t row_bytes, const uint32_t* idx,
                                    uint32_t waves, uint32_t rows_per_wave,
                                    uint32_t* out, void* stream);

/* Reverse Cuthill-McKee order of the n x n pattern (HOST arrays; row i's
 * columns are its neighbours): perm[i] = the node placed at position i.
 * Components are ordered from pseudo-peripheral start nodes, breadth
 * first, neighbours by increasing degree (ties: smaller index), then the
 * whole order is reversed.  Deterministic.  The builder's symmetric_order
 * 1 uses it.  ACCSPMM_ERR_INVALID_ARG for a bad row_ptr or a column >= n.
 * No reference counterpart. */
accspmm_status accspmm_order_rcm(uint32_t n, const uint32_t* row_ptr,
                                 const uint32_t* col_idx, uint32_t* perm);

/* ------------------------------------------------------------------------ */
/* Serving: resident graphs, HBM-budget admission, continuous batching.     */
/* No reference counterpart.                                                */
/* ------------------------------------------------------------------------ */

typedef struct accspmm_server accspmm_server;

typedef struct accspmm_server_opts {
  uint64_t hbm_budget_bytes;  /* resident graphs + in-flight workspaces     */
  uint32_t max_batch;         /* requests per persistent launch             */
  uint32_t max_queue;         /* pending requests before QUEUE_FULL         */
  uint32_t max_inflight;      /* launches in flight (streams)               */
  uint32_t form;              /* accspmm_form                               */
  uint32_t one_per_launch;    /* 1: baseline mode, one launch per request   */
  uint64_t large_work;        /* requests with nnz * K at or above this run
                                 in launches of their own, one at a time
                                 (0: 2^30); smaller ones are batched         */
} accspmm_server_opts;

void accspmm_server_opts_default(accspmm_server_opts* o);
accspmm_status accspmm_server_create(const accspmm_server_opts* opts,
                                     accspmm_server** out);
/* Registers a resident graph built from a device CSR; charges its bytes
 * against the budget (ACCSPMM_ERR_BUDGET if it does not fit). */
accspmm_status accspmm_server_register(accspmm_server* s,
                                       const accspmm_csr_view* csr,
                                       uint32_t* graph_id);
/* Same, built with the given options (row window, reorder_rows; NULL, and
 * accspmm_server_register: the automatic layout, row_window 0 and
 * reorder_rows 2; AdpBME groups are never built for served graphs). */
accspmm_status accspmm_server_register_opts(accspmm_server* s,
                                            const accspmm_csr_view* csr,
                                            const accspmm_build_opts* opts,
                                            uint32_t* graph_id);
/* Registers a resident graph saved by accspmm_metcf_save (warm start). */
accspmm_status accspmm_server_register_file(accspmm_server* s, const char* path,
                                            uint32_t* graph_id);
accspmm_status accspmm_server_unregister(accspmm_server* s, uint32_t graph_id);
/* Submits C = A_g * B (device buffers, caller-owned until completion). */
accspmm_status accspmm_server_submit(accspmm_server* s, uint32_t graph_id,
                                     const float* B, uint32_t ldb, float* C,
                                     uint32_t ldc, uint32_t K,
                                     uint64_t* request_id);
/* Non-blocking: ACCSPMM_OK when done (latency in seconds written, from
 * submit to completion), ACCSPMM_ERR_NOT_READY otherwise. */
accspmm_status accspmm_server_poll(accspmm_server* s, uint64_t request_id,
                                   double* latency_s);
/* Blocks until every submitted request has completed (pending requests of a
 * paused server are not waited for). */
accspmm_status accspmm_server_drain(accspmm_server* s);
/* Stops (paused != 0) or resumes admitting queued requests; launches
 * already on the GPU finish either way. */
accspmm_status accspmm_server_pause(accspmm_server* s, int paused);
typedef struct accspmm_server_stats {
  uint64_t submitted, completed, launches, rejected_budget;
  uint64_t resident_bytes, peak_inflight_bytes;
  double busy_s;
} accspmm_server_stats;
accspmm_status accspmm_server_get_stats(accspmm_server* s,
                                        accspmm_server_stats* st);
accspmm_status accspmm_server_destroy(accspmm_server* s);

#ifdef __cplusplus
}
#endif
#endif /* ACCSPMM_H_ */

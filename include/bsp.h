/*
 * bsp.h - C-ABI of the MI355X (gfx950) SpMV / block-SpMM library.
 *
 * Two layers live behind this header:
 *   1. A device-level API: device pointers, explicit HIP streams, status
 *      codes, no hidden host<->device copies. Formats (BSR, DBSR, DASP, ...)
 *      are built on the GPU into library-owned device buffers.
 *   2. The reference call surface (functions named bsp_<reference name>):
 *      host pointers in and out, `repeat` timed launches, average time in
 *      whole microseconds. Each one replaces the reference function cited
 *      beside it; include/benchsp_compat.hpp maps the reference's exact C++
 *      signatures (double& avg_time, CSRFormat<>, dbsrMat) onto them.
 *
 * Every function returns bsp_status (0 = success) unless it mirrors a
 * reference function returning int/void; none of them calls exit().
 * The message of the last error on the calling thread is bsp_last_error().
 *
 * Streams: a bsp_stream is a hipStream_t (NULL = the null stream).
 * Concurrency: handles are not internally synchronised; do not use one handle
 * from two host threads at once.
 */
#ifndef BSP_H_
#define BSP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum bsp_status {
  BSP_OK = 0,
  BSP_ERR_INVALID = 1,     /* bad argument: null pointer, negative size, ... */
  BSP_ERR_HIP = 2,         /* a HIP runtime call failed */
  BSP_ERR_ALLOC = 3,       /* device or host allocation failed */
  BSP_ERR_UNSORTED = 4,    /* CSR column indices not ascending within a row */
  BSP_ERR_UNSUPPORTED = 5, /* valid input this build does not handle */
  BSP_ERR_TOO_LARGE = 6    /* sizes beyond 32-bit index range or a fixed limit */
} bsp_status;

typedef enum bsp_dtype { BSP_F64 = 0, BSP_F32 = 1, BSP_F16 = 2, BSP_BF16 = 3 } bsp_dtype;

typedef void* bsp_stream; /* hipStream_t */

const char* bsp_status_string(bsp_status s);
const char* bsp_last_error(void);
/* Synchronous device-to-host copy of `bytes` (for reading format views). */
bsp_status bsp_copy_to_host(void* dst, const void* src, size_t bytes);
/* Library version and the gfx target its kernels were built for. */
const char* bsp_version(void);

/* ------------------------------------------------------------------------ */
/* Device CSR descriptor (all pointers are device memory, not owned).        */
/* rowptr has nrows+1 int32 entries; nrows, ncols and nnz are at most
 * 2^31 - 2^20 (BSP_ERR_TOO_LARGE beyond), which leaves in-kernel index
 * offsets room within int32. Columns ascending per row. */
typedef struct bsp_csr {
  int64_t nrows, ncols, nnz;
  const int32_t* rowptr;
  const int32_t* colidx;
  const void* val; /* element type `dtype` */
  bsp_dtype dtype;
} bsp_csr;

/* ------------------------------------------------------------------------ */
/* CSR SpMV (FP64).                                                          */
typedef enum bsp_csr_method {
  /* Library kernel: nnz-balanced row blocks staged through LDS, long rows
   * on whole workgroups. The default. */
  BSP_CSR_ADAPTIVE = 0,
  /* Thread per row. Replaces spmv_fp64 (src/csrspmv.cu:39-57). */
  BSP_CSR_SCALAR = 1,
  /* Thread per 16-nnz chunk, atomic-free two-pass combine. Replaces
   * spmv_fp64_balance (src/csrspmv.cu:59-102). */
  BSP_CSR_BALANCED = 2,
  /* Rows <= 128 nnz thread-per-row, longer rows wave-per-row. Replaces
   * MySpmvWithPrefp64 (src/spmv.cu:28-42,77-108,131-196,323-397). */
  BSP_CSR_SPLIT = 3
} bsp_csr_method;

typedef struct bsp_spmv_plan_s* bsp_spmv_plan;

/* Tuning of the adaptive method (zero-initialise for the defaults).
 * seg_cols: column-segment width. The adaptive plan regroups the entries of
 *   a matrix whose columns are scattered into column segments of this many
 *   columns (8*seg_cols bytes of x stay L2-resident per pass) and runs one
 *   pass per segment. 0 = automatic (segments of 2^19 columns when more
 *   than a quarter of the entries lie farther than 2^18 columns from the
 *   diagonal and rows are sorted), < 0 = never segment. */
typedef struct bsp_csr_options {
  int64_t seg_cols;
  /* adaptive block window in rows+entries, 128..2048 (others: 1024 for an
   * unsegmented plan, 2048 for a segmented one) */
  int32_t chunk;
  /* segmented plans: 0 = automatic (2, 4 or 8 segments run in one launch,
   * each pinned to a group of XCDs, partial results added by a second
   * kernel; automatic width then targets 2 MB of x per segment),
   * 1 = one launch per segment, later ones accumulating into y,
   * 2 = as 0, with each segment cut into blocks of fewer than 6,144 entries
   *   stored in column order (plus a 2-byte row-order slot per entry), so a
   *   wave's x gathers share cache lines; chunk is not used. Plans that do
   *   not segment into 2, 4 or 8 pieces run as with 0 / 1. */
  int32_t seg_mode;
  /* unsegmented plans: > 0 = use 16-bit column offsets where possible,
   * 0 (default) or < 0 = int32 indices. When every block's column indices
   * lie within int16 of its first row (banded matrices such as stencils)
   * and the values are 16-byte aligned, the plan then keeps 16-bit offsets
   * from each block's first row instead of the int32 indices (2 bytes per
   * entry less to read; y bit-identical) */
  int32_t narrow_cols;
  int32_t reserved32;
  int64_t reserved[5];
} bsp_csr_options;

/* Validates a device CSR on the GPU before it is handed to any plan or
 * builder (which trust their input): rowptr starts at 0, never decreases and
 * ends at nnz, every column index lies in [0, ncols). BSP_ERR_INVALID
 * otherwise. *sorted (may be NULL) = 1 when columns are strictly ascending
 * in every row, as the column-segmented plan needs. Synchronises the stream. */
bsp_status bsp_csr_check(const bsp_csr* A, bsp_stream s, int32_t* sorted);

/* Build the per-matrix schedule for `method` on the GPU (stream-ordered).
 * The plan keeps the descriptor's device pointers; they must outlive it
 * (a segmented adaptive plan also holds its own regrouped copy of col/val:
 * call bsp_csr_plan_create again after changing the values). */
bsp_status bsp_csr_plan_create(const bsp_csr* A, int method, bsp_stream s, bsp_spmv_plan* out);
bsp_status bsp_csr_plan_create_ex(const bsp_csr* A, int method, const bsp_csr_options* opt,
                                  bsp_stream s, bsp_spmv_plan* out);
/* Column-segment width of an adaptive plan (0 = not segmented) and the
 * number of segments. */
bsp_status bsp_csr_plan_segments(bsp_spmv_plan p, int64_t* seg_cols, int64_t* nseg);
/* 1 when an adaptive plan reads 16-bit column offsets (bsp_csr_options
 * narrow_cols), else 0. */
bsp_status bsp_csr_plan_narrow(bsp_spmv_plan p, int32_t* narrow);
/* Measurement aid, no counterpart in the reference: the data movement of an
 * XCD-pinned column-segmented adaptive plan without its arithmetic. Each
 * workgroup streams 2048 entries of its segment's column indices (and
 * values, with_values != 0) on the segment's XCDs and gathers x at each,
 * storing one sum per workgroup in sink. sink == NULL returns the number of
 * workgroups in *grid only. Stream-ordered, one launch. BSP_ERR_UNSUPPORTED
 * for other plans. */
bsp_status bsp_csr_plan_probe(bsp_spmv_plan p, int with_values, const double* x, double* sink,
                              int64_t sink_len, int64_t* grid, bsp_stream s);
/* Measurement aid, no counterpart in the reference: `gathers` random 8-byte
 * reads of x alone (no streams), x split into `slices` equal slices (1, 2, 4
 * or 8; a power-of-two length each), workgroup i gathering from slice
 * (i % 8) % slices, i.e. on the XCDs that own it, as a segmented plan's
 * launch does: the rate at which the L2 serves random gathers. One sum per
 * workgroup goes to sink; sink == NULL returns the workgroup count only. */
bsp_status bsp_gather_probe(const double* x, int64_t n, int32_t slices, int64_t gathers,
                            double* sink, int64_t sink_len, int64_t* grid, bsp_stream s);
/* Measurement aid, no counterpart in the reference: about `nbytes` of
 * whole-line reads from the same slices as bsp_gather_probe (slice length a
 * power of two of at least 128 doubles): each wave reads 1 KB runs (16 bytes
 * per lane) at random 1 KB-aligned positions of its slice, 32 per lane: the
 * rate at which the L2 feeds the CUs full lines, against the 64-byte sector
 * each 8-byte gather of bsp_gather_probe moves. One sum per workgroup goes
 * to sink; sink == NULL returns the workgroup count only. */
bsp_status bsp_line_probe(const double* x, int64_t n, int32_t slices, int64_t nbytes,
                          double* sink, int64_t sink_len, int64_t* grid, bsp_stream s);
/* y = A*x. x has ncols, y has nrows doubles (device). Stream-ordered. */
bsp_status bsp_csr_spmv(bsp_spmv_plan p, const double* x, double* y, bsp_stream s);
/* Number of workgroups one bsp_csr_spmv launch uses (its main kernel). */
bsp_status bsp_csr_plan_info(bsp_spmv_plan p, int64_t* nblocks, int64_t* nlong);
bsp_status bsp_plan_destroy(bsp_spmv_plan p);
/* y = A*x for an FP64 CSR that allocates nothing: no counterpart in the
 * reference. For callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The
 * adaptive schedule with 2048-unit windows is derived on every call into a
 * fixed 4 MB device array per GPU (held until the process ends), so y is
 * bit-identical to bsp_csr_spmv on a plan from bsp_csr_plan_create_ex with
 * chunk 2048 and seg_cols < 0. Limit: nrows + nnz <= 2^31
 * (BSP_ERR_TOO_LARGE beyond). Calls that overlap on one GPU (two streams,
 * or two host threads) share the array and race; serialise them.
 * Stream-ordered; no host synchronisation. */
bsp_status bsp_csr_spmv_noalloc(const bsp_csr* A, const double* x, double* y, bsp_stream s);

/* ------------------------------------------------------------------------ */
/* Block CSR (BSR) with tm x tk dense blocks, built on the GPU.              */
/* Structure equals csrToBcsr<T,Y> (utils/utils.hpp:9-124): blocks ordered by
 * block row then ascending block column, values row-major within a block
 * (offset (r%tm)*tk + c%tk), zero padded. For 4x4 blocks the AmgT extras of
 * CSR2BSR_GPU / BSR_BALANCED_PREPROCESS_GPU (src/amgT.cu:523-1200) are also
 * built: the 16-bit occupancy map (bit r*4+c, nblocks+1 entries, last 0), the
 * 64-blocks-per-warp balance arrays, avg_nnz = nnz/nblocks and
 * stand = sqrt(mean((blocks per block row - mean)^2)). Duplicate (row, col)
 * entries are not supported. */
typedef struct bsp_bsr_s* bsp_bsr;
typedef struct bsp_bsr_view {
  int64_t nrows, ncols, nnz;    /* of the source CSR */
  int64_t brows, bcols, nblocks;
  int32_t tm, tk;
  const int32_t* rowptr;        /* brows+1 */
  const int32_t* colidx;        /* nblocks */
  const double* val;            /* nblocks*tm*tk */
  const uint16_t* map;          /* nblocks+1 (4x4 only, else NULL) */
  int32_t warpnum;              /* 4x4 only */
  const int32_t* rowptr_by_warp; /* brows+1 */
  const int32_t* rowidx_by_warp; /* warpnum */
  double avg_nnz, stand;
} bsp_bsr_view;

typedef enum bsp_bsr_kernel {
  BSP_BSR_AUTO = 0,
  /* Wave per block row, lane per (block, row) slot, VALU FP64. Any tile.
   * Counterpart of bsr_spmv_cc_fp64 / bcsr_spmv_kernel_column_by_column
   * (src/amgT.cu:276-322, src/bcsr.cu:161-222). */
  BSP_BSR_VALU = 1,
  /* 4x4 or 8x8 blocks on v_mfma_f64_16x16x4f64: four 4x4 blocks (or two
   * 8x8 blocks, four of their columns per product) of a block row stacked
   * into the 16x4 A operand, x segments as B columns, the block-diagonal of
   * the 16x16 result accumulated. Counterpart of bsr_spmv_tc_fp64
   * (src/amgT.cu:210-274, mma.m8n8k4 with 2 blocks). */
  BSP_BSR_MFMA = 2,
  /* 4x4 blocks, load-balanced as the reference's balanced kernels
   * (src/amgT.cu:464-521 on MMA, :160-209 on CUDA cores with the bitmap):
   * one wave per 64-block chunk of the rowptr_by_warp / rowidx_by_warp
   * arrays, so a long block row is spread over several waves. Chunk sums of
   * a block row are added in chunk order by a second pass (no atomics, the
   * same bits every run). MFMA: the chunk's blocks four per product, as
   * BSP_BSR_MFMA. VALU: only entries whose map bit is set are multiplied. */
  BSP_BSR_MFMA_BALANCED = 3,
  BSP_BSR_VALU_BALANCED = 4
} bsp_bsr_kernel;

bsp_status bsp_bsr_build(const bsp_csr* A, int tm, int tk, bsp_stream s, bsp_bsr* out);
bsp_status bsp_bsr_view_get(bsp_bsr h, bsp_bsr_view* v);
/* y = A*x (x: ncols, y: nrows doubles). */
bsp_status bsp_bsr_spmv(bsp_bsr h, int kernel, const double* x, double* y, bsp_stream s);
bsp_status bsp_bsr_destroy(bsp_bsr h);

/* ------------------------------------------------------------------------ */
/* DBSR (include/dbsr.h): rows sorted by (length, index); each group of four
 * sorted rows is packed position-wise into 4x4 "blocks" (element z of sorted
 * row i -> block blcptr[i/4] + z/4, slot (i%4)*4 + z%4) with a column index
 * per slot; plus the 64-blocks-per-warp schedule of balanceDbsr
 * (dbsr.h:94-169, reproduced pass for pass). Built on the GPU. */
typedef struct bsp_dbsr_s* bsp_dbsr;
typedef struct bsp_dbsr_view {
  int64_t nrows, ncols, nnz, brows, nblocks;
  const int32_t* recorder; /* nrows: recorderMap[sorted] = original row */
  const int32_t* blcptr;   /* brows+1 */
  const int32_t* colidx;   /* nblocks*16 */
  const double* val;       /* nblocks*16 */
  const uint16_t* map;     /* nblocks */
  int32_t warpnum;
  const int32_t* prefix;   /* warpnum+1 */
  /* 1 where the reference's second balance pass would run past the warp
   * list of its first pass (undefined there); the list is then extended. */
  int32_t overflow;
} bsp_dbsr_view;

bsp_status bsp_dbsr_build(const bsp_csr* A, bsp_stream s, bsp_dbsr* out);
bsp_status bsp_dbsr_view_get(bsp_dbsr h, bsp_dbsr_view* v);
/* sorted_order != 0: y[i] belongs to sorted row i (the reference's output
 * order, un-permuted by the caller through recorderMap); otherwise y is in
 * the original row order. */
bsp_status bsp_dbsr_spmv(bsp_dbsr h, const double* x, double* y, int sorted_order, bsp_stream s);
typedef enum bsp_dbsr_kernel {
  /* groups at most 8 blocks wide two per wave, wider ones a wave each; per
   * lane one slot's value and column per block (VALU FP64) */
  BSP_DBSR_VALU = 0,
  /* a wave per 4-row group on v_mfma_f64_16x16x4f64, four blocks per
   * product, each element's value and gathered x as A and B operands and
   * the row sums read off the result's diagonal: the counterpart of the
   * reference's mma.m8n8k4 DBSR kernels (src/amgT.cu:324-461), with x used
   * (the reference overrides it with 1.0, :356, :425) */
  BSP_DBSR_MFMA = 1
} bsp_dbsr_kernel;
/* bsp_dbsr_spmv with a choice of kernel (bsp_dbsr_spmv = BSP_DBSR_VALU). */
bsp_status bsp_dbsr_spmv_ex(bsp_dbsr h, int kernel, const double* x, double* y, int sorted_order,
                            bsp_stream s);
bsp_status bsp_dbsr_destroy(bsp_dbsr h);

/* ------------------------------------------------------------------------ */
/* DASP (src/dasp.cu:814-1503): rows classified as long (>= block_longest),
 * row-block (5 .. block_longest-1, sorted longest first, stable), short
 * (1, 2, 3, 4) and empty; each class packed into its padded layout; y is
 * produced in order_rid order. Built on the GPU, arrays bit-exact with the
 * reference's host builder (MMA_M 8, MMA_K 4, rowloop 1). */
typedef struct bsp_dasp_s* bsp_dasp;
typedef struct bsp_dasp_view {
  int64_t nrows, ncols, nnz;
  int32_t row_long, row_block, row_zero;
  int32_t short_row_1, short_row_2, short_row_3, short_row_4, common_13;
  int32_t fill0_nnz_short13, fill0_nnz_short34, fill0_nnz_short22, fill0_nnz_short;
  int32_t threadblock13, threadblock34, threadblock22;
  int32_t blocknum, fill0_nnz_reg, nnz_irreg;
  int32_t warp_number, fill0_nnz_long, nnz_long;
  const int32_t* order_rid;   /* nrows */
  const double* short_val;    /* fill0_nnz_short */
  const int32_t* short_cid;
  const int32_t* long_rpt;    /* row_long+1, in 64-entry warps */
  const double* long_val;     /* fill0_nnz_long */
  const int32_t* long_cid;
  const int32_t* rid_by_warp; /* warp_number */
  const int32_t* block_ptr;   /* blocknum+1 */
  const int32_t* irreg_rpt;   /* row_block+1 */
  const double* reg_val;      /* fill0_nnz_reg */
  const int32_t* reg_cid;
  const double* irreg_val;    /* nnz_irreg */
  const int32_t* irreg_cid;
} bsp_dasp_view;

/* threshold in (0, 1]: a 4-column slab of an 8-row block stays regular while
 * it holds >= threshold*32 entries (the reference passes 0.75, 256). */
bsp_status bsp_dasp_build(const bsp_csr* A, double threshold, int32_t block_longest, bsp_stream s,
                          bsp_dasp* out);
bsp_status bsp_dasp_view_get(bsp_dasp h, bsp_dasp_view* v);
/* y = A*x. permuted != 0: y[k] is row order_rid[k] (the reference's Y_val);
 * otherwise y is in the original row order. */
bsp_status bsp_dasp_spmv(bsp_dasp h, const double* x, double* y, int permuted, bsp_stream s);
typedef enum bsp_dasp_kernel {
  BSP_DASP_VALU = 0, /* one FP64 FMA per entry, wave-level row sums */
  /* long rows and row blocks on v_mfma_f64_16x16x4f64, as the reference's
   * dasp_spmv2 on mma.m8n8k4 (src/dasp.cu:405-812): one entry per lane, the
   * value as the A operand and its gathered x as B, row sums read off the
   * diagonal of the result (an 8-row block's two 4-column slabs stacked as
   * A rows 0..7 and 8..15); short rows on VALU */
  BSP_DASP_MFMA = 1
} bsp_dasp_kernel;
/* bsp_dasp_spmv with a choice of kernel (bsp_dasp_spmv = BSP_DASP_VALU). */
bsp_status bsp_dasp_spmv_ex(bsp_dasp h, int kernel, const double* x, double* y, int permuted,
                            bsp_stream s);
bsp_status bsp_dasp_destroy(bsp_dasp h);

/* ------------------------------------------------------------------------ */
/* Diagonal layout for banded and stencil matrices. No counterpart in the
 * reference, whose answer to index bytes is the bitmap-block BSR of
 * src/amgT.cu:160-209, 781-834: here every distinct diagonal (column - row)
 * is one contiguous run of values over the rows, with a 32-bit mask per row
 * of the diagonals the row holds, so the SpMV reads no column indices
 * (8 bytes per stored slot + 4 per row). y is bit-identical to the
 * row-sequential CSR sum of src/ref.cpp:26-37 (ascending columns, products
 * rounded before each add). BSP_ERR_UNSUPPORTED when the matrix has more
 * than 32 distinct diagonals, BSP_ERR_UNSORTED when a row's columns are not
 * strictly ascending, BSP_ERR_INVALID for a malformed pattern (all checked
 * on the GPU). */
typedef struct bsp_dia_s* bsp_dia;
bsp_status bsp_dia_build(const bsp_csr* A, bsp_stream s, bsp_dia* out);
/* Number of diagonals, their offsets ascending (offsets may be NULL, else
 * room for 32) and the bytes one SpMV streams for the matrix. */
bsp_status bsp_dia_info(bsp_dia h, int32_t* ndiag, int32_t* offsets, int64_t* bytes);
bsp_status bsp_dia_spmv(bsp_dia h, const double* x, double* y, bsp_stream s);
bsp_status bsp_dia_destroy(bsp_dia h);

/* ------------------------------------------------------------------------ */
/* Per-matrix SpMV autotuner. No counterpart in the reference, where format
 * and thresholds are fixed per call site (bench/spmvfp64.cpp:135-177; DASP
 * threshold 0.75 and block_longest 256 at :177; AmgT's stand / avgnz cuts,
 * src/amgT.cu:1355-1488): the tuner builds each candidate layout of one FP64
 * matrix on the GPU, times its SpMV and keeps the fastest. Candidates:
 * adaptive CSR with windows 512 / 1024 / 2048 (and, when the automatic plan
 * segments columns, the unsegmented plan too); 4x4 and 8x8 BSR on VALU and
 * MFMA when their blocks are at least half full on average (bsp_block_info); DBSR; DASP
 * with threshold 0.5 / 0.75 / 1.0 x block_longest 64 / 256; the diagonal
 * layout when the matrix has at most 32 diagonals. Timing is
 * cache-warm (the same matrix and x every launch), so it ranks the layouts'
 * kernels, not a cold HBM read. The choice is a plain record: store it and
 * rebuild the same layout later with bsp_tuned_from_record, no timing. */
typedef enum bsp_tune_format {
  BSP_TUNE_CSR = 0,  /* adaptive plan: chunk, seg_cols as in bsp_csr_options;
                        kernel 1 = sorted-block segments (seg_mode 2) */
  BSP_TUNE_BSR4 = 1, /* 4x4 BSR: kernel = any bsp_bsr_kernel but AUTO */
  BSP_TUNE_DBSR = 2, /* kernel = BSP_DBSR_VALU or BSP_DBSR_MFMA */
  BSP_TUNE_DASP = 3, /* threshold, block_longest as in bsp_dasp_build; kernel = bsp_dasp_kernel */
  BSP_TUNE_BSR8 = 4, /* 8x8 BSR: kernel = BSP_BSR_VALU or BSP_BSR_MFMA */
  BSP_TUNE_DIA = 5   /* diagonal layout (bsp_dia_build) */
} bsp_tune_format;

typedef struct bsp_tune_record {
  int32_t format;        /* bsp_tune_format */
  int32_t kernel;        /* bsp_bsr_kernel (BSR4, BSR8), bsp_dbsr_kernel, bsp_dasp_kernel */
  int32_t chunk;         /* CSR window, 128..2048 */
  int32_t block_longest; /* DASP */
  int64_t seg_cols;      /* CSR: 0 = automatic, < 0 = never segment */
  double threshold;      /* DASP */
  double time_us;        /* mean SpMV time measured by the tuner (0 when replayed) */
  int64_t reserved[4];
} bsp_tune_record;

typedef struct bsp_tune_options {
  int32_t formats; /* bit mask of 1 << bsp_tune_format to try; 0 = all */
  int32_t reps;    /* timed SpMVs per candidate after two warmups; 0 = 10 */
  int64_t reserved[7];
} bsp_tune_options;

typedef struct bsp_tuned_s* bsp_tuned;
/* Times every candidate (stream-ordered, synchronises `s`); a candidate
 * whose build fails (unsorted columns for a segmented plan, memory) is
 * skipped. BSP_ERR_UNSUPPORTED when no candidate could be built. */
bsp_status bsp_autotune(const bsp_csr* A, const bsp_tune_options* opt, bsp_stream s,
                        bsp_tuned* out);
/* The layout `rec` names, built without timing (rec->time_us is ignored). */
bsp_status bsp_tuned_from_record(const bsp_csr* A, const bsp_tune_record* rec, bsp_stream s,
                                 bsp_tuned* out);
bsp_status bsp_tuned_record(bsp_tuned h, bsp_tune_record* rec);
/* The candidates timed, in the order tried: up to `cap` records; *count is
 * the number tried (0 for a handle built from a record). */
bsp_status bsp_tuned_trials(bsp_tuned h, bsp_tune_record* recs, int32_t cap, int32_t* count);
/* y = A*x in the original row order, with the chosen layout. */
bsp_status bsp_tuned_spmv(bsp_tuned h, const double* x, double* y, bsp_stream s);
bsp_status bsp_tuned_destroy(bsp_tuned h);

/* ------------------------------------------------------------------------ */
/* Differentiable CSR SpMM: Y = A*X with dense row-major X (ncols x F) and Y
 * (nrows x F, FP32), FP32 accumulation. A's values are FP32 or BF16, dense
 * operands FP32 or BF16; F is 64, 128 or 256. No counterpart in the
 * reference (its only SpMM is the block kernel of src/smat.cu); this is the
 * forward and backward of a GNN-style layer:
 *   bsp_spmm    Y  = A * X
 *   bsp_spmm_t  dX = A^T * dY   (over the pattern's CSC, values read through
 *                                the CSR permutation: no transposed copy)
 *   bsp_sddmm   dvals[j] = dY[row_j, :] . X[col_j, :]   (A's pattern, CSR order)
 * The plan (work lists, CSC permutation) is built on the GPU from the
 * pattern alone (A.val may be NULL); the values (nnz entries in CSR order,
 * FP32 or BF16) are passed to every call. Dense operands and outputs must
 * be 16-byte aligned (BSP_ERR_INVALID otherwise). Results are deterministic
 * (no atomics). */
typedef struct bsp_spmm_plan_s* bsp_spmm_plan;
bsp_status bsp_spmm_plan_create(const bsp_csr* A, bsp_stream s, bsp_spmm_plan* out);
bsp_status bsp_spmm(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* X,
                    bsp_dtype xdtype, int64_t F, float* Y, bsp_stream s);
bsp_status bsp_spmm_t(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* dY,
                      bsp_dtype dtype, int64_t F, float* dX, bsp_stream s);
bsp_status bsp_sddmm(bsp_spmm_plan p, const void* dY, const void* X, bsp_dtype dtype, int64_t F,
                     float* dvals, bsp_stream s);
/* The whole backward in one traversal of the CSC: dX as bsp_spmm_t and
 * dvals as bsp_sddmm, each bit-identical to those calls, every gathered row
 * of dY serving both products (dY and X share `dtype`). */
bsp_status bsp_spmm_t_sddmm(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* dY,
                           const void* X, bsp_dtype dtype, int64_t F, float* dX, float* dvals,
                           bsp_stream s);
bsp_status bsp_spmm_plan_destroy(bsp_spmm_plan p);

/* ------------------------------------------------------------------------ */
/* Block-sparse SpMM on FP16/BF16 MFMA, FP32 accumulation. Replaces the
 * kernel of smatSpmmfp16 (src/smat.cu:22-91, 103-202): 16x16 dense blocks,
 * each scaled by a power of two computed in FP64 on the GPU (values fill
 * (-1, 1) before the one rounding to FP16/BF16), plus the backward.
 * Dense operands are row-major FP16/BF16 (the handle's type): B is K x N,
 * dC is M x N; outputs are FP32 row-major (C: M x N, dB: K x N) and dvals
 * (nblocks x 16 x 16, the gradient with respect to the unscaled block
 * values, block order of the BSR). N is 32, 64, 128 or 256. B, dC and
 * dvals are 16-byte aligned (16-byte loads and stores; BSP_ERR_INVALID
 * otherwise). */
typedef struct bsp_bspmm_s* bsp_bspmm;
/* A: a 16x16 BSR (bsp_bsr_build with tm = tk = 16) that must outlive h. */
bsp_status bsp_bspmm_build(bsp_bsr A, bsp_dtype half_type, bsp_stream s, bsp_bspmm* out);
bsp_status bsp_bspmm_view_get(bsp_bspmm h, const float** scale, const uint16_t** hv,
                              const uint16_t** hvt, int64_t* nblocks);
/* C = A * B */
bsp_status bsp_bspmm_fwd(bsp_bspmm h, const void* B, int64_t N, float* C, bsp_stream s);
/* dB = A^T * dC and dvals = dC * B^T on A's blocks (either output may be NULL) */
bsp_status bsp_bspmm_bwd(bsp_bspmm h, const void* B, const void* dC, int64_t N, float* dB,
                         float* dvals, bsp_stream s);
bsp_status bsp_bspmm_destroy(bsp_bspmm h);

/* ------------------------------------------------------------------------ */
/* Host-side synthetic matrices (generators; see benchsp_amd/csrc/gen.cpp). */
typedef struct bsp_host_csr_s* bsp_host_csr;
bsp_status bsp_gen_uniform(int64_t n, uint64_t seed, double* out);
bsp_status bsp_gen_normal(int64_t n, uint64_t seed, double* out);
bsp_status bsp_gen_stencil5(int64_t nx, int64_t ny, bsp_host_csr* out);
bsp_status bsp_gen_stencil3d(int64_t nx, int64_t ny, int64_t nz, int points, uint64_t seed,
                             bsp_host_csr* out);
bsp_status bsp_gen_powerlaw(int64_t n, int64_t ncols, double mean, double sigma, int64_t cap,
                            uint64_t seed, bsp_host_csr* out);
bsp_status bsp_gen_blocksparse(int64_t nbr, int64_t nbc, int bs, int64_t nblocks, uint64_t seed,
                               bsp_host_csr* out);
bsp_status bsp_gen_graph(int64_t n, double avg_deg, double gamma, int symmetric, uint64_t seed,
                         bsp_host_csr* out);
bsp_status bsp_host_csr_info(bsp_host_csr h, int64_t* nrows, int64_t* ncols, int64_t* nnz);
bsp_status bsp_host_csr_copy(bsp_host_csr h, int64_t* rowptr, int32_t* col, double* val);
void bsp_host_csr_free(bsp_host_csr h);

/* ---- Matrix files and block statistics ------------------------------------ */
/* Replaces the drivers' reader and assembly (bench/spmvfp64.cpp:36-93, which
 * use fast_matrix_market): a MatrixMarket coordinate file (real, integer or
 * pattern = 1.0; general, symmetric or skew-symmetric, the mirrored
 * off-diagonal entries added) into a host CSR sorted by (row, column).
 * Duplicate entries are all kept, in file order. */
bsp_status bsp_mtx_read(const char* path, bsp_host_csr* out);
/* COO (0-based) to host CSR, the same assembly; vals may be NULL (1.0). */
bsp_status bsp_coo_to_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* rows,
                          const int32_t* cols, const double* vals, bsp_host_csr* out);
/* Replaces bench/blockinfo.cpp:96-119: b x b blocks holding an entry
 * (`blocks`, avg = nnz / blocks) and blocks after packing each row's entries
 * to the left (`blocks_packed` = sum over groups of b rows of the longest
 * row's ceil(len / b)). Computed on the GPU from a device CSR's pattern
 * (values are not read). The batch autotuner uses it to decide which
 * matrices get a 4x4 BSR candidate. */
typedef struct bsp_blockinfo {
  int64_t nnz, blocks, blocks_packed;
  double avg, avg_packed;
} bsp_blockinfo;
bsp_status bsp_block_info(const bsp_csr* A, int block_len, bsp_stream s, bsp_blockinfo* out);

/* ------------------------------------------------------------------------ */
/* Reference call surface (host pointers).                                   */
/* Every function uploads its inputs, builds what the reference builds (on the
 * GPU), times `repeat` launches each followed by a stream synchronisation
 * (host clock, as the reference does) and downloads the result. avg_time
 * receives whole microseconds per launch like the reference. Results are
 * y = A*x exactly once per call: the reference's accumulation into an
 * un-zeroed output (atomicAdd across repeats) is not reproduced.          */

/* Layout of CSRFormat<double, int32_t> (include/matrixFormat.hpp:4-12). */
typedef struct bsp_CSRFormat_f64 {
  int32_t* rowPtr;
  int32_t* colIdx;
  double* values;
  int32_t row, col;
  int32_t nnz;
} bsp_CSRFormat_f64;

/* Layout of dbsrMat (include/matrixFormat.hpp:86-106). */
typedef struct bsp_dbsrMat {
  int row, col, nnz, blc_row, blc_col, blc_num;
  double stand, avg_nnz;
  int32_t* blcPtr;
  int32_t* colIdx;
  double* blcVal;
  unsigned short* blcMap;
  int32_t blocknnz;
  int32_t warpnum;
  int32_t* rowPtrbyWarp;
  int32_t* rowIdxbyWarp;
  int32_t* recorderMap;
  int32_t* prefix;
} bsp_dbsrMat;

/* Replaces csrTodbsrWithRecorder (include/dbsr.h:208-253, with csrTodbsr
 * :10-93 and balanceDbsr :94-169): builds on the GPU and fills `out` with
 * malloc'ed host arrays (blcPtr, colIdx, blcVal, blcMap, recorderMap,
 * prefix) bit-identical to the reference builder's; free them with
 * bsp_dbsrMat_free. blc_num is set to the number of blocks, blc_col to 0.
 * Returns 0 on success. */
int bsp_csrTodbsrWithRecorder(const bsp_CSRFormat_f64* csr, bsp_dbsrMat* out);
void bsp_dbsrMat_free(bsp_dbsrMat* m);
/* Replaces csrTodbsr (include/dbsr.h:10-93): DBSR of the CSR with rows in
 * their own order, built on the GPU. Sets the fields the reference sets
 * (row, col, nnz, blc_row, blocknnz, blcPtr, colIdx, blcVal, blcMap) plus
 * blc_num and avg_nnz; the struct is cleared first, so recorderMap and
 * prefix are null (the reference leaves them unset) until balanceDbsr.
 * Returns 0 on success. */
int bsp_csrTodbsr(const bsp_CSRFormat_f64* csr, bsp_dbsrMat* out);
/* Replaces balanceDbsr (include/dbsr.h:94-169): packs the block rows of
 * m->blcPtr into 64-block warps on the GPU and sets m->prefix (malloc'ed,
 * warpnum + 1 entries; a previous prefix is freed) and m->warpnum. Where the
 * reference's second pass would run past its warp list (undefined there),
 * the list is extended by one warp. Returns 0 on success. */
int bsp_balanceDbsr(bsp_dbsrMat* m);
/* Replaces csrRecorderSelf (include/dbsr.h:170-206): `out` receives the CSR
 * with rows sorted by (length, index), built on the GPU; its arrays are
 * malloc'ed (free with bsp_free). Returns 0 on success. */
int bsp_csrRecorderSelf(const bsp_CSRFormat_f64* csr, bsp_CSRFormat_f64* out);
/* The arithmetic of csrToBcsr<T,Y> (utils/utils.hpp:9-124) with int64
 * indices and FP64 values: tile_m x tile_k blocks (1..64), block columns
 * ascending, row-major zero-padded values, built on the GPU. The three
 * arrays are malloc'ed (free with bsp_free): bcsrRowPtr ceil(row/tile_m)+1,
 * bcsrColIdx and bcsrVal (tile_m*tile_k per block) of *nonzeroBlocks
 * blocks. include/benchsp_compat.hpp wraps it as the reference template.
 * Returns 0 on success. */
int bsp_csrToBcsr(const int64_t* rowPtr, const int64_t* colIdx, const double* values, int64_t row,
                  int64_t col, int tile_m, int tile_k, int64_t* nonzeroBlocks,
                  int64_t** bcsrRowPtr, int64_t** bcsrColIdx, double** bcsrVal);
/* Frees memory the library malloc'ed for the caller. */
void bsp_free(void* p);
/* Uploads a host DBSR (as filled above) into a device handle. */
bsp_status bsp_dbsr_from_host(const bsp_dbsrMat* m, bsp_stream s, bsp_dbsr* out);

/* Replaces int amgT_spmv_fp64_dbsr(dbsrMat*, double* hX, double* hY, int32_t
 * warmup, int32_t repeat, double& avg_time) (include/amgT.h:4,
 * src/amgT.cu:1202-1290): y in sorted-row order (the caller un-permutes
 * through recorderMap). */
int bsp_amgT_spmv_fp64_dbsr(bsp_dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                            int32_t repeat, double* avg_time);
/* Replaces amgT_spmv_fp64_dbsr_balance (include/amgT.h:6, src/amgT.cu:1292-1353);
 * the same kernel (it balances by block rows itself, not by `prefix`). */
int bsp_amgT_spmv_fp64_dbsr_balance(bsp_dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                                    int32_t repeat, double* avg_time);
/* Replaces int amgT_spmv_fp64(...) (include/amgT.h:3, src/amgT.cu:1355-1488):
 * CSR -> 4x4 BSR on the GPU, then the balanced kernels the reference runs
 * (its stand is forced to 12, amgT.cu:1176): BSP_BSR_MFMA_BALANCED when
 * nnz/blocks >= 10, BSP_BSR_VALU_BALANCED otherwise (amgT.cu:1394-1407),
 * 5 warmups, `repeat` timed launches. */
int bsp_amgT_spmv_fp64(int32_t* hA_csrOffsets, int32_t* hA_columns, double* hA_values, double* hX,
                       double* hY, int32_t A_num_rows, int32_t A_num_cols, int32_t A_nnz,
                       int repeat, double* avg_time);
/* Replaces void bcsr_spmv_fp64(...) (include/bcsr.h:3, src/bcsr.cu:224-302):
 * 3x3 BCSR built on the GPU, `repeat` launches. */
void bsp_bcsr_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values,
                        double* hX, double* hY, int64_t A_num_rows, int64_t A_num_cols,
                        int64_t A_nnz, int repeat);
/* Replaces void spmv_all(...) (include/dasp.h:1-2, src/dasp.cu:814-1823): DASP
 * layout built on the GPU, 100 warmups + 10 timed launches as the reference
 * (it ignores `repeat`); Y in order_rid order and order_rid (rowA entries)
 * returned. filename and NUM are unused, as in the reference. */
void bsp_spmv_all(char* filename, double* csrValA, int* csrRowPtrA, int* csrColIdxA, double* X_val,
                  double* Y_val, int* order_rid, int rowA, int colA, int nnzA, int NUM,
                  double threshold, int block_longest, double* avg_time);
/* Replaces void MySpmvWithPrefp64(...) (include/spmv.h:3, src/spmv.cu:323-397):
 * short/long split at 128; outvalues keeps full FP64 (the reference truncates
 * to float). */
void bsp_MySpmvWithPrefp64(int64_t* rowPtr, int64_t* colIdx, int64_t rows, int64_t cols,
                           int64_t nnz, double* values, double* x, double* outvalues);
/* Replaces void MySpmvWithPrefp16(...) (include/spmv.h:2, src/spmv.cu:197-272):
 * values and x rounded to FP16, products and sums in FP32 (the reference
 * sums in half), y rounded to FP16 and returned as float. */
void bsp_MySpmvWithPrefp16(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                           int32_t nnz, float* values, float* x, float* outvalues);
/* Replaces void smatSpmmfp16(...) (include/smat.h:1, src/smat.cu:103-202):
 * C (row x n, FP32, row-major) = A * B with B = 1 (col x n) as in the
 * reference, on the FP16 block-SpMM kernel; C is written to outvalues (the
 * reference never writes it). n: any multiple of 32. */
void bsp_smatSpmmfp16(int* rowPtr, int* colIdx, int row, int col, int n, int nnz, float* values,
                      float* outvalues);
/* Replaces int cusparse_spmv_fp64(...) (include/cusp.h:4, src/cusparse.cu:104-197)
 * with rocSPARSE: 64-bit indices, default algorithm, 10 warmups + `repeat`
 * timed calls. The vendor comparator, not a library kernel. */
int bsp_cusparse_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values,
                           double* hX, double* hY, int64_t A_num_rows, int64_t A_num_cols,
                           int64_t A_nnz, int repeat, double* avg_time);
/* Device-level FP16 split SpMV: plan from bsp_csr_plan_create(method
 * BSP_CSR_SPLIT, FP16 values); x FP16 (ncols), y FP32 (nrows). */
bsp_status bsp_csr_spmv_f16(bsp_spmv_plan p, const uint16_t* x, float* y, bsp_stream s);
/* ---- Batched SpMV + SpMM over many matrices --------------------------------
 * No counterpart in the reference (its drivers run one matrix per process).
 * A batch borrows `count` device CSRs (FP64 values, int32 indices). One step
 * is y = A_m x_m (FP64) and Y_m = A_m X_m (A in FP16, X 8 columns of FP16,
 * FP32 accumulation and output) for every m. Vectors are concatenated: x_m
 * starts at col_off[m] of x, y_m at row_off[m] of y; X and Y are row-major
 * with 8 entries per row (X: 16-byte aligned). The tiled schedule
 * (BSP_BATCH_PERSISTENT) runs each operation as one launch over tiles of all
 * matrices, each workgroup a contiguous range of tiles; the plain schedule
 * launches a thread-per-row kernel per matrix and operation. With tune = 1
 * the builder times, per matrix, the CSR tile windows {128, 256, 512, 1024}
 * rows+entries for the SpMV and {128, 256, 512} for the SpMM (which set rows
 * per workgroup and lanes per row) and, for the SpMV, the format and block
 * size: 4x4 and 8x8 BSR wherever their blocks are at least 11/16 full on
 * average, and DASP (threshold 0.75, block_longest 256; its fused launch's
 * workgroups become tiles) for matrices of at least 1024 entries, and
 * keeps a candidate 10% faster than the best before it; for both
 * operations the diagonal layout (bsp_dia_build) of every matrix with at
 * most 8 diagonals (stencils), a thread per row, also against the best
 * before it; then, per
 * operation, it times the whole batch with each uniform window and with the
 * per-matrix choice, at several launch sizes, and keeps the fastest.
 * tune = 0 uses windows of 1024 (SpMV) and 512 (SpMM); tune = 2 puts every
 * eligible matrix in BSR (8x8 where eligible, else 4x4), tune = 3 every
 * matrix in DASP and tune = 4 every matrix with at most 8 diagonals in the
 * diagonal layout for both operations, without timing.
 * When a step has both operations, the SpMM runs on a side stream of the
 * batch, forked from and joined back to the caller's stream with events.
 * A captured graph holds one step; replaying it needs no host work. */
typedef struct bsp_batch_s* bsp_batch;
typedef struct bsp_batch_options {
  int32_t tune;
  /* workgroups of each tiled launch, capped at the tile count: 0 = 16x as
   * many as the device holds at once (tune = 1: the fastest of 1x, 4x, 16x
   * and one per tile); the resident count gives the persistent form, each
   * workgroup one long range of tiles */
  int32_t grid;
  int64_t reserved[7];
} bsp_batch_options;
enum {
  BSP_BATCH_CSR = 0,
  BSP_BATCH_BSR4 = 1,
  BSP_BATCH_BSR8 = 2,
  BSP_BATCH_DASP = 3,
  BSP_BATCH_DIA = 4
};
enum { BSP_BATCH_PERSISTENT = 0, BSP_BATCH_PLAIN = 1 };
bsp_status bsp_batch_create(const bsp_csr* mats, int64_t count, const bsp_batch_options* opt,
                            bsp_stream s, bsp_batch* out);
/* row_off, col_off: count+1 entries each. */
bsp_status bsp_batch_layout(bsp_batch b, int64_t* row_off, int64_t* col_off);
/* Per matrix: the SpMV's format (BSP_BATCH_CSR/BSR4/BSR8/DASP/DIA) and the
 * CSR windows chosen (count each). */
bsp_status bsp_batch_choice(bsp_batch b, int32_t* format, int32_t* spmv_window,
                            int32_t* spmm_window);
/* Per matrix: the SpMM's format, BSP_BATCH_CSR or BSP_BATCH_DIA (count). */
bsp_status bsp_batch_spmm_format(bsp_batch b, int32_t* format);
bsp_status bsp_batch_info(bsp_batch b, int64_t* spmv_tiles, int64_t* spmm_tiles, int64_t* grid);
/* One step on stream s; x/y or X/Y may both be NULL to skip that operation. */
bsp_status bsp_batch_run(bsp_batch b, int schedule, const double* x, double* y, const uint16_t* X,
                         float* Y, bsp_stream s);
/* Captures one step (same arguments) into a graph owned by the batch,
 * replacing any earlier one; bsp_batch_graph_launch replays it on s. */
bsp_status bsp_batch_graph_capture(bsp_batch b, int schedule, const double* x, double* y,
                                   const uint16_t* X, float* Y);
bsp_status bsp_batch_graph_launch(bsp_batch b, bsp_stream s);
bsp_status bsp_batch_destroy(bsp_batch b);

/* Replaces int cusparse_spmv(...) (include/cusp.h:3, src/cusparse.cu:35-102)
 * with rocSPARSE: FP32, one call, y = A*x. */
int bsp_cusparse_spmv(int* hA_csrOffsets, int64_t* hA_columns, float* hA_values, float* hX,
                      float* hY, int64_t A_num_rows, int64_t A_num_cols, int64_t A_nnz);
/* Replaces int cusparse_spmv_fp16(...) (include/cusp.h:5, src/cusparse.cu:200-270)
 * with rocSPARSE: A and x FP16 (binary16 bit patterns), FP32 compute and
 * accumulation, y rounded to FP16 (round to nearest even) into hY. */
int bsp_cusparse_spmv_fp16(int* hA_csrOffsets, int32_t* hA_columns, const uint16_t* hA_values,
                           const uint16_t* hX, uint16_t* hY, int32_t A_num_rows,
                           int32_t A_num_cols, int32_t A_nnz);
/* Replaces void cusparse_spmv_fp16_warpper(...) (include/cusp.h:6,
 * src/cusparse.cu:272-286): float inputs rounded to FP16, then as above; the
 * FP16 result is widened back into hY32 (the reference drops it). */
void bsp_cusparse_spmv_fp16_warpper(int* hA_csrOffsets, int32_t* hA_columns, float* hA_values32,
                                    float* hX32, float* hY32, int32_t A_num_rows,
                                    int32_t A_num_cols, int32_t A_nnz);

/* Replace the CPU reference loops of include/ref.h:2-4 (src/ref.cpp:3-37) with
 * GPU thread-per-row kernels that add each row left to right with no fused
 * multiply-add, so y is bit-identical to the CPU loop: FP64; FP32; FP16
 * (inputs rounded to FP16, FP32 sum, y rounded to FP16, returned as float).
 * avg_time (FP64 only) receives the kernel's host-timed microseconds. */
void bsp_ref_spmv_fp64(int64_t* A_csrOffsets, int64_t* A_columns, double* A_values, double* x,
                       double* y, int64_t row, int64_t col, double* avg_time);
void bsp_ref_spmv_fp32(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                       float* y, int32_t row, int32_t col);
void bsp_ref_spmv_fp16(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                       float* y, int32_t row, int32_t col);

/* Device-level rocSPARSE SpMV with int32 indices and the given CSR
 * algorithm (rocsparse_spmv_alg codes: 0 default, 2 adaptive, 3 rowsplit,
 * 7 LRB; BSP_ERR_INVALID otherwise), prepared once (buffer and
 * analysis) and run per call: the vendor point reported next to the
 * library's kernels. Value types by A->dtype: F64 (x, y FP64), F32 (x, y
 * FP32), F16 (x FP16, y FP32, FP32 compute). */
typedef struct bsp_rocsparse_plan_s* bsp_rocsparse_plan;
bsp_status bsp_rocsparse_plan_create(const bsp_csr* A, int alg, bsp_stream s,
                                     bsp_rocsparse_plan* out);
bsp_status bsp_rocsparse_spmv(bsp_rocsparse_plan p, const void* x, void* y, bsp_stream s);
bsp_status bsp_rocsparse_plan_destroy(bsp_rocsparse_plan p);

/* Device-level rocSPARSE SpMM (counterpart of cusparse_spmm_fp16,
 * src/cusparse.cu:288-390, generalised): C = op(A) B with A a device CSR of
 * FP32, FP16 or BF16 values, B row-major (ncols x N, or nrows x N when
 * transpose) of the same type, C row-major FP32, FP32 compute; prepared once,
 * run per call. The vendor point next to the library's SpMM kernels. `alg`
 * is a rocsparse_spmm_alg code for CSR: 0 default, 1 csr, 4 row split,
 * 5 nnz split, 9 merge path (BSP_ERR_INVALID otherwise; rocSPARSE itself
 * may refuse a combination, reported as BSP_ERR_HIP). */
typedef struct bsp_rocsparse_spmm_plan_s* bsp_rocsparse_spmm_plan;
bsp_status bsp_rocsparse_spmm_plan_create(const bsp_csr* A, int64_t N, int transpose, int alg,
                                          bsp_stream s, bsp_rocsparse_spmm_plan* out);
bsp_status bsp_rocsparse_spmm(bsp_rocsparse_spmm_plan p, const void* B, float* C, bsp_stream s);
bsp_status bsp_rocsparse_spmm_plan_destroy(bsp_rocsparse_spmm_plan p);

/* Replaces int spmv_fp64_warpper(...) (include/csrspmv.h:2, src/csrspmv.cu:104-134):
 * thread-per-row CSR, 10 warmups, `repeat` timed launches; y = A*x. */
int bsp_spmv_fp64_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                          int32_t nnz, double* values, double* x, double* outvalues, int repeat,
                          double* avg_time);
/* Replaces int spmv_fp64_balance_warpper(...) (include/csrspmv.h:3,
 * src/csrspmv.cu:136-195). y = A*x exactly once per call (the reference's
 * atomicAdd into an un-zeroed y is not reproduced). */
int bsp_spmv_fp64_balance_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                                  int32_t nnz, double* values, double* x, double* outvalues,
                                  int repeat, double* avg_time);

#ifdef __cplusplus
}
#endif
#endif /* BSP_H_ */

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
/* Library version and the gfx target its kernels were built for. */
const char* bsp_version(void);

/* ------------------------------------------------------------------------ */
/* Device CSR descriptor (all pointers are device memory, not owned).        */
/* rowptr has nrows+1 int32 entries (nnz < 2^31); columns ascending per row. */
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

/* Build the per-matrix schedule for `method` on the GPU (stream-ordered).
 * The plan keeps the descriptor's device pointers; they must outlive it. */
bsp_status bsp_csr_plan_create(const bsp_csr* A, int method, bsp_stream s, bsp_spmv_plan* out);
/* y = A*x. x has ncols, y has nrows doubles (device). Stream-ordered. */
bsp_status bsp_csr_spmv(bsp_spmv_plan p, const double* x, double* y, bsp_stream s);
/* Number of workgroups one bsp_csr_spmv launch uses (its main kernel). */
bsp_status bsp_csr_plan_info(bsp_spmv_plan p, int64_t* nblocks, int64_t* nlong);
bsp_status bsp_plan_destroy(bsp_spmv_plan p);

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

/* ------------------------------------------------------------------------ */
/* Reference call surface (host pointers).                                   */

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

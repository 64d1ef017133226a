// C-ABI entry points: status plumbing, CSR plans and the reference-surface
// host wrappers for the CSR kernels.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstring>
#include <memory>
#include <string>

#include "bsp_hip.hpp"
#include "csr_plan.hpp"

namespace bsp {
static thread_local std::string g_last_error;
bsp_status fail(bsp_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}
}  // namespace bsp

using namespace bsp;

namespace {

bsp_status check_csr(const bsp_csr* A) {
  BSP_REQUIRE(A, "null CSR descriptor");
  BSP_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, "negative CSR size");
  if (!sizes_fit(A->nrows, A->ncols, A->nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(A->nrows == 0 || A->rowptr, "null rowptr");
  BSP_REQUIRE(A->nnz == 0 || (A->colidx && A->val), "null colidx/val");
  return BSP_OK;
}

// Host-wrapper scaffolding shared by the reference-surface functions: upload
// a host CSR (int32 indices), x and y, run `body` warmup + repeat times with a
// stream sync after each launch (the reference times launch+sync pairs on the
// host), download y. avg_time gets whole microseconds per launch like the
// reference (integer division of the elapsed microseconds by repeat).
struct HostRun {
  DevBuf<int32_t> rp, ci;
  DevBuf<double> v, x, y;
  bsp_csr A{};
  hipStream_t s = nullptr;
};

bsp_status upload_csr(HostRun& h, const int32_t* rowPtr, const int32_t* colIdx, int64_t rows,
                      int64_t cols, int64_t nnz, const double* values, const double* x) {
  BSP_TRY(check_host_csr(rowPtr, colIdx, rows, cols, nnz));
  BSP_REQUIRE((nnz == 0 || values) && (cols == 0 || x), "null host pointer");
  BSP_TRY(h.rp.alloc(rows + 1));
  BSP_TRY(h.ci.alloc(nnz));
  BSP_TRY(h.v.alloc(nnz));
  BSP_TRY(h.x.alloc(cols));
  BSP_TRY(h.y.alloc(rows));
  BSP_HIP(hipMemcpy(h.rp.get(), rowPtr, sizeof(int32_t) * (rows + 1), hipMemcpyHostToDevice));
  if (nnz) {
    BSP_HIP(hipMemcpy(h.ci.get(), colIdx, sizeof(int32_t) * nnz, hipMemcpyHostToDevice));
    BSP_HIP(hipMemcpy(h.v.get(), values, sizeof(double) * nnz, hipMemcpyHostToDevice));
  }
  if (cols) BSP_HIP(hipMemcpy(h.x.get(), x, sizeof(double) * cols, hipMemcpyHostToDevice));
  BSP_HIP(hipMemset(h.y.get(), 0, sizeof(double) * (rows ? rows : 1)));
  h.A = bsp_csr{rows, cols, nnz, h.rp.get(), h.ci.get(), h.v.get(), BSP_F64};
  return BSP_OK;
}

template <class F>
bsp_status timed_loop(int warmup, int repeat, hipStream_t s, double* avg_time, F body) {
  for (int i = 0; i < warmup; ++i) {
    BSP_TRY(body());
    BSP_HIP(hipStreamSynchronize(s));
  }
  auto t0 = std::chrono::high_resolution_clock::now();
  for (int i = 0; i < repeat; ++i) {
    BSP_TRY(body());
    BSP_HIP(hipStreamSynchronize(s));
  }
  auto t1 = std::chrono::high_resolution_clock::now();
  if (avg_time) {
    long long us = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
    *avg_time = repeat > 0 ? (double)(us / repeat) : 0.0;
  }
  return BSP_OK;
}

int csr_host_wrapper(int method, int warmup, int32_t* rowPtr, int32_t* colIdx, int32_t rows,
                     int32_t cols, int32_t nnz, double* values, double* x, double* outvalues,
                     int repeat, double* avg_time) {
  HostRun h;
  if (upload_csr(h, rowPtr, colIdx, rows, cols, nnz, values, x) != BSP_OK) return 1;
  BSP_REQUIRE(outvalues || rows == 0, "null output");
  bsp_spmv_plan p = nullptr;
  if (bsp_csr_plan_create(&h.A, method, h.s, &p) != BSP_OK) return 1;
  std::unique_ptr<bsp_spmv_plan_s> own(p);
  bsp_status st = timed_loop(warmup, repeat, h.s, avg_time,
                             [&] { return csr_plan_run(p, h.x.get(), h.y.get(), h.s); });
  if (st != BSP_OK) return 1;
  if (rows && hipMemcpy(outvalues, h.y.get(), sizeof(double) * rows, hipMemcpyDeviceToHost) !=
                  hipSuccess)
    return 1;
  return 0;
}

}  // namespace

extern "C" {

const char* bsp_status_string(bsp_status s) {
  switch (s) {
    case BSP_OK: return "ok";
    case BSP_ERR_INVALID: return "invalid argument";
    case BSP_ERR_HIP: return "HIP runtime error";
    case BSP_ERR_ALLOC: return "allocation failed";
    case BSP_ERR_UNSORTED: return "CSR columns not ascending within a row";
    case BSP_ERR_UNSUPPORTED: return "unsupported input";
    case BSP_ERR_TOO_LARGE: return "size exceeds supported range";
  }
  return "unknown status";
}

const char* bsp_last_error(void) { return g_last_error.c_str(); }

bsp_status bsp_copy_to_host(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return BSP_OK;
  BSP_REQUIRE(dst && src, "null pointer");
  BSP_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return BSP_OK;
}

const char* bsp_version(void) { return "bsp 0.1 gfx950"; }

bsp_status bsp_csr_plan_create_ex(const bsp_csr* A, int method, const bsp_csr_options* opt,
                                  bsp_stream s, bsp_spmv_plan* out) {
  BSP_TRY(check_csr(A));
  BSP_REQUIRE(out, "null output handle");
  BSP_REQUIRE(A->dtype == BSP_F64 || (A->dtype == BSP_F16 && method == BSP_CSR_SPLIT),
              "CSR SpMV plans need FP64 values (FP16 with the split method)");
  auto p = std::make_unique<bsp_spmv_plan_s>();
  p->A = *A;
  p->method = method;
  if (opt) p->opt = *opt;
  BSP_TRY(csr_plan_build(p.get(), (hipStream_t)s));
  *out = p.release();
  return BSP_OK;
}

bsp_status bsp_csr_plan_create(const bsp_csr* A, int method, bsp_stream s, bsp_spmv_plan* out) {
  return bsp_csr_plan_create_ex(A, method, nullptr, s, out);
}

bsp_status bsp_csr_plan_segments(bsp_spmv_plan p, int64_t* seg_cols, int64_t* nseg) {
  BSP_REQUIRE(p, "null plan");
  if (seg_cols) *seg_cols = p->seg_width;
  if (nseg) *nseg = (int64_t)p->seg.size();
  return BSP_OK;
}

bsp_status bsp_csr_plan_probe(bsp_spmv_plan p, int with_values, const double* x, double* sink,
                             int64_t sink_len, int64_t* grid, bsp_stream s) {
  BSP_REQUIRE(p && (x || !sink), "null argument");
  return csr_plan_probe(p, with_values, x, sink, sink_len, grid, (hipStream_t)s);
}

bsp_status bsp_gather_probe(const double* x, int64_t n, int32_t slices, int64_t gathers,
                            double* sink, int64_t sink_len, int64_t* grid, bsp_stream s) {
  BSP_REQUIRE(x || !sink, "null argument");
  return gather_probe(x, n, slices, gathers, sink, sink_len, grid, (hipStream_t)s);
}

bsp_status bsp_line_probe(const double* x, int64_t n, int32_t slices, int64_t nbytes,
                          double* sink, int64_t sink_len, int64_t* grid, bsp_stream s) {
  BSP_REQUIRE(x || !sink, "null argument");
  return line_probe(x, n, slices, nbytes, sink, sink_len, grid, (hipStream_t)s);
}

bsp_status bsp_csr_check(const bsp_csr* A, bsp_stream s, int32_t* sorted) {
  BSP_TRY(check_csr_pattern(A));
  if (sorted) *sorted = 1;
  if (A->nrows == 0) {
    BSP_REQUIRE(A->nnz == 0, "a CSR without rows has no entries");
    return BSP_OK;
  }
  int f = 0;
  BSP_TRY(csr_check(*A, (hipStream_t)s, &f));
  if (f & 1) return fail(BSP_ERR_INVALID, "rowptr must start at 0, never decrease and end at nnz");
  if (f & 2) return fail(BSP_ERR_INVALID, "column index out of range");
  if (sorted) *sorted = (f & 4) ? 0 : 1;
  return BSP_OK;
}

bsp_status bsp_csr_plan_narrow(bsp_spmv_plan p, int32_t* narrow) {
  BSP_REQUIRE(p && narrow, "null argument");
  *narrow = p->narrow ? 1 : 0;
  return BSP_OK;
}

bsp_status bsp_csr_spmv(bsp_spmv_plan p, const double* x, double* y, bsp_stream s) {
  BSP_REQUIRE(p, "null plan");
  BSP_REQUIRE(p->A.dtype == BSP_F64, "plan has FP16 values: use bsp_csr_spmv_f16");
  BSP_REQUIRE((p->A.ncols == 0 || x) && (p->A.nrows == 0 || y), "null x/y");
  return csr_plan_run(p, x, y, (hipStream_t)s);
}

bsp_status bsp_csr_spmv_f16(bsp_spmv_plan p, const uint16_t* x, float* y, bsp_stream s) {
  BSP_REQUIRE(p, "null plan");
  BSP_REQUIRE(p->A.dtype == BSP_F16, "plan has no FP16 values");
  BSP_REQUIRE((p->A.ncols == 0 || x) && (p->A.nrows == 0 || y), "null x/y");
  return csr_plan_run_f16(p, x, y, (hipStream_t)s);
}

bsp_status bsp_csr_plan_info(bsp_spmv_plan p, int64_t* nblocks, int64_t* nlong) {
  BSP_REQUIRE(p, "null plan");
  if (nblocks) *nblocks = p->nblk;
  if (nlong) *nlong = p->nlong;
  return BSP_OK;
}

bsp_status bsp_plan_destroy(bsp_spmv_plan p) {
  delete p;
  return BSP_OK;
}

int bsp_spmv_fp64_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                          int32_t nnz, double* values, double* x, double* outvalues, int repeat,
                          double* avg_time) {
  return csr_host_wrapper(BSP_CSR_SCALAR, 10, rowPtr, colIdx, rows, cols, nnz, values, x,
                          outvalues, repeat, avg_time);
}

int bsp_spmv_fp64_balance_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                                  int32_t nnz, double* values, double* x, double* outvalues,
                                  int repeat, double* avg_time) {
  return csr_host_wrapper(BSP_CSR_BALANCED, 10, rowPtr, colIdx, rows, cols, nnz, values, x,
                          outvalues, repeat, avg_time);
}

}  // extern "C"

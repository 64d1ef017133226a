// Reference call surface: host-pointer wrappers with the semantics of the
// reference's host functions (upload, build on the GPU, warmup + timed
// launches each followed by a synchronisation, download), plus the rocSPARSE
// comparator. Only public C-ABI entry points of the library are used here.
#include <hip/hip_runtime.h>
#include <rocsparse/rocsparse.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <vector>

#include "bsp_hip.hpp"

using namespace bsp;

namespace {

constexpr int kTB = 256;

__global__ void i64_to_i32(const int64_t* __restrict__ a, int64_t n, int32_t* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (int32_t)a[i];
}
__global__ void f32_to_f64(const float* __restrict__ a, int64_t n, double* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (double)a[i];
}
__global__ void f32_to_f16(const float* __restrict__ a, int64_t n, uint16_t* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = __builtin_bit_cast(uint16_t, (_Float16)a[i]);
}
__global__ void fill_u16(uint16_t* __restrict__ a, int64_t n, uint16_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ void f16_to_f32(const uint16_t* __restrict__ a, int64_t n, float* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)__builtin_bit_cast(_Float16, a[i]);
}
__global__ void f32_round_f16(const float* __restrict__ a, int64_t n, uint16_t* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = __builtin_bit_cast(uint16_t, (_Float16)a[i]);
}

// Thread per row, left to right, no contraction: the reference CPU loop's
// result bit for bit (src/ref.cpp:3-24). HALF rounds inputs and y to FP16.
template <bool HALF>
__global__ __launch_bounds__(kTB) void rows_seq_f32(const int32_t* __restrict__ rowptr,
                                                    const int32_t* __restrict__ col,
                                                    const float* __restrict__ val,
                                                    const float* __restrict__ x,
                                                    float* __restrict__ y, int64_t n) {
#pragma clang fp contract(off)
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float s = 0.f;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) {
    float a = val[j], b = x[col[j]];
    if (HALF) {
      a = (float)(_Float16)a;
      b = (float)(_Float16)b;
    }
    s += a * b;
  }
  y[r] = HALF ? (float)(_Float16)s : s;
}

template <class T>
bsp_status up(DevBuf<T>& d, const T* h, int64_t n) {
  BSP_TRY(d.alloc(n));
  if (n > 0) BSP_HIP(hipMemcpy(d.get(), h, sizeof(T) * n, hipMemcpyHostToDevice));
  return BSP_OK;
}

template <class T>
bsp_status down(T* h, const T* d, int64_t n) {
  if (n > 0) BSP_HIP(hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost));
  return BSP_OK;
}

// host-clock timing of launch+sync pairs, whole microseconds per launch
template <class F>
bsp_status timed(int warmup, int repeat, double* avg_time, F body) {
  for (int i = 0; i < warmup; ++i) {
    BSP_TRY(body());
    BSP_HIP(hipDeviceSynchronize());
  }
  auto t0 = std::chrono::high_resolution_clock::now();
  for (int i = 0; i < repeat; ++i) {
    BSP_TRY(body());
    BSP_HIP(hipDeviceSynchronize());
  }
  auto t1 = std::chrono::high_resolution_clock::now();
  if (avg_time) {
    long long us = std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count();
    *avg_time = repeat > 0 ? (double)(us / repeat) : 0.0;
  }
  return BSP_OK;
}

struct DevCsr {
  DevBuf<int32_t> rp, ci;
  DevBuf<double> v;
  bsp_csr A{};
};

bsp_status upload_i32(DevCsr& c, const int32_t* rp, const int32_t* ci, const double* v,
                      int64_t rows, int64_t cols, int64_t nnz) {
  BSP_TRY(check_host_csr(rp, ci, rows, cols, nnz));
  BSP_REQUIRE(nnz == 0 || v, "null host pointer");
  BSP_TRY(up(c.rp, rp, rows + 1));
  BSP_TRY(up(c.ci, ci, nnz));
  BSP_TRY(up(c.v, v, nnz));
  c.A = bsp_csr{rows, cols, nnz, c.rp.get(), c.ci.get(), c.v.get(), BSP_F64};
  return BSP_OK;
}

bsp_status upload_i64(DevCsr& c, const int64_t* rp, const int64_t* ci, const double* v,
                      int64_t rows, int64_t cols, int64_t nnz) {
  // a pattern that passes has every index below 2^31, so narrowing is exact
  BSP_TRY(check_host_csr(rp, ci, rows, cols, nnz));
  BSP_REQUIRE(nnz == 0 || v, "null host pointer");
  DevBuf<int64_t> t;
  BSP_TRY(up(t, rp, rows + 1));
  BSP_TRY(c.rp.alloc(rows + 1));
  i64_to_i32<<<cdiv(rows + 1, kTB), kTB>>>(t.get(), rows + 1, c.rp.get());
  BSP_TRY(up(t, ci, nnz));
  BSP_TRY(c.ci.alloc(nnz));
  if (nnz) i64_to_i32<<<cdiv(nnz, kTB), kTB>>>(t.get(), nnz, c.ci.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(up(c.v, v, nnz));
  BSP_HIP(hipDeviceSynchronize());
  c.A = bsp_csr{rows, cols, nnz, c.rp.get(), c.ci.get(), c.v.get(), BSP_F64};
  return BSP_OK;
}

template <class T>
T* host_copy(const T* d, int64_t n) {
  T* h = (T*)std::malloc(sizeof(T) * (n > 0 ? n : 1));
  if (h && n > 0 && hipMemcpy(h, d, sizeof(T) * n, hipMemcpyDeviceToHost) != hipSuccess) {
    std::free(h);
    return nullptr;
  }
  return h;
}

bsp_status dbsr_spmv_host(bsp_dbsrMat* m, double* hX, double* hY, int warmup, int repeat,
                          double* avg_time) {
  BSP_REQUIRE(m && hX && hY, "null argument");
  bsp_dbsr h = nullptr;
  BSP_TRY(bsp_dbsr_from_host(m, nullptr, &h));
  std::unique_ptr<bsp_dbsr_s, bsp_status (*)(bsp_dbsr)> own(h, bsp_dbsr_destroy);
  DevBuf<double> x, y;
  BSP_TRY(up(x, hX, m->col));
  BSP_TRY(y.alloc(m->row));
  BSP_TRY(timed(warmup, repeat, avg_time,
                [&] { return bsp_dbsr_spmv(h, x.get(), y.get(), 1, nullptr); }));
  return down(hY, y.get(), m->row);
}

}  // namespace

// rocsparse_spmv with explicit stages is the API of this ROCm release.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wdeprecated-declarations"

struct bsp_rocsparse_plan_s {
  rocsparse_handle handle = nullptr;
  rocsparse_spmat_descr mat = nullptr;
  rocsparse_dnvec_descr x = nullptr, y = nullptr;
  rocsparse_spmv_alg alg = rocsparse_spmv_alg_default;
  rocsparse_datatype ty = rocsparse_datatype_f64_r;  // y and compute type
  DevBuf<char> buffer;
  size_t bytes = 0;
  bsp_csr A{};
  rocsparse_status call(rocsparse_spmv_stage stage, void* buf) {
    const double d1 = 1.0, d0 = 0.0;
    const float f1 = 1.f, f0 = 0.f;
    const bool dbl = ty == rocsparse_datatype_f64_r;
    return rocsparse_spmv(handle, rocsparse_operation_none, dbl ? (const void*)&d1 : &f1, mat, x,
                          dbl ? (const void*)&d0 : &f0, y, ty, alg, stage, &bytes, buf);
  }
  ~bsp_rocsparse_plan_s() {
    if (x) rocsparse_destroy_dnvec_descr(x);
    if (y) rocsparse_destroy_dnvec_descr(y);
    if (mat) rocsparse_destroy_spmat_descr(mat);
    if (handle) rocsparse_destroy_handle(handle);
  }
};

#define BSP_RS(call)                                                                         \
  do {                                                                                       \
    rocsparse_status e_ = (call);                                                            \
    if (e_ != rocsparse_status_success)                                                      \
      return fail(BSP_ERR_HIP, std::string(#call) + ": rocsparse status " + std::to_string(e_)); \
  } while (0)

extern "C" {

bsp_status bsp_rocsparse_plan_create(const bsp_csr* A, int alg, bsp_stream s,
                                     bsp_rocsparse_plan* out) {
  BSP_REQUIRE(A && out, "null argument");
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(A->dtype == BSP_F64 || A->dtype == BSP_F32 || A->dtype == BSP_F16,
              "rocSPARSE comparator takes FP64, FP32 or FP16 values");
  BSP_REQUIRE(alg == 0 || alg == 2 || alg == 3 || alg == 7,
              "rocSPARSE CSR algorithm must be 0 (default), 2 (adaptive), 3 (rowsplit) or 7 "
              "(LRB)");
  auto p = std::make_unique<bsp_rocsparse_plan_s>();
  p->A = *A;
  p->alg = alg == 2   ? rocsparse_spmv_alg_csr_adaptive
           : alg == 3 ? rocsparse_spmv_alg_csr_rowsplit
           : alg == 7 ? rocsparse_spmv_alg_csr_lrb
                      : rocsparse_spmv_alg_default;
  const rocsparse_datatype ta = A->dtype == BSP_F64   ? rocsparse_datatype_f64_r
                                : A->dtype == BSP_F32 ? rocsparse_datatype_f32_r
                                                      : rocsparse_datatype_f16_r;
  p->ty = A->dtype == BSP_F64 ? rocsparse_datatype_f64_r : rocsparse_datatype_f32_r;
  BSP_RS(rocsparse_create_handle(&p->handle));
  BSP_RS(rocsparse_set_stream(p->handle, (hipStream_t)s));
  BSP_RS(rocsparse_create_csr_descr(&p->mat, A->nrows, A->ncols, A->nnz, (void*)A->rowptr,
                                    (void*)A->colidx, (void*)A->val, rocsparse_indextype_i32,
                                    rocsparse_indextype_i32, rocsparse_index_base_zero, ta));
  DevBuf<double> xd, yd;  // placeholders until the first run sets x and y
  BSP_TRY(xd.alloc(A->ncols));
  BSP_TRY(yd.alloc(A->nrows));
  BSP_RS(rocsparse_create_dnvec_descr(&p->x, A->ncols, xd.get(), ta));
  BSP_RS(rocsparse_create_dnvec_descr(&p->y, A->nrows, yd.get(), p->ty));
  BSP_RS(p->call(rocsparse_spmv_stage_buffer_size, nullptr));
  BSP_TRY(p->buffer.alloc(p->bytes));
  BSP_RS(p->call(rocsparse_spmv_stage_preprocess, p->buffer.get()));
  BSP_HIP(hipStreamSynchronize((hipStream_t)s));
  *out = p.release();
  return BSP_OK;
}

bsp_status bsp_rocsparse_spmv(bsp_rocsparse_plan p, const void* x, void* y, bsp_stream s) {
  BSP_REQUIRE(p && x && y, "null argument");
  BSP_RS(rocsparse_set_stream(p->handle, (hipStream_t)s));
  BSP_RS(rocsparse_dnvec_set_values(p->x, (void*)x));
  BSP_RS(rocsparse_dnvec_set_values(p->y, y));
  BSP_RS(p->call(rocsparse_spmv_stage_compute, p->buffer.get()));
  return BSP_OK;
}

bsp_status bsp_rocsparse_plan_destroy(bsp_rocsparse_plan p) {
  delete p;
  return BSP_OK;
}

}  // extern "C"

struct bsp_rocsparse_spmm_plan_s {
  rocsparse_handle handle = nullptr;
  rocsparse_spmat_descr mat = nullptr;
  rocsparse_dnmat_descr B = nullptr, C = nullptr;
  rocsparse_spmm_alg alg = rocsparse_spmm_alg_default;
  rocsparse_operation op = rocsparse_operation_none;
  DevBuf<char> buffer;
  size_t bytes = 0;
  ~bsp_rocsparse_spmm_plan_s() {
    if (B) rocsparse_destroy_dnmat_descr(B);
    if (C) rocsparse_destroy_dnmat_descr(C);
    if (mat) rocsparse_destroy_spmat_descr(mat);
    if (handle) rocsparse_destroy_handle(handle);
  }
  rocsparse_status call(rocsparse_spmm_stage stage, void* buf) {
    const float one = 1.f, zero = 0.f;
    return rocsparse_spmm(handle, op, rocsparse_operation_none, &one, mat, B, &zero, C,
                          rocsparse_datatype_f32_r, alg, stage, &bytes, buf);
  }
};

extern "C" {

bsp_status bsp_rocsparse_spmm_plan_create(const bsp_csr* A, int64_t N, int transpose, int alg,
                                          bsp_stream s, bsp_rocsparse_spmm_plan* out) {
  BSP_REQUIRE(A && out && N > 0, "bad argument");
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(A->dtype == BSP_F32 || A->dtype == BSP_F16 || A->dtype == BSP_BF16,
              "rocSPARSE SpMM comparator takes FP32, FP16 or BF16 values");
  BSP_REQUIRE(alg == 0 || alg == 1 || alg == 4 || alg == 5 || alg == 9,
              "rocSPARSE CSR SpMM algorithm must be 0 (default), 1 (csr), 4 (row split), "
              "5 (nnz split) or 9 (merge path)");
  auto p = std::make_unique<bsp_rocsparse_spmm_plan_s>();
  p->alg = alg == 1   ? rocsparse_spmm_alg_csr
           : alg == 4 ? rocsparse_spmm_alg_csr_row_split
           : alg == 5 ? rocsparse_spmm_alg_csr_nnz_split
           : alg == 9 ? rocsparse_spmm_alg_csr_merge_path
                      : rocsparse_spmm_alg_default;
  p->op = transpose ? rocsparse_operation_transpose : rocsparse_operation_none;
  const rocsparse_datatype ta = A->dtype == BSP_F32   ? rocsparse_datatype_f32_r
                                : A->dtype == BSP_F16 ? rocsparse_datatype_f16_r
                                                      : rocsparse_datatype_bf16_r;
  const int64_t brows = transpose ? A->nrows : A->ncols, crows = transpose ? A->ncols : A->nrows;
  BSP_RS(rocsparse_create_handle(&p->handle));
  BSP_RS(rocsparse_set_stream(p->handle, (hipStream_t)s));
  BSP_RS(rocsparse_create_csr_descr(&p->mat, A->nrows, A->ncols, A->nnz, (void*)A->rowptr,
                                    (void*)A->colidx, (void*)A->val, rocsparse_indextype_i32,
                                    rocsparse_indextype_i32, rocsparse_index_base_zero, ta));
  DevBuf<char> bd, cd;  // placeholders until the first run sets B and C
  BSP_TRY(bd.alloc(brows * N * 4));
  BSP_TRY(cd.alloc(crows * N * 4));
  BSP_RS(rocsparse_create_dnmat_descr(&p->B, brows, N, N, bd.get(), ta, rocsparse_order_row));
  BSP_RS(rocsparse_create_dnmat_descr(&p->C, crows, N, N, cd.get(), rocsparse_datatype_f32_r,
                                      rocsparse_order_row));
  BSP_RS(p->call(rocsparse_spmm_stage_buffer_size, nullptr));
  BSP_TRY(p->buffer.alloc(p->bytes));
  BSP_RS(p->call(rocsparse_spmm_stage_preprocess, p->buffer.get()));
  BSP_HIP(hipStreamSynchronize((hipStream_t)s));
  *out = p.release();
  return BSP_OK;
}

bsp_status bsp_rocsparse_spmm(bsp_rocsparse_spmm_plan p, const void* B, float* C, bsp_stream s) {
  BSP_REQUIRE(p && B && C, "null argument");
  BSP_RS(rocsparse_set_stream(p->handle, (hipStream_t)s));
  BSP_RS(rocsparse_dnmat_set_values(p->B, (void*)B));
  BSP_RS(rocsparse_dnmat_set_values(p->C, C));
  BSP_RS(p->call(rocsparse_spmm_stage_compute, p->buffer.get()));
  return BSP_OK;
}

bsp_status bsp_rocsparse_spmm_plan_destroy(bsp_rocsparse_spmm_plan p) {
  delete p;
  return BSP_OK;
}

int bsp_cusparse_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values,
                           double* hX, double* hY, int64_t A_num_rows, int64_t A_num_cols,
                           int64_t A_nnz, int repeat, double* avg_time) {
  DevCsr c;
  if (upload_i64(c, hA_csrOffsets, hA_columns, hA_values, A_num_rows, A_num_cols, A_nnz) != BSP_OK)
    return 1;
  if (!hX || !hY) return 1;
  DevBuf<double> x, y;
  if (up(x, hX, A_num_cols) != BSP_OK || y.alloc(A_num_rows) != BSP_OK) return 1;
  bsp_rocsparse_plan p = nullptr;
  if (bsp_rocsparse_plan_create(&c.A, 0, nullptr, &p) != BSP_OK) return 1;
  std::unique_ptr<bsp_rocsparse_plan_s> own(p);
  if (timed(10, repeat, avg_time, [&] { return bsp_rocsparse_spmv(p, x.get(), y.get(), nullptr); }) !=
      BSP_OK)
    return 1;
  return down(hY, y.get(), A_num_rows) == BSP_OK ? 0 : 1;
}

int bsp_csrTodbsrWithRecorder(const bsp_CSRFormat_f64* csr, bsp_dbsrMat* out) {
  if (!csr || !out) return 1;
  DevCsr c;
  if (upload_i32(c, csr->rowPtr, csr->colIdx, csr->values, csr->row, csr->col, csr->nnz) != BSP_OK)
    return 1;
  bsp_dbsr h = nullptr;
  if (bsp_dbsr_build(&c.A, nullptr, &h) != BSP_OK) return 1;
  std::unique_ptr<bsp_dbsr_s, bsp_status (*)(bsp_dbsr)> own(h, bsp_dbsr_destroy);
  bsp_dbsr_view v;
  bsp_dbsr_view_get(h, &v);
  std::memset(out, 0, sizeof(*out));
  out->row = (int)v.nrows;
  out->col = (int)v.ncols;
  out->nnz = (int)v.nnz;
  out->blc_row = (int)v.brows;
  out->blc_num = (int)v.nblocks;
  out->avg_nnz = v.nblocks ? (double)v.nnz / (double)v.nblocks : 0.0;
  out->blocknnz = (int32_t)(v.nblocks * 16);
  out->warpnum = v.warpnum;
  out->blcPtr = host_copy(v.blcptr, v.brows + 1);
  out->colIdx = host_copy(v.colidx, v.nblocks * 16);
  out->blcVal = host_copy(v.val, v.nblocks * 16);
  out->blcMap = host_copy(v.map, v.nblocks);
  out->recorderMap = host_copy(v.recorder, v.nrows);
  out->prefix = host_copy(v.prefix, (int64_t)v.warpnum + 1);
  if (!out->blcPtr || !out->colIdx || !out->blcVal || !out->blcMap || !out->recorderMap ||
      !out->prefix) {
    bsp_dbsrMat_free(out);
    return 1;
  }
  return 0;
}

int bsp_csrTodbsr(const bsp_CSRFormat_f64* csr, bsp_dbsrMat* out) {
  if (!csr || !out) return 1;
  DevCsr c;
  if (upload_i32(c, csr->rowPtr, csr->colIdx, csr->values, csr->row, csr->col, csr->nnz) != BSP_OK)
    return 1;
  bsp_dbsr h = nullptr;
  if (dbsr_build_ordered(c.A, false, nullptr, &h) != BSP_OK) return 1;
  std::unique_ptr<bsp_dbsr_s, bsp_status (*)(bsp_dbsr)> own(h, bsp_dbsr_destroy);
  bsp_dbsr_view v;
  bsp_dbsr_view_get(h, &v);
  std::memset(out, 0, sizeof(*out));  // recorderMap and prefix stay null
  out->row = (int)v.nrows;
  out->col = (int)v.ncols;
  out->nnz = (int)v.nnz;
  out->blc_row = (int)v.brows;
  out->blc_num = (int)v.nblocks;
  out->avg_nnz = v.nblocks ? (double)v.nnz / (double)v.nblocks : 0.0;
  out->blocknnz = (int32_t)(v.nblocks * 16);
  out->blcPtr = host_copy(v.blcptr, v.brows + 1);
  out->colIdx = host_copy(v.colidx, v.nblocks * 16);
  out->blcVal = host_copy(v.val, v.nblocks * 16);
  out->blcMap = host_copy(v.map, v.nblocks);
  if (!out->blcPtr || !out->colIdx || !out->blcVal || !out->blcMap) {
    bsp_dbsrMat_free(out);
    fail(BSP_ERR_ALLOC, "host allocation failed");
    return 1;
  }
  return 0;
}

int bsp_balanceDbsr(bsp_dbsrMat* m) {
  if (!m) return 1;
  if (m->blc_row < 0 || !m->blcPtr || m->blcPtr[0] != 0) {
    fail(BSP_ERR_INVALID, "balanceDbsr: blcPtr missing or not starting at 0");
    return 1;
  }
  for (int64_t g = 0; g < m->blc_row; ++g)
    if (m->blcPtr[g + 1] < m->blcPtr[g]) {
      fail(BSP_ERR_INVALID, "balanceDbsr: blcPtr decreases at block row " + std::to_string(g));
      return 1;
    }
  int32_t* prefix = nullptr;
  int32_t wn = 0;
  if (dbsr_balance_host(m->blcPtr, m->blc_row, &prefix, &wn) != BSP_OK) return 1;
  std::free(m->prefix);
  m->prefix = prefix;
  m->warpnum = wn;
  return 0;
}

int bsp_csrRecorderSelf(const bsp_CSRFormat_f64* csr, bsp_CSRFormat_f64* out) {
  if (!csr || !out) return 1;
  DevCsr c;
  if (upload_i32(c, csr->rowPtr, csr->colIdx, csr->values, csr->row, csr->col, csr->nnz) != BSP_OK)
    return 1;
  DevBuf<int32_t> rp, ci;
  DevBuf<double> v;
  if (rp.alloc((int64_t)csr->row + 1) != BSP_OK || ci.alloc(csr->nnz) != BSP_OK ||
      v.alloc(csr->nnz) != BSP_OK)
    return 1;
  if (csr_sort_rows(c.A, nullptr, rp.get(), ci.get(), v.get()) != BSP_OK) return 1;
  int32_t* hrp = host_copy(rp.get(), (int64_t)csr->row + 1);
  int32_t* hci = host_copy(ci.get(), csr->nnz);
  double* hv = host_copy(v.get(), csr->nnz);
  if (!hrp || !hci || !hv) {
    std::free(hrp);
    std::free(hci);
    std::free(hv);
    fail(BSP_ERR_ALLOC, "host allocation failed");
    return 1;
  }
  out->row = csr->row;
  out->col = csr->col;
  out->nnz = csr->nnz;
  out->rowPtr = hrp;
  out->colIdx = hci;
  out->values = hv;
  return 0;
}

int bsp_csrToBcsr(const int64_t* rowPtr, const int64_t* colIdx, const double* values, int64_t row,
                  int64_t col, int tile_m, int tile_k, int64_t* nonzeroBlocks,
                  int64_t** bcsrRowPtr, int64_t** bcsrColIdx, double** bcsrVal) {
  if (!nonzeroBlocks || !bcsrRowPtr || !bcsrColIdx || !bcsrVal) return 1;
  const int64_t nnz = rowPtr && row >= 0 ? rowPtr[row] : -1;
  DevCsr c;
  if (upload_i64(c, rowPtr, colIdx, values, row, col, nnz) != BSP_OK) return 1;
  bsp_bsr b = nullptr;
  if (bsp_bsr_build(&c.A, tile_m, tile_k, nullptr, &b) != BSP_OK) return 1;
  std::unique_ptr<bsp_bsr_s, bsp_status (*)(bsp_bsr)> own(b, bsp_bsr_destroy);
  bsp_bsr_view v;
  bsp_bsr_view_get(b, &v);
  std::vector<int32_t> rp32(v.brows + 1), ci32(v.nblocks);
  const int64_t nval = v.nblocks * (int64_t)tile_m * tile_k;
  int64_t* orp = (int64_t*)std::malloc(sizeof(int64_t) * (v.brows + 1));
  int64_t* oci = (int64_t*)std::malloc(sizeof(int64_t) * (v.nblocks ? v.nblocks : 1));
  double* ov = (double*)std::malloc(sizeof(double) * (nval ? nval : 1));
  if (!orp || !oci || !ov || down(rp32.data(), v.rowptr, v.brows + 1) != BSP_OK ||
      down(ci32.data(), v.colidx, v.nblocks) != BSP_OK || down(ov, v.val, nval) != BSP_OK) {
    std::free(orp);
    std::free(oci);
    std::free(ov);
    return 1;
  }
  for (int64_t i = 0; i <= v.brows; ++i) orp[i] = rp32[i];
  for (int64_t i = 0; i < v.nblocks; ++i) oci[i] = ci32[i];
  *nonzeroBlocks = v.nblocks;
  *bcsrRowPtr = orp;
  *bcsrColIdx = oci;
  *bcsrVal = ov;
  return 0;
}

void bsp_free(void* p) { std::free(p); }

void bsp_dbsrMat_free(bsp_dbsrMat* m) {
  if (!m) return;
  std::free(m->blcPtr);
  std::free(m->colIdx);
  std::free(m->blcVal);
  std::free(m->blcMap);
  std::free(m->recorderMap);
  std::free(m->prefix);
  m->blcPtr = m->colIdx = m->recorderMap = m->prefix = nullptr;
  m->blcVal = nullptr;
  m->blcMap = nullptr;
}

int bsp_amgT_spmv_fp64_dbsr(bsp_dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                            int32_t repeat, double* avg_time) {
  return dbsr_spmv_host(dbsr, hX, hY, warmup, repeat, avg_time) == BSP_OK ? 0 : 1;
}

int bsp_amgT_spmv_fp64_dbsr_balance(bsp_dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                                    int32_t repeat, double* avg_time) {
  return dbsr_spmv_host(dbsr, hX, hY, warmup, repeat, avg_time) == BSP_OK ? 0 : 1;
}

int bsp_amgT_spmv_fp64(int32_t* hA_csrOffsets, int32_t* hA_columns, double* hA_values, double* hX,
                       double* hY, int32_t A_num_rows, int32_t A_num_cols, int32_t A_nnz,
                       int repeat, double* avg_time) {
  DevCsr c;
  if (upload_i32(c, hA_csrOffsets, hA_columns, hA_values, A_num_rows, A_num_cols, A_nnz) != BSP_OK)
    return 1;
  if (!hX || !hY) return 1;
  bsp_bsr b = nullptr;
  if (bsp_bsr_build(&c.A, 4, 4, nullptr, &b) != BSP_OK) return 1;
  std::unique_ptr<bsp_bsr_s, bsp_status (*)(bsp_bsr)> own(b, bsp_bsr_destroy);
  bsp_bsr_view v;
  bsp_bsr_view_get(b, &v);
  // BSR_BALANCED_PREPROCESS_GPU sets stand = 12 (src/amgT.cu:1176), so the
  // reference always takes its balanced branch (:1394-1407): MMA when
  // avgnz >= 10, CUDA cores with the bitmap otherwise
  const int kernel = v.avg_nnz >= 10.0 ? BSP_BSR_MFMA_BALANCED : BSP_BSR_VALU_BALANCED;
  DevBuf<double> x, y;
  if (up(x, hX, A_num_cols) != BSP_OK || y.alloc(A_num_rows) != BSP_OK) return 1;
  if (timed(5, repeat, avg_time, [&] { return bsp_bsr_spmv(b, kernel, x.get(), y.get(), nullptr); }) !=
      BSP_OK)
    return 1;
  return down(hY, y.get(), A_num_rows) == BSP_OK ? 0 : 1;
}

void bsp_bcsr_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values,
                        double* hX, double* hY, int64_t A_num_rows, int64_t A_num_cols,
                        int64_t A_nnz, int repeat) {
  DevCsr c;
  if (upload_i64(c, hA_csrOffsets, hA_columns, hA_values, A_num_rows, A_num_cols, A_nnz) != BSP_OK)
    return;
  if (!hX || !hY) return;
  bsp_bsr b = nullptr;
  if (bsp_bsr_build(&c.A, 3, 3, nullptr, &b) != BSP_OK) return;
  std::unique_ptr<bsp_bsr_s, bsp_status (*)(bsp_bsr)> own(b, bsp_bsr_destroy);
  DevBuf<double> x, y;
  if (up(x, hX, A_num_cols) != BSP_OK || y.alloc(A_num_rows) != BSP_OK) return;
  double t = 0;
  if (timed(0, repeat > 0 ? repeat : 1, &t,
            [&] { return bsp_bsr_spmv(b, BSP_BSR_VALU, x.get(), y.get(), nullptr); }) != BSP_OK)
    return;
  down(hY, y.get(), A_num_rows);
}

void bsp_spmv_all(char* filename, double* csrValA, int* csrRowPtrA, int* csrColIdxA, double* X_val,
                  double* Y_val, int* order_rid, int rowA, int colA, int nnzA, int NUM,
                  double threshold, int block_longest, double* avg_time) {
  (void)filename;
  (void)NUM;
  DevCsr c;
  if (upload_i32(c, csrRowPtrA, csrColIdxA, csrValA, rowA, colA, nnzA) != BSP_OK) return;
  if (!X_val || !Y_val || !order_rid) return;
  bsp_dasp d = nullptr;
  if (bsp_dasp_build(&c.A, threshold, block_longest, nullptr, &d) != BSP_OK) return;
  std::unique_ptr<bsp_dasp_s, bsp_status (*)(bsp_dasp)> own(d, bsp_dasp_destroy);
  DevBuf<double> x, y;
  if (up(x, X_val, colA) != BSP_OK || y.alloc(rowA) != BSP_OK) return;
  if (timed(100, 10, avg_time, [&] { return bsp_dasp_spmv(d, x.get(), y.get(), 1, nullptr); }) !=
      BSP_OK)
    return;
  bsp_dasp_view v;
  bsp_dasp_view_get(d, &v);
  down(Y_val, y.get(), rowA);
  down(order_rid, v.order_rid, rowA);
}

void bsp_MySpmvWithPrefp64(int64_t* rowPtr, int64_t* colIdx, int64_t rows, int64_t cols,
                           int64_t nnz, double* values, double* x, double* outvalues) {
  DevCsr c;
  if (upload_i64(c, rowPtr, colIdx, values, rows, cols, nnz) != BSP_OK) return;
  if (!x || !outvalues) return;
  bsp_spmv_plan p = nullptr;
  if (bsp_csr_plan_create(&c.A, BSP_CSR_SPLIT, nullptr, &p) != BSP_OK) return;
  std::unique_ptr<bsp_spmv_plan_s, bsp_status (*)(bsp_spmv_plan)> own(p, bsp_plan_destroy);
  DevBuf<double> xd, yd;
  if (up(xd, x, cols) != BSP_OK || yd.alloc(rows) != BSP_OK) return;
  double t = 0;
  if (timed(0, 10, &t, [&] { return bsp_csr_spmv(p, xd.get(), yd.get(), nullptr); }) != BSP_OK)
    return;
  down(outvalues, yd.get(), rows);
}

void bsp_MySpmvWithPrefp16(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                           int32_t nnz, float* values, float* x, float* outvalues) {
  if (check_host_csr(rowPtr, colIdx, rows, cols, nnz) != BSP_OK || (nnz && !values) || !x ||
      !outvalues)
    return;
  DevBuf<int32_t> rp, ci;
  DevBuf<float> v32, x32;
  DevBuf<uint16_t> v16, x16;
  DevBuf<float> y;
  if (up(rp, rowPtr, rows + 1) != BSP_OK || up(ci, colIdx, nnz) != BSP_OK ||
      up(v32, values, nnz) != BSP_OK || up(x32, x, cols) != BSP_OK || v16.alloc(nnz) != BSP_OK ||
      x16.alloc(cols) != BSP_OK || y.alloc(rows) != BSP_OK)
    return;
  if (nnz) f32_to_f16<<<cdiv(nnz, kTB), kTB>>>(v32.get(), nnz, v16.get());
  if (cols) f32_to_f16<<<cdiv(cols, kTB), kTB>>>(x32.get(), cols, x16.get());
  if (hipDeviceSynchronize() != hipSuccess) return;
  bsp_csr A{rows, cols, nnz, rp.get(), ci.get(), v16.get(), BSP_F16};
  bsp_spmv_plan p = nullptr;
  if (bsp_csr_plan_create(&A, BSP_CSR_SPLIT, nullptr, &p) != BSP_OK) return;
  std::unique_ptr<bsp_spmv_plan_s, bsp_status (*)(bsp_spmv_plan)> own(p, bsp_plan_destroy);
  double t = 0;
  if (timed(0, 10, &t, [&] { return bsp_csr_spmv_f16(p, x16.get(), y.get(), nullptr); }) != BSP_OK)
    return;
  down(outvalues, y.get(), rows);
}

void bsp_smatSpmmfp16(int* rowPtr, int* colIdx, int row, int col, int n, int nnz, float* values,
                      float* outvalues) {
  if (check_host_csr(rowPtr, colIdx, row, col, nnz) != BSP_OK || n <= 0 || n % 32 != 0 ||
      !outvalues || (nnz && !values))
    return;
  DevCsr c;
  DevBuf<float> v32;
  if (up(c.rp, rowPtr, (int64_t)row + 1) != BSP_OK || up(c.ci, colIdx, nnz) != BSP_OK ||
      up(v32, values, nnz) != BSP_OK || c.v.alloc(nnz) != BSP_OK)
    return;
  if (nnz) f32_to_f64<<<cdiv(nnz, kTB), kTB>>>(v32.get(), nnz, c.v.get());
  c.A = bsp_csr{row, col, nnz, c.rp.get(), c.ci.get(), c.v.get(), BSP_F64};
  bsp_bsr b = nullptr;
  if (bsp_bsr_build(&c.A, 16, 16, nullptr, &b) != BSP_OK) return;
  std::unique_ptr<bsp_bsr_s, bsp_status (*)(bsp_bsr)> ownb(b, bsp_bsr_destroy);
  bsp_bspmm h = nullptr;
  if (bsp_bspmm_build(b, BSP_F16, nullptr, &h) != BSP_OK) return;
  std::unique_ptr<bsp_bspmm_s, bsp_status (*)(bsp_bspmm)> ownh(h, bsp_bspmm_destroy);
  // B = 1 (col x n) as in the reference (src/smat.cu:142-144), in column
  // chunks of at most 256
  const int chunk = n % 256 == 0 ? 256 : n % 128 == 0 ? 128 : n % 64 == 0 ? 64 : 32;
  DevBuf<uint16_t> B;
  DevBuf<float> Cc, Cfull;
  if (B.alloc((int64_t)col * chunk) != BSP_OK || Cc.alloc((int64_t)row * chunk) != BSP_OK ||
      Cfull.alloc((int64_t)row * n) != BSP_OK)
    return;
  const uint16_t one = __builtin_bit_cast(uint16_t, (_Float16)1.0f);
  fill_u16<<<cdiv((int64_t)col * chunk, kTB), kTB>>>(B.get(), (int64_t)col * chunk, one);
  for (int c0 = 0; c0 < n; c0 += chunk) {
    if (bsp_bspmm_fwd(h, B.get(), chunk, Cc.get(), nullptr) != BSP_OK) return;
    if (row > 0 &&
        hipMemcpy2D(Cfull.get() + c0, sizeof(float) * n, Cc.get(), sizeof(float) * chunk,
                    sizeof(float) * chunk, row, hipMemcpyDeviceToDevice) != hipSuccess)
      return;
  }
  if (hipDeviceSynchronize() != hipSuccess) return;
  down(outvalues, Cfull.get(), (int64_t)row * n);
}

int bsp_cusparse_spmv(int* hA_csrOffsets, int64_t* hA_columns, float* hA_values, float* hX,
                      float* hY, int64_t A_num_rows, int64_t A_num_cols, int64_t A_nnz) {
  if (check_host_csr(hA_csrOffsets, hA_columns, A_num_rows, A_num_cols, A_nnz) != BSP_OK ||
      (A_nnz && !hA_values) || !hX || !hY)
    return 1;
  DevBuf<int32_t> rp, ci;
  DevBuf<int64_t> c64;
  DevBuf<float> v, x, y;
  if (up(rp, hA_csrOffsets, A_num_rows + 1) != BSP_OK || up(c64, hA_columns, A_nnz) != BSP_OK ||
      ci.alloc(A_nnz) != BSP_OK || up(v, hA_values, A_nnz) != BSP_OK ||
      up(x, hX, A_num_cols) != BSP_OK || y.alloc(A_num_rows) != BSP_OK)
    return 1;
  if (A_nnz) i64_to_i32<<<cdiv(A_nnz, kTB), kTB>>>(c64.get(), A_nnz, ci.get());
  bsp_csr A{A_num_rows, A_num_cols, A_nnz, rp.get(), ci.get(), v.get(), BSP_F32};
  bsp_rocsparse_plan p = nullptr;
  if (bsp_rocsparse_plan_create(&A, 0, nullptr, &p) != BSP_OK) return 1;
  std::unique_ptr<bsp_rocsparse_plan_s> own(p);
  if (bsp_rocsparse_spmv(p, x.get(), y.get(), nullptr) != BSP_OK) return 1;
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  return down(hY, y.get(), A_num_rows) == BSP_OK ? 0 : 1;
}

int bsp_cusparse_spmv_fp16(int* hA_csrOffsets, int32_t* hA_columns, const uint16_t* hA_values,
                           const uint16_t* hX, uint16_t* hY, int32_t A_num_rows,
                           int32_t A_num_cols, int32_t A_nnz) {
  if (check_host_csr(hA_csrOffsets, hA_columns, A_num_rows, A_num_cols, A_nnz) != BSP_OK ||
      (A_nnz && !hA_values) || !hX || !hY)
    return 1;
  DevBuf<int32_t> rp, ci;
  DevBuf<uint16_t> v, x, yh;
  DevBuf<float> y;
  if (up(rp, hA_csrOffsets, (int64_t)A_num_rows + 1) != BSP_OK || up(ci, hA_columns, A_nnz) != BSP_OK ||
      up(v, hA_values, A_nnz) != BSP_OK || up(x, hX, A_num_cols) != BSP_OK ||
      y.alloc(A_num_rows) != BSP_OK || yh.alloc(A_num_rows) != BSP_OK)
    return 1;
  bsp_csr A{A_num_rows, A_num_cols, A_nnz, rp.get(), ci.get(), v.get(), BSP_F16};
  bsp_rocsparse_plan p = nullptr;
  if (bsp_rocsparse_plan_create(&A, 0, nullptr, &p) != BSP_OK) return 1;
  std::unique_ptr<bsp_rocsparse_plan_s> own(p);
  if (bsp_rocsparse_spmv(p, x.get(), y.get(), nullptr) != BSP_OK) return 1;
  if (A_num_rows) f32_round_f16<<<cdiv(A_num_rows, kTB), kTB>>>(y.get(), A_num_rows, yh.get());
  if (hipDeviceSynchronize() != hipSuccess) return 1;
  return down(hY, yh.get(), A_num_rows) == BSP_OK ? 0 : 1;
}

void bsp_cusparse_spmv_fp16_warpper(int* hA_csrOffsets, int32_t* hA_columns, float* hA_values32,
                                    float* hX32, float* hY32, int32_t A_num_rows,
                                    int32_t A_num_cols, int32_t A_nnz) {
  if (check_host_csr(hA_csrOffsets, hA_columns, A_num_rows, A_num_cols, A_nnz) != BSP_OK ||
      !hX32 || !hY32 || (A_nnz && !hA_values32))
    return;
  std::vector<uint16_t> hv(A_nnz), hx(A_num_cols), hy(A_num_rows);
  DevBuf<float> t32;
  DevBuf<uint16_t> t16;
  // round to FP16 on the GPU (round to nearest even, as the (half) cast does)
  auto to_half = [&](const float* src, int64_t n, uint16_t* dst) {
    if (n == 0) return true;
    if (up(t32, src, n) != BSP_OK || t16.alloc(n) != BSP_OK) return false;
    f32_to_f16<<<cdiv(n, kTB), kTB>>>(t32.get(), n, t16.get());
    return down(dst, t16.get(), n) == BSP_OK;
  };
  if (!to_half(hA_values32, A_nnz, hv.data()) || !to_half(hX32, A_num_cols, hx.data())) return;
  if (bsp_cusparse_spmv_fp16(hA_csrOffsets, hA_columns, hv.data(), hx.data(), hy.data(), A_num_rows,
                             A_num_cols, A_nnz) != 0)
    return;
  DevBuf<float> yf;
  if (A_num_rows == 0 || up(t16, hy.data(), A_num_rows) != BSP_OK || yf.alloc(A_num_rows) != BSP_OK)
    return;
  f16_to_f32<<<cdiv(A_num_rows, kTB), kTB>>>(t16.get(), A_num_rows, yf.get());
  down(hY32, yf.get(), A_num_rows);
}

void bsp_ref_spmv_fp64(int64_t* A_csrOffsets, int64_t* A_columns, double* A_values, double* x,
                       double* y, int64_t row, int64_t col, double* avg_time) {
  if (!A_csrOffsets || row < 0 || !x || !y) return;
  const int64_t nnz = A_csrOffsets[row];
  DevCsr c;
  if (upload_i64(c, A_csrOffsets, A_columns, A_values, row, col, nnz) != BSP_OK) return;
  bsp_spmv_plan p = nullptr;
  if (bsp_csr_plan_create(&c.A, BSP_CSR_SCALAR, nullptr, &p) != BSP_OK) return;
  std::unique_ptr<bsp_spmv_plan_s, bsp_status (*)(bsp_spmv_plan)> own(p, bsp_plan_destroy);
  DevBuf<double> xd, yd;
  if (up(xd, x, col) != BSP_OK || yd.alloc(row) != BSP_OK) return;
  double t = 0;
  if (timed(0, 1, &t, [&] { return bsp_csr_spmv(p, xd.get(), yd.get(), nullptr); }) != BSP_OK)
    return;
  if (avg_time) *avg_time = t;
  down(y, yd.get(), row);
}

static void ref_rows_f32(bool half, int* rowptr, int32_t* cols, float* vals, float* x, float* y,
                         int32_t row, int32_t col) {
  if (!rowptr || row < 0 || col < 0 || !x || !y) return;
  const int64_t nnz = rowptr[row];
  if (check_host_csr(rowptr, cols, row, col, nnz) != BSP_OK || (nnz && !vals)) return;
  DevBuf<int32_t> rp, ci;
  DevBuf<float> v, xd, yd;
  if (up(rp, rowptr, (int64_t)row + 1) != BSP_OK || up(ci, cols, nnz) != BSP_OK ||
      up(v, vals, nnz) != BSP_OK || up(xd, x, col) != BSP_OK || yd.alloc(row) != BSP_OK)
    return;
  if (row == 0) return;
  if (half)
    rows_seq_f32<true><<<cdiv(row, kTB), kTB>>>(rp.get(), ci.get(), v.get(), xd.get(), yd.get(), row);
  else
    rows_seq_f32<false><<<cdiv(row, kTB), kTB>>>(rp.get(), ci.get(), v.get(), xd.get(), yd.get(), row);
  if (hipDeviceSynchronize() != hipSuccess) return;
  down(y, yd.get(), row);
}

void bsp_ref_spmv_fp32(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                       float* y, int32_t row, int32_t col) {
  ref_rows_f32(false, A_csrOffsets, A_columns, A_values, x, y, row, col);
}

void bsp_ref_spmv_fp16(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                       float* y, int32_t row, int32_t col) {
  ref_rows_f32(true, A_csrOffsets, A_columns, A_values, x, y, row, col);
}

}  // extern "C"

#pragma clang diagnostic pop

// Diagonal layout for banded and stencil matrices. Every distinct diagonal
// (column - row = d) of the matrix is stored as one contiguous run over the
// rows, with a per-row mask of the diagonals the row really holds, so the
// SpMV streams values with no column indices: a 27-point stencil moves
// 8 bytes per stored slot plus 4 per row where CSR moves 12 per entry.
//
// No counterpart in the reference (its layouts are CSR, BCSR, DBSR, DASP
// and the AmgT BSR, whose bitmap blocks at src/amgT.cu:160-209 are its
// answer to index bytes). The row sum adds the row's entries in ascending
// column order with separately rounded products, the order of the
// reference's row-sequential CSR loop (src/ref.cpp:26-37), so y is
// bit-identical to it.
#include <algorithm>
#include <memory>

#include "csr_plan.hpp"
#include "dia.hpp"

namespace bsp {
namespace {

constexpr int kTB = 256;

// present[d + nrows - 1] = 1 for every diagonal d holding an entry; a thread
// per row (the rows of an eligible matrix are at most kDiaMax long; longer
// ones stop at kDiaMax + 1 entries, which already rules the matrix out).
__global__ void dia_mark(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                         int64_t n, int32_t* __restrict__ present, int32_t* __restrict__ too_long) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = rowptr[r], e = rowptr[r + 1];
  if (e - s > kDiaMax) {
    *too_long = 1;
    return;
  }
  for (int j = s; j < e; ++j) present[col[j] - r + (n - 1)] = 1;
}

// Diagonal offsets in ascending order from the scanned presence flags.
__global__ void dia_list(const int32_t* __restrict__ present, const int32_t* __restrict__ pos,
                         int64_t len, int64_t n, int32_t* __restrict__ off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < len && present[i] && pos[i] < kDiaMax) off[pos[i]] = (int32_t)(i - (n - 1));
}

// Values into their diagonal slots and the row masks; a thread per row.
__global__ void dia_fill(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                         const double* __restrict__ v, int64_t n, const int32_t* __restrict__ pos,
                         double* __restrict__ val, uint32_t* __restrict__ mask) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  uint32_t m = 0;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) {
    const int k = pos[col[j] - r + (n - 1)];
    val[(int64_t)k * n + r] = v[j];
    m |= 1u << k;
  }
  mask[r] = m;
}

// y[r] = sum over the row's K diagonals, ascending, of val[k][r] * x[r + d_k];
// a thread per row, K a compile-time count (one instantiation per count, so
// the offsets arrive in SGPRs at once and nothing branches). Every value
// slot is loaded (coalesced: consecutive rows of one diagonal; absent slots
// hold 0) and x at the clamped column, all before the sum, which keeps
// s + a*b only where the row's mask bit is set: an absent slot never
// contributes, whatever x holds there.
template <int K>
__global__ __launch_bounds__(kTB) void dia_spmv_f64(DiaOffsets off,
                                                    const double* __restrict__ val,
                                                    const uint32_t* __restrict__ mask,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y, int32_t n,
                                                    int32_t ncols) {
#pragma clang fp contract(off)
  const int32_t r = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  double a[K > 0 ? K : 1], b[K > 0 ? K : 1];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    a[k] = ld_stream(&val[(int64_t)k * n + r]);
    b[k] = x[min(max(r + off.d[k], 0), ncols - 1)];
  }
  const uint32_t m = mask[r];
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t = s + a[k] * b[k];
    s = ((m >> k) & 1u) ? t : s;
  }
  y[r] = s;
}

// The same sums two rows per lane (rows 2t, 2t + 1): one 16-byte value load
// per diagonal, so a wave's load is 1 KB of one diagonal, and one 8-byte mask
// load. Needs an even row count (each diagonal's run then starts 16-byte
// aligned) and a 16-byte aligned y. Each row's arithmetic is the one-row
// kernel's, so y is the same bits. 128-lane workgroups: 27-point 64^3
// measured 11.1 us against 12.8 us for a row per lane (a read of the same
// bytes alone: 11.0 us).
constexpr int kTB2 = 128;
typedef double dbl2v __attribute__((ext_vector_type(2)));
typedef uint32_t u2v __attribute__((ext_vector_type(2)));

template <int K>
__global__ __launch_bounds__(kTB2) void dia_spmv2_f64(DiaOffsets off,
                                                      const double* __restrict__ val,
                                                      const uint32_t* __restrict__ mask,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, int32_t n,
                                                      int32_t ncols) {
#pragma clang fp contract(off)
  const int32_t r = 2 * (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  dbl2v a[K > 0 ? K : 1];
  double b0[K > 0 ? K : 1], b1[K > 0 ? K : 1];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    a[k] = ld_stream(reinterpret_cast<const dbl2v*>(&val[(int64_t)k * n + r]));
    const int32_t c = r + off.d[k];
    b0[k] = x[min(max(c, 0), ncols - 1)];
    b1[k] = x[min(max(c + 1, 0), ncols - 1)];
  }
  const u2v m = *reinterpret_cast<const u2v*>(&mask[r]);
  double s0 = 0.0, s1 = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const double t0 = s0 + a[k].x * b0[k];
    s0 = ((m.x >> k) & 1u) ? t0 : s0;
    const double t1 = s1 + a[k].y * b1[k];
    s1 = ((m.y >> k) & 1u) ? t1 : s1;
  }
  *reinterpret_cast<dbl2v*>(&y[r]) = dbl2v{s0, s1};
}

template <int K>
void launch_dia(const bsp_dia_s* h, const double* x, double* y, bool pairs, hipStream_t s) {
  if constexpr (K <= kDiaMax) {
    if (h->ndiag == K) {
      const int64_t n = h->nrows;
      if (pairs)
        dia_spmv2_f64<K><<<cdiv(n / 2, kTB2), kTB2, 0, s>>>(h->off, h->val.get(), h->mask.get(),
                                                             x, y, (int32_t)n, (int32_t)h->ncols);
      else
        dia_spmv_f64<K><<<cdiv(n, kTB), kTB, 0, s>>>(h->off, h->val.get(), h->mask.get(), x, y,
                                                     (int32_t)n, (int32_t)h->ncols);
      return;
    }
    launch_dia<K + 1>(h, x, y, pairs, s);
  }
}

}  // namespace
}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_dia_build(const bsp_csr* A, bsp_stream stream, bsp_dia* out) {
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(out, "null output handle");
  BSP_REQUIRE(A->dtype == BSP_F64 && (A->nnz == 0 || A->val), "FP64 values required");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = A->nrows;
  int flags = 0;
  BSP_TRY(csr_check(*A, s, &flags));
  if (flags & 3) return fail(BSP_ERR_INVALID, "malformed CSR (row pointers or column range)");
  if (flags & 4)
    return fail(BSP_ERR_UNSORTED, "the diagonal layout needs strictly ascending columns");
  auto h = std::make_unique<bsp_dia_s>();
  h->nrows = n;
  h->ncols = A->ncols;
  h->nnz = A->nnz;
  if (n > 0 && A->nnz > 0) {
    const int64_t len = n + A->ncols - 1;  // offsets -(n-1) .. ncols-1
    DevBuf<int32_t> present, pos, off, too_long;
    BSP_TRY(present.alloc(len + 1));
    BSP_TRY(pos.alloc(len + 1));
    BSP_TRY(off.alloc(kDiaMax));
    BSP_TRY(too_long.alloc(1));
    BSP_HIP(hipMemsetAsync(present.get(), 0, sizeof(int32_t) * (len + 1), s));
    BSP_HIP(hipMemsetAsync(too_long.get(), 0, sizeof(int32_t), s));
    dia_mark<<<cdiv(n, kTB), kTB, 0, s>>>(A->rowptr, A->colidx, n, present.get(), too_long.get());
    BSP_HIP(hipGetLastError());
    int32_t tl = 0;
    BSP_TRY(fetch(too_long.get(), &tl, s));
    if (tl) return fail(BSP_ERR_UNSUPPORTED, "a row is longer than 32 entries: too many diagonals");
    BSP_TRY(exclusive_scan_i32(present.get(), pos.get(), len + 1, s));
    int32_t k = 0;
    BSP_TRY(fetch(pos.get() + len, &k, s));
    if (k > kDiaMax) return fail(BSP_ERR_UNSUPPORTED, "more than 32 distinct diagonals");
    dia_list<<<cdiv(len, kTB), kTB, 0, s>>>(present.get(), pos.get(), len, n, off.get());
    BSP_HIP(hipGetLastError());
    h->ndiag = k;
    BSP_HIP(hipMemcpyAsync(h->off.d, off.get(), sizeof(int32_t) * k, hipMemcpyDeviceToHost, s));
    BSP_TRY(h->val.alloc((size_t)k * n));
    BSP_TRY(h->mask.alloc(n));
    BSP_HIP(hipMemsetAsync(h->val.get(), 0, sizeof(double) * (size_t)k * n, s));
    dia_fill<<<cdiv(n, kTB), kTB, 0, s>>>(A->rowptr, A->colidx, static_cast<const double*>(A->val),
                                          n, pos.get(), h->val.get(), h->mask.get());
    BSP_HIP(hipGetLastError());
    BSP_HIP(hipStreamSynchronize(s));
  } else if (n > 0) {
    BSP_TRY(h->mask.alloc(n));
    BSP_HIP(hipMemsetAsync(h->mask.get(), 0, sizeof(uint32_t) * n, s));
    BSP_HIP(hipStreamSynchronize(s));
  }
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_dia_info(bsp_dia h, int32_t* ndiag, int32_t* offsets, int64_t* bytes) {
  BSP_REQUIRE(h && ndiag, "null argument");
  *ndiag = h->ndiag;
  if (offsets) std::copy(h->off.d, h->off.d + h->ndiag, offsets);
  if (bytes) *bytes = 8 * (int64_t)h->ndiag * h->nrows + 4 * h->nrows;
  return BSP_OK;
}

bsp_status bsp_dia_spmv(bsp_dia h, const double* x, double* y, bsp_stream stream) {
  BSP_REQUIRE(h, "null handle");
  BSP_REQUIRE((x || h->ncols == 0) && (y || h->nrows == 0), "null vector");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = h->nrows;
  if (n == 0) return BSP_OK;
  BSP_REQUIRE(cdiv(n, kTB) * kTB < kDispatchLimit, "too many rows for one launch");
  launch_dia<0>(h, x, y, n % 2 == 0 && ((uintptr_t)y & 15) == 0, s);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_dia_destroy(bsp_dia h) {
  delete h;
  return BSP_OK;
}

}  // extern "C"

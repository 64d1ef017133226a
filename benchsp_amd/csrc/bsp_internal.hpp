// Internal helpers shared by the library's translation units.
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/bsp.h"

namespace bsp {

// Host CSR produced by the generators (int64 row offsets, int32 columns).
struct HostCsr {
  int64_t nrows = 0, ncols = 0;
  std::vector<int64_t> rowptr;
  std::vector<int32_t> col;
  std::vector<double> val;
};

// Records `msg` as the calling thread's last error and returns `s`.
bsp_status fail(bsp_status s, const std::string& msg);

}  // namespace bsp

#define BSP_REQUIRE(cond, msg)                                     \
  do {                                                             \
    if (!(cond)) return ::bsp::fail(BSP_ERR_INVALID, (msg));       \
  } while (0)

namespace bsp {

// Largest row count, column count and entry count the int32 kernels accept:
// 2^31 - 2^20, so that an index plus any in-kernel offset (a window of
// entries, a lane or a workgroup stride, all below 2^20) still fits int32.
constexpr int64_t kMaxIndex = (int64_t(1) << 31) - (int64_t(1) << 20);

inline bool sizes_fit(int64_t nrows, int64_t ncols, int64_t nnz) {
  return nrows <= kMaxIndex && ncols <= kMaxIndex && nnz <= kMaxIndex;
}

// Sizes and index pointers of a CSR descriptor (values not checked): sizes
// must be non-negative and at most kMaxIndex (BSP_ERR_TOO_LARGE otherwise).
inline bsp_status check_csr_pattern(const bsp_csr* A) {
  BSP_REQUIRE(A, "null CSR descriptor");
  BSP_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, "negative CSR size");
  if (!sizes_fit(A->nrows, A->ncols, A->nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(A->nrows == 0 || A->rowptr, "null rowptr");
  BSP_REQUIRE(A->nnz == 0 || A->colidx, "null colidx");
  return BSP_OK;
}

// Full O(rows + nnz) check of a caller's host CSR before anything is
// uploaded, so that a malformed pattern is an error status instead of an
// out-of-bounds access on the GPU: sizes non-negative and within kMaxIndex,
// rowptr[0] == 0, rowptr non-decreasing, rowptr[rows] == nnz, and every
// column index in [0, cols). Works for int32 and int64 index arrays (an
// int64 value that passes fits int32).
template <class RP, class CI>
bsp_status check_host_csr(const RP* rp, const CI* ci, int64_t rows, int64_t cols, int64_t nnz) {
  BSP_REQUIRE(rows >= 0 && cols >= 0 && nnz >= 0, "negative CSR size");
  if (!sizes_fit(rows, cols, nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(rp, "null rowptr");
  BSP_REQUIRE(nnz == 0 || ci, "null column index array");
  BSP_REQUIRE(rp[0] == 0, "rowptr[0] must be 0");
  for (int64_t i = 0; i < rows; ++i)
    if (rp[i + 1] < rp[i])
      return fail(BSP_ERR_INVALID, "rowptr decreases at row " + std::to_string(i));
  BSP_REQUIRE((int64_t)rp[rows] == nnz, "rowptr[rows] must equal nnz");
  for (int64_t j = 0; j < nnz; ++j)
    if (ci[j] < 0 || (int64_t)ci[j] >= cols)
      return fail(BSP_ERR_INVALID, "column index out of range at entry " + std::to_string(j));
  return BSP_OK;
}

}  // namespace bsp

// C++ drop-in for the reference's public headers: the same free-function
// names and signatures as include/csrspmv.h, amgT.h, bcsr.h, dasp.h, spmv.h,
// smat.h, cusp.h, ref.h, and the types of include/matrixFormat.hpp (CSRFormat,
// BCSRFormat, dbsrMat), include/dbsr.h (csrTodbsr, balanceDbsr, csrRecorderSelf,
// csrTodbsrWithRecorder) and utils/utils.hpp (csrToBcsr), each an
// inline forward to the C-ABI in bsp.h. A driver written against the
// reference headers includes this header instead and links libbsp.so.
//
// Differences of behaviour (the reference's are defects):
//  - amgT_spmv_fp64* return y = A*x; the reference accumulates into hY.
//  - MySpmvWithPrefp64 keeps FP64 output; the reference truncates to float.
//  - smatSpmmfp16 writes C (row x n, FP32) to outvalues; the reference does not.
//  - every function reports failure (non-zero return, or std::runtime_error
//    for the void builders) instead of exiting.
// MySpmvWithPre (include/spmv.h:1) is declared by the reference but never
// defined; it is not provided.
#pragma once

#include <hip/hip_fp16.h>

#include <cstddef>
#include <cstdint>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "bsp.h"

// include/common.h:7-8
#ifndef WARP_CAPACITY
#define WARP_CAPACITY 64
#endif
#ifndef WARP_NUM_SPMV
#define WARP_NUM_SPMV 4
#endif

// include/matrixFormat.hpp:5-13
template <typename T, typename Y>
class CSRFormat {
 public:
  Y* rowPtr;
  Y* colIdx;
  T* values;
  Y row, col;
  Y nnz;
};

// include/matrixFormat.hpp:15-30
template <typename T, typename Y>
class BCSRFormat {
 public:
  Y* bcsrRowPtr;
  Y* bcsrcolIdx;
  Y* relativeBlockIndexMapping;
  T* values;
  Y row, col, blockNum;
  int TILE_M;
  int TILE_N;
  int TILE_K;
  Y nnz;
  Y colRegions, rowRegions, nonzeroBlocks, elemCount;
  Y block_row;
  Y block_col;
};

// include/matrixFormat.hpp:86-106
class dbsrMat {
 public:
  int row;
  int col;
  int nnz;
  int blc_row;
  int blc_col;
  int blc_num;
  double stand;
  double avg_nnz;
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
};

namespace bsp_compat_detail {
#define BSP_SAME_FIELD(A, B, f) static_assert(offsetof(A, f) == offsetof(B, f), #f)
using Csr64 = CSRFormat<double, int32_t>;
static_assert(sizeof(Csr64) == sizeof(bsp_CSRFormat_f64), "CSRFormat layout");
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, rowPtr);
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, colIdx);
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, values);
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, row);
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, col);
BSP_SAME_FIELD(Csr64, bsp_CSRFormat_f64, nnz);
static_assert(sizeof(dbsrMat) == sizeof(bsp_dbsrMat), "dbsrMat layout");
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, blc_num);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, stand);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, avg_nnz);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, blcPtr);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, blcMap);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, blocknnz);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, warpnum);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, rowPtrbyWarp);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, recorderMap);
BSP_SAME_FIELD(dbsrMat, bsp_dbsrMat, prefix);
#undef BSP_SAME_FIELD
static_assert(sizeof(half) == sizeof(uint16_t), "half is binary16");

inline bsp_dbsrMat* c(dbsrMat* m) { return reinterpret_cast<bsp_dbsrMat*>(m); }
}  // namespace bsp_compat_detail

// include/csrspmv.h:2-3
inline int spmv_fp64_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                             int32_t nnz, double* values, double* x, double* outvalues,
                             int repeat, double& avg_time) {
  return bsp_spmv_fp64_warpper(rowPtr, colIdx, rows, cols, nnz, values, x, outvalues, repeat,
                               &avg_time);
}
inline int spmv_fp64_balance_warpper(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                                     int32_t nnz, double* values, double* x, double* outvalues,
                                     int repeat, double& avg_time) {
  return bsp_spmv_fp64_balance_warpper(rowPtr, colIdx, rows, cols, nnz, values, x, outvalues,
                                       repeat, &avg_time);
}

// include/amgT.h:3-6
inline int amgT_spmv_fp64(int32_t* hA_csrOffsets, int32_t* hA_columns, double* hA_values,
                          double* hX, double* hY, int32_t A_num_rows, int32_t A_num_cols,
                          int32_t A_nnz, int repeat, double& avg_time) {
  return bsp_amgT_spmv_fp64(hA_csrOffsets, hA_columns, hA_values, hX, hY, A_num_rows, A_num_cols,
                            A_nnz, repeat, &avg_time);
}
inline int amgT_spmv_fp64_dbsr(dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                               int32_t repeat, double& avg_time) {
  return bsp_amgT_spmv_fp64_dbsr(bsp_compat_detail::c(dbsr), hX, hY, warmup, repeat, &avg_time);
}
inline int amgT_spmv_fp64_dbsr_balance(dbsrMat* dbsr, double* hX, double* hY, int32_t warmup,
                                       int32_t repeat, double& avg_time) {
  return bsp_amgT_spmv_fp64_dbsr_balance(bsp_compat_detail::c(dbsr), hX, hY, warmup, repeat,
                                         &avg_time);
}

// include/bcsr.h:3
inline void bcsr_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values32,
                           double* hX32, double* hY32, int64_t A_num_rows, int64_t A_num_cols,
                           int64_t A_nnz, int repeat) {
  bsp_bcsr_spmv_fp64(hA_csrOffsets, hA_columns, hA_values32, hX32, hY32, A_num_rows, A_num_cols,
                     A_nnz, repeat);
}

// include/dasp.h:1-2
inline void spmv_all(char* filename, double* csrValA, int* csrRowPtrA, int* csrColIdxA,
                     double* X_val, double* Y_val, int* order_rid, int rowA, int colA, int nnzA,
                     int NUM, double threshold, int block_longest, double& avg_time) {
  bsp_spmv_all(filename, csrValA, csrRowPtrA, csrColIdxA, X_val, Y_val, order_rid, rowA, colA,
               nnzA, NUM, threshold, block_longest, &avg_time);
}

// include/spmv.h:2-3
inline void MySpmvWithPrefp16(int32_t* rowPtr, int32_t* colIdx, int32_t rows, int32_t cols,
                              int32_t nnz, float* values, float* x, float* outvalues) {
  bsp_MySpmvWithPrefp16(rowPtr, colIdx, rows, cols, nnz, values, x, outvalues);
}
inline void MySpmvWithPrefp64(int64_t* rowPtr, int64_t* colIdx, int64_t rows, int64_t cols,
                              int64_t nnz, double* values, double* x, double* outvalues) {
  bsp_MySpmvWithPrefp64(rowPtr, colIdx, rows, cols, nnz, values, x, outvalues);
}

// include/smat.h:1
inline void smatSpmmfp16(int* rowPtr, int* colIdx, int row, int col, int n, int nnz,
                         float* values, float* outvalues) {
  bsp_smatSpmmfp16(rowPtr, colIdx, row, col, n, nnz, values, outvalues);
}

// include/cusp.h:3-6 (rocSPARSE underneath)
inline int cusparse_spmv(int* A_csrOffsets, int64_t* A_columns, float* A_values, float* x,
                         float* y, int64_t row, int64_t col, int64_t nnz) {
  return bsp_cusparse_spmv(A_csrOffsets, A_columns, A_values, x, y, row, col, nnz);
}
inline int cusparse_spmv_fp64(int64_t* hA_csrOffsets, int64_t* hA_columns, double* hA_values,
                              double* hX, double* hY, int64_t A_num_rows, int64_t A_num_cols,
                              int64_t A_nnz, int repeat, double& avg_time) {
  return bsp_cusparse_spmv_fp64(hA_csrOffsets, hA_columns, hA_values, hX, hY, A_num_rows,
                                A_num_cols, A_nnz, repeat, &avg_time);
}
inline int cusparse_spmv_fp16(int* hA_csrOffsets, int32_t* hA_columns, half* hA_values, half* hX,
                              half* hY, int32_t A_num_rows, int32_t A_num_cols, int32_t A_nnz) {
  return bsp_cusparse_spmv_fp16(hA_csrOffsets, hA_columns,
                                reinterpret_cast<const uint16_t*>(hA_values),
                                reinterpret_cast<const uint16_t*>(hX),
                                reinterpret_cast<uint16_t*>(hY), A_num_rows, A_num_cols, A_nnz);
}
inline void cusparse_spmv_fp16_warpper(int* hA_csrOffsets, int32_t* hA_columns, float* hA_values32,
                                       float* hX32, float* hY32, int32_t A_num_rows,
                                       int32_t A_num_cols, int32_t A_nnz) {
  bsp_cusparse_spmv_fp16_warpper(hA_csrOffsets, hA_columns, hA_values32, hX32, hY32, A_num_rows,
                                 A_num_cols, A_nnz);
}

// include/ref.h:2-4 (GPU kernels bit-identical to the CPU loops)
inline void ref_spmv_fp16(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                          float* y, int32_t row, int32_t col) {
  bsp_ref_spmv_fp16(A_csrOffsets, A_columns, A_values, x, y, row, col);
}
inline void ref_spmv_fp32(int* A_csrOffsets, int32_t* A_columns, float* A_values, float* x,
                          float* y, int32_t row, int32_t col) {
  bsp_ref_spmv_fp32(A_csrOffsets, A_columns, A_values, x, y, row, col);
}
inline void ref_spmv_fp64(int64_t* A_csrOffsets, int64_t* A_columns, double* A_values, double* x,
                          double* y, int64_t row, int64_t col, double& avg_time) {
  bsp_ref_spmv_fp64(A_csrOffsets, A_columns, A_values, x, y, row, col, &avg_time);
}

// include/dbsr.h:10-93: arrays are malloc'ed; bsp_dbsrMat_free releases them
inline void csrTodbsr(CSRFormat<double, int32_t>* csr, dbsrMat* bsr) {
  if (bsp_csrTodbsr(reinterpret_cast<const bsp_CSRFormat_f64*>(csr), bsp_compat_detail::c(bsr)) != 0)
    throw std::runtime_error(std::string("csrTodbsr: ") + bsp_last_error());
}

// include/dbsr.h:94-169
inline void balanceDbsr(dbsrMat* dbsr) {
  if (bsp_balanceDbsr(bsp_compat_detail::c(dbsr)) != 0)
    throw std::runtime_error(std::string("balanceDbsr: ") + bsp_last_error());
}

// include/dbsr.h:170-206: recsr's arrays are malloc'ed (bsp_free)
inline void csrRecorderSelf(CSRFormat<double, int32_t>* csr, CSRFormat<double, int32_t>* recsr) {
  if (bsp_csrRecorderSelf(reinterpret_cast<const bsp_CSRFormat_f64*>(csr),
                          reinterpret_cast<bsp_CSRFormat_f64*>(recsr)) != 0)
    throw std::runtime_error(std::string("csrRecorderSelf: ") + bsp_last_error());
}

// utils/utils.hpp:9-124. Reads TILE_M and TILE_K from bcsr; fills blockNum,
// colRegions, rowRegions, row, col, bcsrRowPtr, bcsrcolIdx, values,
// relativeBlockIndexMapping (blockNum entries, -1 for an empty block, as the
// reference's dense block map), nonzeroBlocks and elemCount. Every array is
// malloc'ed. Values pass through FP64 unchanged (half, float and double are
// exact there).
template <typename T, typename Y>
void csrToBcsr(CSRFormat<T, Y>* csr, BCSRFormat<T, Y>* bcsr) {
  const int64_t rows = csr->row, cols = csr->col, nnz = rows > 0 ? (int64_t)csr->rowPtr[rows] : 0;
  const int tm = bcsr->TILE_M, tk = bcsr->TILE_K;
  if (tm < 1 || tk < 1) throw std::runtime_error("csrToBcsr: TILE_M and TILE_K must be >= 1");
  auto to_f64 = [](T v) { return static_cast<double>(static_cast<float>(v)); };
  std::vector<int64_t> rp(rows + 1), ci(nnz);
  std::vector<double> v(nnz);
  for (int64_t i = 0; i <= rows; ++i) rp[i] = csr->rowPtr[i];
  for (int64_t j = 0; j < nnz; ++j) {
    ci[j] = csr->colIdx[j];
    if constexpr (std::is_same<T, double>::value)
      v[j] = csr->values[j];
    else
      v[j] = to_f64(csr->values[j]);
  }
  int64_t nb = 0, *brp = nullptr, *bci = nullptr;
  double* bv = nullptr;
  if (bsp_csrToBcsr(rp.data(), ci.data(), v.data(), rows, cols, tm, tk, &nb, &brp, &bci, &bv) != 0)
    throw std::runtime_error(std::string("csrToBcsr: ") + bsp_last_error());
  const int64_t rreg = (rows + tm - 1) / tm, creg = (cols + tk - 1) / tk;
  bcsr->row = (Y)rows;
  bcsr->col = (Y)cols;
  bcsr->rowRegions = (Y)rreg;
  bcsr->colRegions = (Y)creg;
  bcsr->blockNum = (Y)(rreg * creg);
  bcsr->nonzeroBlocks = (Y)nb;
  bcsr->elemCount = (Y)(nb * tm * tk);
  bcsr->bcsrRowPtr = (Y*)std::malloc(sizeof(Y) * (rreg + 1));
  bcsr->bcsrcolIdx = (Y*)std::malloc(sizeof(Y) * (nb ? nb : 1));
  bcsr->values = (T*)std::malloc(sizeof(T) * (nb ? nb * tm * tk : 1));
  bcsr->relativeBlockIndexMapping = (Y*)std::malloc(sizeof(Y) * (rreg * creg ? rreg * creg : 1));
  if (!bcsr->bcsrRowPtr || !bcsr->bcsrcolIdx || !bcsr->values || !bcsr->relativeBlockIndexMapping)
    throw std::runtime_error("csrToBcsr: host allocation failed");
  for (int64_t i = 0; i <= rreg; ++i) bcsr->bcsrRowPtr[i] = (Y)brp[i];
  for (int64_t i = 0; i < rreg * creg; ++i) bcsr->relativeBlockIndexMapping[i] = Y(-1);
  for (int64_t br = 0; br < rreg; ++br)
    for (int64_t k = brp[br]; k < brp[br + 1]; ++k) {
      bcsr->bcsrcolIdx[k] = (Y)bci[k];
      bcsr->relativeBlockIndexMapping[br * creg + bci[k]] = (Y)k;
    }
  for (int64_t i = 0; i < nb * tm * tk; ++i) {
    if constexpr (std::is_same<T, double>::value)
      bcsr->values[i] = bv[i];
    else
      bcsr->values[i] = T(static_cast<float>(bv[i]));
  }
  bsp_free(brp);
  bsp_free(bci);
  bsp_free(bv);
}

// include/dbsr.h:208-253: arrays are malloc'ed; bsp_dbsrMat_free releases them
inline void csrTodbsrWithRecorder(CSRFormat<double, int32_t>* csr, dbsrMat* bsr) {
  if (bsp_csrTodbsrWithRecorder(reinterpret_cast<const bsp_CSRFormat_f64*>(csr),
                                bsp_compat_detail::c(bsr)) != 0)
    throw std::runtime_error(std::string("csrTodbsrWithRecorder: ") + bsp_last_error());
}

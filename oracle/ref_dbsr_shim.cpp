// Exposes the reference's own header-only DBSR builder (include/dbsr.h,
// compiled from the reference tree, never copied) through a C ABI so tests
// can compare the oracle's restatement with it on small matrices.
// TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <cstdio>
#include "dbsr.h"

extern "C" {
struct ref_dbsr_out {
  int32_t brows, nblocks, warpnum, blocknnz;
  int32_t *blcPtr, *colIdx, *recorderMap, *prefix;
  double* blcVal;
  unsigned short* blcMap;
};

// Runs csrTodbsrWithRecorder (dbsr.h:208-253) on a CSR given as int32 arrays.
// The arrays the reference allocates are handed back as they are.
int ref_csr_to_dbsr(int32_t row, int32_t col, int32_t nnz, int32_t* rowPtr, int32_t* colIdx,
                    double* values, ref_dbsr_out* out) {
  CSRFormat<double, int32_t> csr;
  csr.row = row;
  csr.col = col;
  csr.nnz = nnz;
  csr.rowPtr = rowPtr;
  csr.colIdx = colIdx;
  csr.values = values;
  dbsrMat d;
  // The reference prints statistics to stdout; keep the test output clean.
  fflush(stdout);
  FILE* saved = stdout;
  FILE* devnull = fopen("/dev/null", "w");
  if (devnull) stdout = devnull;
  csrTodbsrWithRecorder(&csr, &d);
  if (devnull) { fflush(stdout); stdout = saved; fclose(devnull); }
  out->brows = d.blc_row;
  out->blocknnz = d.blocknnz;
  out->nblocks = d.blocknnz / 16;
  out->warpnum = d.warpnum;
  out->blcPtr = d.blcPtr;
  out->colIdx = d.colIdx;
  out->blcVal = d.blcVal;
  out->blcMap = d.blcMap;
  out->recorderMap = d.recorderMap;
  out->prefix = d.prefix;
  return 0;
}
}

extern "C" {
// Runs csrTodbsr (dbsr.h:10-93) and then balanceDbsr (:94-169) on the CSR as
// given, without the row sort of csrTodbsrWithRecorder.
int ref_csr_to_dbsr_plain(int32_t row, int32_t col, int32_t nnz, int32_t* rowPtr,
                          int32_t* colIdx, double* values, ref_dbsr_out* out) {
  CSRFormat<double, int32_t> csr;
  csr.row = row;
  csr.col = col;
  csr.nnz = nnz;
  csr.rowPtr = rowPtr;
  csr.colIdx = colIdx;
  csr.values = values;
  dbsrMat d;
  d.recorderMap = nullptr;
  fflush(stdout);
  FILE* saved = stdout;
  FILE* devnull = fopen("/dev/null", "w");
  if (devnull) stdout = devnull;
  csrTodbsr(&csr, &d);
  balanceDbsr(&d);
  if (devnull) { fflush(stdout); stdout = saved; fclose(devnull); }
  out->brows = d.blc_row;
  out->blocknnz = d.blocknnz;
  out->nblocks = d.blocknnz / 16;
  out->warpnum = d.warpnum;
  out->blcPtr = d.blcPtr;
  out->colIdx = d.colIdx;
  out->blcVal = d.blcVal;
  out->blcMap = d.blcMap;
  out->recorderMap = nullptr;
  out->prefix = d.prefix;
  return 0;
}

// Runs csrRecorderSelf (dbsr.h:170-206): the CSR with rows sorted by
// (length, index); arrays handed back as the reference allocates them.
int ref_csr_recorder_self(int32_t row, int32_t col, int32_t nnz, int32_t* rowPtr,
                          int32_t* colIdx, double* values, int32_t** outRowPtr,
                          int32_t** outColIdx, double** outValues) {
  CSRFormat<double, int32_t> csr, rec;
  csr.row = row;
  csr.col = col;
  csr.nnz = nnz;
  csr.rowPtr = rowPtr;
  csr.colIdx = colIdx;
  csr.values = values;
  csrRecorderSelf(&csr, &rec);
  *outRowPtr = rec.rowPtr;
  *outColIdx = rec.colIdx;
  *outValues = rec.values;
  return 0;
}
}

// Per-matrix CSR SpMV schedule (what bsp_spmv_plan points at).
#pragma once

#include <vector>

#include "bsp_hip.hpp"

namespace bsp {
// One column segment of the adaptive layout: a CSR over all rows holding the
// entries whose column falls in [c0, c0 + width), with its row-block cut.
struct CsrSegment {
  const int32_t* rowptr = nullptr;
  const int32_t* col = nullptr;
  const double* val = nullptr;
  int64_t nnz = 0;
  DevBuf<int32_t> own_rowptr;
  DevBuf<int32_t> blk;  // nblk+1 {first row, first entry} pairs
  int64_t nblk = 0;
  int64_t nlong = 0;
};
}  // namespace bsp

struct bsp_spmv_plan_s {
  bsp_csr A{};       // borrowed device pointers
  int method = 0;    // bsp_csr_method
  bsp_csr_options opt{};
  int64_t nblk = 0;  // balanced: chunks; split: short rows; adaptive: total blocks
  int64_t nlong = 0; // split: long rows; adaptive: long-row blocks
  bsp::DevBuf<int32_t> a, b;  // method-specific index arrays
  bsp::DevBuf<double> work;   // balanced: chunk partials
  // adaptive
  std::vector<bsp::CsrSegment> seg;
  bsp::DevBuf<int32_t> seg_col;  // entries regrouped by column segment
  bsp::DevBuf<double> seg_val;
  int64_t seg_width = 0;         // 0: one segment (the input arrays as they are)
  int chunk = 1024;              // adaptive block window (rows + entries)
  bool xcd_mode = false;         // all segments in one XCD-pinned launch
  int64_t xcd_grid = 0;
  bsp::DevBuf<double> part;      // per-segment partial y (xcd_mode)
  bsp::DevBuf<char> segtab;      // device table of segment descriptors (xcd_mode)
  bool sorted = false;           // sorted-block segments (seg_mode 2)
  bsp::DevBuf<uint16_t> slot;    // sorted blocks: each entry's row-order position in its block
  bsp::DevBuf<int16_t> col16;    // 16-bit column offsets from each block's first row
  bool narrow = false;           // unsegmented plan reading col16
};

namespace bsp {
using CsrPlan = bsp_spmv_plan_s;
bsp_status csr_plan_build(CsrPlan* p, hipStream_t s);
bsp_status csr_plan_run(const CsrPlan* p, const double* x, double* y, hipStream_t s);
bsp_status csr_plan_run_f16(const CsrPlan* p, const uint16_t* x, float* y, hipStream_t s);
// GPU validation of a CSR's pattern: *flags bit 1 = bad row pointers, bit 2 =
// column out of range, bit 4 = columns not strictly ascending (synchronises).
bsp_status csr_check(const bsp_csr& A, hipStream_t s, int* flags);
bsp_status csr_plan_probe(const CsrPlan* p, int with_values, const double* x, double* sink,
                          int64_t sink_len, int64_t* grid, hipStream_t s);
bsp_status gather_probe(const double* x, int64_t n, int nslices, int64_t gathers, double* sink,
                        int64_t sink_len, int64_t* grid, hipStream_t s);
bsp_status line_probe(const double* x, int64_t n, int nslices, int64_t nbytes, double* sink,
                      int64_t sink_len, int64_t* grid, hipStream_t s);
}  // namespace bsp

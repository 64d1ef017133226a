// Per-matrix CSR SpMV schedule (what bsp_spmv_plan points at).
#pragma once

#include "bsp_hip.hpp"

struct bsp_spmv_plan_s {
  bsp_csr A{};       // borrowed device pointers
  int method = 0;    // bsp_csr_method
  int64_t nblk = 0;  // adaptive: blocks; balanced: chunks; split: short rows
  int64_t nlong = 0; // split: long rows; adaptive: long-row blocks (info only)
  bsp::DevBuf<int32_t> a, b;  // method-specific index arrays
  bsp::DevBuf<double> work;   // balanced: chunk partials
};

namespace bsp {
using CsrPlan = bsp_spmv_plan_s;
bsp_status csr_plan_build(CsrPlan* p, hipStream_t s);
bsp_status csr_plan_run(const CsrPlan* p, const double* x, double* y, hipStream_t s);
}  // namespace bsp

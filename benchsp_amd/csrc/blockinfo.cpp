// Block statistics of a CSR matrix (bench/blockinfo.cpp:96-119): the number
// of b x b blocks holding at least one entry, and the number of blocks when
// each row's entries are packed to the left (entry k of a row goes to block
// column k / b), i.e. per group of b rows the longest row's ceil(len / b).
// Both on the GPU from the pattern alone; the first with the BSR builder's
// key sort, no values placed.
#include <hip/hip_runtime.h>

#include "bsp_hip.hpp"

using namespace bsp;

namespace {

__global__ void packed_blocks(const int32_t* __restrict__ rowptr, int64_t nrows, int b,
                              int64_t* __restrict__ out) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t ng = (nrows + b - 1) / b;
  if (g >= ng) return;
  int64_t m = 0;
  const int64_t r1 = min(nrows, (g + 1) * b);
  for (int64_t r = g * b; r < r1; ++r) {
    const int64_t len = rowptr[r + 1] - rowptr[r];
    m = max(m, (len + b - 1) / b);
  }
  out[g] = m;
}

}  // namespace

extern "C" bsp_status bsp_block_info(const bsp_csr* A, int block_len, bsp_stream stream,
                                     bsp_blockinfo* out) {
  BSP_REQUIRE(A && out, "null argument");
  BSP_REQUIRE(block_len >= 1 && block_len <= 64, "block_len must be in [1, 64]");
  BSP_TRY(check_csr_pattern(A));
  hipStream_t s = (hipStream_t)stream;
  bsp_blockinfo r{};
  r.nnz = A->nnz;
  if (A->nrows > 0 && A->nnz > 0) {
    BSP_TRY(bsr_count_blocks(*A, block_len, block_len, s, &r.blocks));
    const int64_t ng = (A->nrows + block_len - 1) / block_len;
    DevBuf<int64_t> per, sum;
    BSP_TRY(per.alloc(ng));
    BSP_TRY(sum.alloc(1));
    packed_blocks<<<cdiv(ng, 256), 256, 0, s>>>(A->rowptr, A->nrows, block_len, per.get());
    BSP_HIP(hipGetLastError());
    BSP_TRY(reduce_sum_i64(per.get(), sum.get(), ng, s));
    BSP_TRY(fetch(sum.get(), &r.blocks_packed, s));
  }
  r.avg = r.blocks ? (double)r.nnz / (double)r.blocks : 0.0;
  r.avg_packed = r.blocks_packed ? (double)r.nnz / (double)r.blocks_packed : 0.0;
  *out = r;
  return BSP_OK;
}

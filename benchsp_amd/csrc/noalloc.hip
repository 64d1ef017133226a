// FP64 CSR SpMV that allocates nothing: the adaptive schedule (2048-unit
// merge-path windows) derived on every call into a fixed device array, for
// callers that cannot allocate device memory (code under graph capture,
// tight serving loops).
//
// The adaptive plan (csr_spmv.hip) cuts rows into blocks: rows whose
// merge-path coordinate u(r) = rowptr[r] + r falls in the same window of
// CHUNK units share a block, and a row longer than CHUNK/2 (and the row after
// it) starts a new one. Here the only schedule state is the first row of
// every window, ws[w] = min{r : u(r) >= w*CHUNK}, written by one thread-per-
// row pass into `scratch`; one workgroup per window then walks that window's
// blocks in order. A window holds at most two long rows (each spans more than
// CHUNK/2 + 1 units of u), so it has at most five blocks. Every block is run
// by the same block routine as the plan's kernel, so y is bit-identical to
// an unsegmented adaptive plan with a 2048-unit window.
#include <mutex>

#include "bsp_hip.hpp"
#include "csr_adaptive.hpp"

// Windows the scratch array holds: n + nnz <= 2^31 (4 MB per GPU).
constexpr int64_t kNaMaxWindows = int64_t(1) << 20;

// Per-GPU window starts; each GPU the library runs on holds its own copy
// until the process ends. Host code uses the address from
// hipGetSymbolAddress (na_scratch_ptr), never this name.
static __device__ __align__(256) int32_t scratch[kNaMaxWindows + 1];

namespace bsp {
namespace {

constexpr int kNaChunk = 2048;
constexpr int kMaxDevices = 64;
std::mutex g_mu;
int32_t* g_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device (with that device current).
bsp_status na_scratch_ptr(int32_t** out) {
  int dev = 0;
  BSP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(BSP_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[dev]) {
    void* p = nullptr;
    BSP_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(BSP_ERR_HIP, "hipGetSymbolAddress returned a null scratch address");
    g_ptr[dev] = static_cast<int32_t*>(p);
  }
  *out = g_ptr[dev];
  return BSP_OK;
}

// ws[w] = first row r with u(r) >= w*CHUNK, for w = 0..nwin (ws[nwin] = n).
// Thread r writes the windows in (k(r-1), k(r)], k(r) = u(r)/CHUNK; thread n
// writes (k(n-1), nwin]. Each entry is written exactly once.
__global__ void na_window_starts(const int32_t* __restrict__ rowptr, int64_t n, int64_t nwin,
                                 int32_t* __restrict__ ws) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  const int64_t kp = r == 0 ? -1 : ((int64_t)rowptr[r - 1] + (r - 1)) / kNaChunk;
  const int64_t kr = r == n ? nwin : ((int64_t)rowptr[r] + r) / kNaChunk;
  for (int64_t w = kp + 1; w <= kr; ++w) ws[w] = (int32_t)r;
}

__global__ __launch_bounds__(kWG) void na_csr_windows_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
    const int32_t* __restrict__ ws, int64_t nwin) {
  constexpr int LONG = Ad<kNaChunk>::LONG;
  __shared__ int nlong;
  __shared__ int longs[2];
  const int64_t w = xcd_remap(blockIdx.x, nwin);
  const int r0 = ws[w], r1 = ws[w + 1];
  if (r0 >= r1) return;  // uniform: a window where no row starts
  if (threadIdx.x == 0) nlong = 0;
  __syncthreads();
  for (int r = r0 + (int)threadIdx.x; r < r1; r += kWG)
    if (rowptr[r + 1] - rowptr[r] > LONG) {
      const int k = atomicAdd(&nlong, 1);
      if (k < 2) longs[k] = r;
    }
  __syncthreads();
  const int nl = min(nlong, 2);
  int l0 = nl > 0 ? longs[0] : r1, l1 = nl > 1 ? longs[1] : r1;
  if (l1 < l0) {
    const int t = l0;
    l0 = l1;
    l1 = t;
  }
  // blocks in row order: [r0, l0), {l0}, [l0 + 1, l1), {l1}, [l1 + 1, r1)
  int cur = r0;
  for (int i = 0; i <= nl; ++i) {
    const int lr = i == 0 ? l0 : i == 1 ? l1 : r1;
    const int end = i < nl ? lr : r1;
    if (cur < end) {
      adaptive_rows<kNaChunk, false, false>(rowptr, col, val, x, y, cur, rowptr[cur], end,
                                            rowptr[end]);
      __syncthreads();  // the block routine's LDS is reused by the next block
    }
    if (i < nl) {
      adaptive_rows<kNaChunk, false, false>(rowptr, col, val, x, y, lr, rowptr[lr], lr + 1,
                                            rowptr[lr + 1]);
      __syncthreads();
      cur = lr + 1;
    }
  }
}

}  // namespace
}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_csr_spmv_noalloc(const bsp_csr* A, const double* x, double* y, bsp_stream s) {
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(A->dtype == BSP_F64, "FP64 values required");
  BSP_REQUIRE(A->nnz == 0 || A->val, "null values");
  BSP_REQUIRE((A->ncols == 0 || x) && (A->nrows == 0 || y), "null x/y");
  const int64_t n = A->nrows;
  const int64_t nwin = cdiv(n + A->nnz, kNaChunk);
  if (nwin > kNaMaxWindows)
    return fail(BSP_ERR_TOO_LARGE,
                "bsp_csr_spmv_noalloc: rows + entries above 2^31 (its fixed 4 MB schedule array "
                "holds 2^20 windows of 2048); build a plan with bsp_csr_plan_create instead");
  if (n == 0) return BSP_OK;
  int32_t* ws = nullptr;
  BSP_TRY(na_scratch_ptr(&ws));
  const hipStream_t st = (hipStream_t)s;
  na_window_starts<<<cdiv(n + 1, kWG), kWG, 0, st>>>(A->rowptr, n, nwin, ws);
  na_csr_windows_f64<<<nwin, kWG, 0, st>>>(A->rowptr, A->colidx,
                                           static_cast<const double*>(A->val), x, y, ws, nwin);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

}  // extern "C"

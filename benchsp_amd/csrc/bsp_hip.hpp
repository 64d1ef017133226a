// HIP-side helpers: error plumbing, owned device buffers, wave64 reductions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "bsp_internal.hpp"

#define BSP_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      return ::bsp::fail(BSP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define BSP_TRY(call)                      \
  do {                                     \
    bsp_status s_ = (call);                \
    if (s_ != BSP_OK) return s_;           \
  } while (0)

namespace bsp {

constexpr int kWave = 64;

// Owned device allocation (hipMalloc / hipFree).
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { reset(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  bsp_status alloc(size_t count) {
    reset();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      return fail(BSP_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(count * sizeof(T)) +
                                     " B): " + hipGetErrorString(e));
    }
    n = count;
    return BSP_OK;
  }
  T* get() const { return p; }
};

inline int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }

// One dispatch holds fewer than 2^32 work-items. A kernel that gives a wave
// or a workgroup to each item (a row, a block row, a block) takes the index
// of its first item, and the host issues it in slices of at most 2^31
// work-items: `launch(first, count)` per slice, one slice at every size that
// fits (a wave per row stays one dispatch up to 2^25 rows).
constexpr int64_t kSliceThreads = int64_t(1) << 31;
// Work-items one dispatch may hold (a launch issued whole must stay below).
constexpr int64_t kDispatchLimit = int64_t(1) << 32;
template <class Launch>
inline void in_slices(int64_t items, int64_t threads_per_item, Launch&& launch) {
  const int64_t per = kSliceThreads / threads_per_item;
  for (int64_t first = 0; first < items; first += per)
    launch(first, items - first < per ? items - first : per);
}

// Device-wide primitives (prims.hip): stream-ordered, temporary storage
// allocated internally.
bsp_status exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, hipStream_t s);
bsp_status exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);
// Stable sort of (key, value) pairs by key, ascending or descending.
bsp_status sort_pairs_i32(const int32_t* keys_in, int32_t* keys_out, const int32_t* vals_in,
                          int32_t* vals_out, int64_t n, bool descending, int end_bit,
                          hipStream_t s);
// Stable ascending sort of (int64 key, int32 value) pairs on bits [0, end_bit).
bsp_status sort_pairs_i64key(const int64_t* keys_in, int64_t* keys_out, const int32_t* vals_in,
                             int32_t* vals_out, int64_t n, int end_bit, hipStream_t s);
// Sum of n int64 into *out (device).
bsp_status reduce_sum_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s);

// Number of occupied tm x tk blocks of a CSR (bsr.hip): the key sort of the
// BSR builder without placing any value.
bsp_status bsr_count_blocks(const bsp_csr& A, int tm, int tk, hipStream_t s, int64_t* nblocks);

// DBSR of an FP64 CSR, rows sorted by (length, index) first (sort_rows) or
// kept in their own order (bsr.hip).
bsp_status dbsr_build_ordered(const bsp_csr& A, bool sort_rows, hipStream_t s, bsp_dbsr* out);
// balanceDbsr on a host block-row pointer: prefix malloc'ed (warpnum + 1).
bsp_status dbsr_balance_host(const int32_t* blcptr, int64_t brows, int32_t** prefix,
                             int32_t* warpnum);
// The CSR with rows sorted by (length, index) into device arrays of n + 1,
// nnz and nnz entries.
bsp_status csr_sort_rows(const bsp_csr& A, hipStream_t s, int32_t* rowptr, int32_t* col,
                         double* val);

// A pair of timing events, destroyed on every exit path (create() first).
struct EventPair {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  EventPair() = default;
  EventPair(const EventPair&) = delete;
  EventPair& operator=(const EventPair&) = delete;
  ~EventPair() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
  bsp_status create() {
    BSP_HIP(hipEventCreate(&e0));
    BSP_HIP(hipEventCreate(&e1));
    return BSP_OK;
  }
};

// Copies one value from device to host (synchronising the stream).
template <class T>
bsp_status fetch(const T* dptr, T* host, hipStream_t s) {
  BSP_HIP(hipMemcpyAsync(host, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

}  // namespace bsp

// ---- device helpers --------------------------------------------------------

// Sum across aligned groups of `width` lanes (width a power of two <= 64);
// every lane of the group receives the group total.
template <int W>
__device__ __forceinline__ double group_sum(double v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double group_sum_dyn(double v, int width) {
  for (int o = width >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}

template <class T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  __builtin_nontemporal_store(v, p);
}

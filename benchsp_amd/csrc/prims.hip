// Device-wide scan / sort / reduce used by the format builders (rocPRIM via
// hipCUB). Kept in one translation unit: the template instantiations are the
// slowest part of the build.
#include <hipcub/hipcub.hpp>

#include "bsp_hip.hpp"

namespace bsp {

namespace {
struct Temp {
  void* p = nullptr;
  ~Temp() { if (p) (void)hipFree(p); }
  bsp_status get(size_t bytes) {
    if (bytes == 0) bytes = 1;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      p = nullptr;
      return fail(BSP_ERR_ALLOC, "temporary storage for device primitive");
    }
    return BSP_OK;
  }
};
}  // namespace

bsp_status exclusive_scan_i32(const int32_t* in, int32_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return BSP_OK;
  size_t bytes = 0;
  BSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  Temp t;
  BSP_TRY(t.get(bytes));
  BSP_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, bytes, in, out, (int)n, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

bsp_status exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
  if (n <= 0) return BSP_OK;
  size_t bytes = 0;
  BSP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, s));
  Temp t;
  BSP_TRY(t.get(bytes));
  BSP_HIP(hipcub::DeviceScan::ExclusiveSum(t.p, bytes, in, out, (int)n, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

bsp_status sort_pairs_i32(const int32_t* keys_in, int32_t* keys_out, const int32_t* vals_in,
                          int32_t* vals_out, int64_t n, bool descending, int end_bit,
                          hipStream_t s) {
  if (n <= 0) return BSP_OK;
  size_t bytes = 0;
  if (descending)
    BSP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, bytes, keys_in, keys_out, vals_in,
                                                         vals_out, (int)n, 0, end_bit, s));
  else
    BSP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                               (int)n, 0, end_bit, s));
  Temp t;
  BSP_TRY(t.get(bytes));
  if (descending)
    BSP_HIP(hipcub::DeviceRadixSort::SortPairsDescending(t.p, bytes, keys_in, keys_out, vals_in,
                                                         vals_out, (int)n, 0, end_bit, s));
  else
    BSP_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, bytes, keys_in, keys_out, vals_in, vals_out,
                                               (int)n, 0, end_bit, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

bsp_status sort_pairs_i64key(const int64_t* keys_in, int64_t* keys_out, const int32_t* vals_in,
                             int32_t* vals_out, int64_t n, int end_bit, hipStream_t s) {
  if (n <= 0) return BSP_OK;
  size_t bytes = 0;
  BSP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, keys_in, keys_out, vals_in, vals_out,
                                             (int)n, 0, end_bit, s));
  Temp t;
  BSP_TRY(t.get(bytes));
  BSP_HIP(hipcub::DeviceRadixSort::SortPairs(t.p, bytes, keys_in, keys_out, vals_in, vals_out,
                                             (int)n, 0, end_bit, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

bsp_status reduce_sum_i64(const int64_t* in, int64_t* out, int64_t n, hipStream_t s) {
  size_t bytes = 0;
  BSP_HIP(hipcub::DeviceReduce::Sum(nullptr, bytes, in, out, (int)n, s));
  Temp t;
  BSP_TRY(t.get(bytes));
  BSP_HIP(hipcub::DeviceReduce::Sum(t.p, bytes, in, out, (int)n, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

}  // namespace bsp

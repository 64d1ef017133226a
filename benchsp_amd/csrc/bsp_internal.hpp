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

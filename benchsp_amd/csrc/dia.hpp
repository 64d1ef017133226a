// The diagonal layout's handle: built by dia.hip, read by its SpMV and by
// the batched step (batch.hip), which runs eligible matrices in it.
#pragma once

#include "bsp_hip.hpp"

namespace bsp {
constexpr int kDiaMax = 32;  // diagonals per matrix (one 32-bit mask per row)
struct DiaOffsets {
  int32_t d[kDiaMax];
};
}  // namespace bsp

struct bsp_dia_s {
  int64_t nrows = 0, ncols = 0, nnz = 0;
  int32_t ndiag = 0;
  bsp::DiaOffsets off{};        // ascending diagonal offsets (column - row)
  bsp::DevBuf<double> val;      // ndiag x nrows, diagonal-major; absent slots 0
  bsp::DevBuf<uint32_t> mask;   // nrows: bit k set when the row holds diagonal k
};

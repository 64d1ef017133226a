// Adaptive CSR block routine shared by the single-matrix plan and the batched
// persistent kernel.
#pragma once

#include "bsp_hip.hpp"

namespace bsp {

// Rows are cut into blocks on the merge-path coordinate u(r) = rowptr[r] + r:
// rows whose start coordinate falls in the same window of CHUNK units share a
// block, and every row longer than CHUNK/2 gets a block of its own. A stream
// block therefore has at most CHUNK rows and fewer than 1.5*CHUNK nonzeros,
// all of which fit in LDS.
//
// Stream block: every lane streams values/column indices (coalesced,
// non-temporal), gathers x and parks the product in LDS; then groups of
// `lpr` lanes (lpr = largest power of two with rows*lpr <= 256) reduce one
// row each. With lpr == 1 a row is summed left to right from separately
// rounded products, which is the reference CPU loop's arithmetic exactly.
// Long block: the whole workgroup strides over the one row.
//
constexpr int kWG = 256;

template <int CHUNK>
struct Ad {
  static constexpr int LONG = CHUNK / 2;
  static constexpr int MAXNNZ = CHUNK + LONG;
  static constexpr int IT = MAXNNZ / kWG;  // value/index loads per lane
};

// Workgroup i is dispatched to XCD i % 8; give each XCD a contiguous range of
// logical work so neighbouring rows (and their x entries) share an L2.
__device__ __forceinline__ int64_t xcd_remap(int64_t bid, int64_t nb) {
  constexpr int64_t X = 8;
  int64_t q = nb / X, rem = nb % X, xcd = bid % X, k = bid / X;
  return xcd * q + (xcd < rem ? xcd : rem) + k;
}

// Rows [r0, r1) holding entries [p0, p1): at most CHUNK rows and at most
// MAXNNZ entries, or one row (of any length).
template <int CHUNK, bool ACCUM, bool NT_Y = true>
__device__ __forceinline__ void adaptive_rows(const int32_t* __restrict__ rowptr,
                                              const int32_t* __restrict__ col,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              const int r0, const int p0, const int r1,
                                              const int p1) {
  using C = Ad<CHUNK>;
  __shared__ double prod[C::MAXNNZ];
  __shared__ int32_t rptr[CHUNK + 1];
  __shared__ double yb[ACCUM ? CHUNK : 1];
  __shared__ double red[kWG / kWave];
  const int tid = threadIdx.x;
  const int nnz = p1 - p0, nrows = r1 - r0;

  if (nrows == 1 && nnz > C::LONG) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = p0 + tid;
    for (; i + 3 * kWG < p1; i += 4 * kWG) {
      a0 += ld_stream(&val[i]) * x[ld_stream(&col[i])];
      a1 += ld_stream(&val[i + kWG]) * x[ld_stream(&col[i + kWG])];
      a2 += ld_stream(&val[i + 2 * kWG]) * x[ld_stream(&col[i + 2 * kWG])];
      a3 += ld_stream(&val[i + 3 * kWG]) * x[ld_stream(&col[i + 3 * kWG])];
    }
    for (; i < p1; i += kWG) a0 += ld_stream(&val[i]) * x[ld_stream(&col[i])];
    const double v = group_sum<kWave>((a0 + a1) + (a2 + a3));
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    if (tid == 0) {
      const double t = (red[0] + red[1]) + (red[2] + red[3]);
      y[r0] = ACCUM ? y[r0] + t : t;
    }
    return;
  }

  // Stream loads (lanes past the block end re-read its last entry), the
  // block's row pointers and, when accumulating, its y entries: all
  // independent, all in flight together.
  int ci[C::IT];
  double vv[C::IT];
  const int iters = (nnz + kWG - 1) / kWG;
  if (nnz > 0) {
#pragma unroll
    for (int k = 0; k < C::IT; ++k) {
      if (k < iters) {
        const int j = min(p0 + tid + k * kWG, p1 - 1);
        ci[k] = ld_stream(&col[j]);
        vv[k] = ld_stream(&val[j]);
      }
    }
  }
  for (int i = tid; i <= nrows; i += kWG) rptr[i] = rowptr[r0 + i] - p0;
  if (ACCUM)
    for (int i = tid; i < nrows; i += kWG) yb[i] = y[r0 + i];
  if (nnz > 0) {
    double xv[C::IT];
#pragma unroll
    for (int k = 0; k < C::IT; ++k)
      if (k < iters) xv[k] = x[ci[k]];
#pragma unroll
    for (int k = 0; k < C::IT; ++k) {
      const int i = tid + k * kWG;
      if (k < iters && i < nnz) prod[i] = vv[k] * xv[k];
    }
  }
  __syncthreads();

  int lpr = 1;
  while (lpr < kWave && nrows * lpr * 2 <= kWG) lpr <<= 1;
  const int g = tid / lpr, l = tid & (lpr - 1), groups = kWG / lpr;
  for (int rr = g; rr < nrows; rr += groups) {
    const int rs = rptr[rr], re = rptr[rr + 1];
    double s = 0.0;
    for (int i = rs + l; i < re; i += lpr) s += prod[i];
    s = group_sum_dyn(s, lpr);
    if (l == 0) {
      const double out = ACCUM ? yb[rr] + s : s;
      if (!ACCUM || re > rs) {
        if (NT_Y)
          st_stream(&y[r0 + rr], out);
        else
          y[r0 + rr] = out;
      }
    }
  }
}

// One block per workgroup. The block's first row and first entry come from
// one 8-byte load each ({row, entry offset} pairs built with the plan), so a
// workgroup's stream loads start after a single round trip.
template <int CHUNK, bool ACCUM, bool NT_Y = true>
__device__ __forceinline__ void adaptive_block(const int32_t* __restrict__ rowptr,
                                               const int32_t* __restrict__ col,
                                               const double* __restrict__ val,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               const int2* __restrict__ blk, int64_t b) {
  const int2 s0 = blk[b], s1 = blk[b + 1];
  adaptive_rows<CHUNK, ACCUM, NT_Y>(rowptr, col, val, x, y, s0.x, s0.y, s1.x, s1.y);
}

}  // namespace bsp

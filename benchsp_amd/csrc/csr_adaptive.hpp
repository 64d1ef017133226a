// Adaptive CSR block routine shared by the single-matrix plan and the batched
// persistent kernel.
#pragma once

#include <type_traits>

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
typedef double dbl2v __attribute__((ext_vector_type(2)));

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
// MAXNNZ entries, or one row (of any length). Column indices are int32, or
// (CI = int16_t) offsets from the block's first row r0.
template <int CHUNK, bool ACCUM, bool NT_Y = true, class CI = int32_t>
__device__ __forceinline__ void adaptive_rows(const int32_t* __restrict__ rowptr,
                                              const CI* __restrict__ col,
                                              const double* __restrict__ val,
                                              const double* __restrict__ x, double* __restrict__ y,
                                              const int r0, const int p0, const int r1,
                                              const int p1) {
  constexpr bool NARROW = std::is_same<CI, int16_t>::value;
  const int cbase = NARROW ? r0 : 0;
  using C = Ad<CHUNK>;
  __shared__ __attribute__((aligned(16))) double prod[C::MAXNNZ + 2];
  __shared__ int32_t rptr[CHUNK + 1];
  __shared__ double yb[ACCUM ? CHUNK : 1];
  __shared__ double red[kWG / kWave];
  const int tid = threadIdx.x;
  const int nnz = p1 - p0, nrows = r1 - r0;

  if (nrows == 1 && nnz > C::LONG) {
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
    int i = p0 + tid;
    for (; i + 3 * kWG < p1; i += 4 * kWG) {
      a0 += ld_stream(&val[i]) * x[cbase + ld_stream(&col[i])];
      a1 += ld_stream(&val[i + kWG]) * x[cbase + ld_stream(&col[i + kWG])];
      a2 += ld_stream(&val[i + 2 * kWG]) * x[cbase + ld_stream(&col[i + 2 * kWG])];
      a3 += ld_stream(&val[i + 3 * kWG]) * x[cbase + ld_stream(&col[i + 3 * kWG])];
    }
    for (; i < p1; i += kWG) a0 += ld_stream(&val[i]) * x[cbase + ld_stream(&col[i])];
    const double v = group_sum<kWave>((a0 + a1) + (a2 + a3));
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    if (tid == 0) {
      const double t = (red[0] + red[1]) + (red[2] + red[3]);
      y[r0] = ACCUM ? y[r0] + t : t;
    }
    return;
  }

  if constexpr (NARROW) {
    // Entries in pairs aligned to even positions: lane t takes entries
    // e, e+1 (e = (p0 & ~1) + 2t + 512k) with one 16-byte value load and one
    // 4-byte load of two 16-bit offsets (the offset array carries one
    // element of padding; a value pair that would cross the block end is
    // loaded as one value). An element of a pair outside [p0, p1) belongs
    // to a neighbouring block: its offset is never used to index x.
    constexpr int IT2 = (C::MAXNNZ + 1 + 2 * kWG - 1) / (2 * kWG);
    const int pa = p0 & ~1;
    const int iters2 = ((p1 - pa + 1) / 2 + kWG - 1) / kWG;
    dbl2v vp[IT2];
    int cw[IT2];
#pragma unroll
    for (int k = 0; k < IT2; ++k) {
      const int e = pa + 2 * (tid + k * kWG);
      if (k < iters2 && e < p1) {
        if (e + 1 < p1) {
          vp[k] = ld_stream(reinterpret_cast<const dbl2v*>(val + e));
        } else {
          vp[k].x = ld_stream(&val[e]);
          vp[k].y = 0.0;
        }
        cw[k] = ld_stream(reinterpret_cast<const int*>(col + e));
      }
    }
    // products indexed from pa (row pointers too), so a lane's pair is one
    // aligned 16-byte LDS store
    for (int i = tid; i <= nrows; i += kWG) rptr[i] = rowptr[r0 + i] - pa;
    if (ACCUM)
      for (int i = tid; i < nrows; i += kWG) yb[i] = y[r0 + i];
    double xa[IT2], xb[IT2];
#pragma unroll
    for (int k = 0; k < IT2; ++k) {
      const int e = pa + 2 * (tid + k * kWG);
      const bool in0 = k < iters2 && e >= p0 && e < p1, in1 = k < iters2 && e + 1 < p1;
      xa[k] = in0 ? x[cbase + (int16_t)(cw[k] & 0xffff)] : 0.0;
      xb[k] = in1 ? x[cbase + (int16_t)((unsigned)cw[k] >> 16)] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < IT2; ++k) {
      const int e = pa + 2 * (tid + k * kWG);
      if (k < iters2 && e < p1)  // an element outside the block stores 0, never read
        *reinterpret_cast<dbl2v*>(&prod[e - pa]) = dbl2v{vp[k].x * xa[k], vp[k].y * xb[k]};
    }
    __syncthreads();
  } else {
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
        ci[k] = cbase + ld_stream(&col[j]);
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
  }

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
template <int CHUNK, bool ACCUM, bool NT_Y = true, class CI = int32_t>
__device__ __forceinline__ void adaptive_block(const int32_t* __restrict__ rowptr,
                                               const CI* __restrict__ col,
                                               const double* __restrict__ val,
                                               const double* __restrict__ x, double* __restrict__ y,
                                               const int2* __restrict__ blk, int64_t b) {
  const int2 s0 = blk[b], s1 = blk[b + 1];
  adaptive_rows<CHUNK, ACCUM, NT_Y, CI>(rowptr, col, val, x, y, s0.x, s0.y, s1.x, s1.y);
}

}  // namespace bsp

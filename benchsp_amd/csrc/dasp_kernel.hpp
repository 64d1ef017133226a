// DASP SpMV over the layout of dasp.hip, shared by the single-matrix launch
// (bsp_dasp_spmv) and the batched step, which runs the same workgroups as
// tiles of its one launch.
#pragma once

#include "bsp_hip.hpp"

namespace bsp {
namespace dasp {

constexpr int kTB = 256;
constexpr int BS = 8;  // rows per row block (dasp.cu:27)

struct DaspArgs {
  const double* x;  // set for the single-matrix launch; dasp_block takes x, y apart
  double* y;
  const int32_t* order;
  int permuted;
  // long
  const double* lv;
  const int32_t* lc;
  const int32_t* lrpt;
  int32_t nl;
  // row blocks
  const double* rv;
  const int32_t* rc;
  const int32_t* bptr;
  const double* iv;
  const int32_t* ic;
  const int32_t* irpt;
  int32_t nb, blocknum;
  // short
  const double* sv;
  const int32_t* sc;
  int32_t s1, c13, s3, s4, s2, nz, f13, f34;
  // workgroup ranges
  int64_t wg_long, wg_reg, wg_short;
};

__device__ __forceinline__ void dasp_store(const DaspArgs& a, double* y, int64_t k, double v) {
  y[a.permuted ? k : a.order[k]] = v;
}

typedef double dbl4m __attribute__((ext_vector_type(4)));

// Diagonal of the 16x16 result of v_mfma_f64_16x16x4f64 (lane l holds
// C[l/16 + 4v][l%16] in acc[v]): C[R][R] sits in acc[R/4] of lane
// 16(R%4) + R; every other lane returns 0.
__device__ __forceinline__ double mfma_diag(const dbl4m& acc, int lane) {
  const int R = lane & 15;
  if ((lane >> 4) != (R & 3)) return 0.0;
  const int v = R >> 2;
  return v == 0 ? acc[0] : v == 1 ? acc[1] : v == 2 ? acc[2] : acc[3];
}

// Workgroup `bid` of the fused launch: [long rows: a workgroup each][row
// blocks: a wave each][short and empty rows: a thread each]. Output position
// k is the row's place in order_rid. x and y are passed apart from the
// layout's arguments, so the batched step reads those from global memory
// as they are and supplies its own vectors.
//
// MF: long rows and row blocks on FP64 MFMA, as the reference's fused
// kernel does on mma.m8n8k4 (src/dasp.cu:405-812). Lane l = 16t + i loads
// four consecutive entries (one 32-byte value and one 16-byte column load)
// and, in product u, supplies entry u as A[i][t] with its gathered x as
// B[t][i], so v_mfma_f64_16x16x4f64's diagonal C[i][i] sums the entries
// lanes (i, 0..3) hold. A row block's step covers eight 32-entry slabs
// (8 rows x 4 columns each): lane (i, t) takes row i%8 of slab 2t + i/8, so
// row g of the block is C[g][g] + C[g+8][g+8]; a long row's step covers four
// 64-entry windows, lane (i, t) entries 4i..4i+3 of window t (the row sum is
// the trace). Short rows stay on VALU in both.
template <bool MF = false>
__device__ __forceinline__ void dasp_block(const DaspArgs& a, const double* __restrict__ x,
                                           double* __restrict__ y, int64_t bid) {
  __shared__ double red[kTB / kWave];
  const int tid = threadIdx.x, lane = tid & (kWave - 1);
  if (MF && bid < a.wg_long) {
    typedef int i4v __attribute__((ext_vector_type(4)));
    const int i = (int)bid;
    const int64_t base = (int64_t)a.lrpt[i] * 64;
    const int64_t nwin = a.lrpt[i + 1] - a.lrpt[i];  // 64-entry windows, zero padded
    const int t = lane >> 4, off = 4 * (lane & 15);
    dbl4m acc = {0.0, 0.0, 0.0, 0.0};
    for (int64_t w0 = 4 * (tid / kWave); w0 < nwin; w0 += 16) {  // four windows per step
      dbl4m v = {0.0, 0.0, 0.0, 0.0};
      double xv[4] = {0.0, 0.0, 0.0, 0.0};
      if (w0 + t < nwin) {
        const int64_t p = base + (w0 + t) * 64 + off;
        v = ld_stream(reinterpret_cast<const dbl4m*>(a.lv + p));
        const i4v c = ld_stream(reinterpret_cast<const i4v*>(a.lc + p));
        xv[0] = x[c.x];
        xv[1] = x[c.y];
        xv[2] = x[c.z];
        xv[3] = x[c.w];
      }
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, xv[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, xv[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.z, xv[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.w, xv[3], acc, 0, 0, 0);
    }
    const double v = group_sum<kWave>(mfma_diag(acc, lane));
    if (lane == 0) red[tid / kWave] = v;
    __syncthreads();
    if (tid == 0) dasp_store(a, y, i, (red[0] + red[1]) + (red[2] + red[3]));
    return;
  }
  if (!MF && bid < a.wg_long) {
    const int i = (int)bid;
    const int64_t base = (int64_t)a.lrpt[i] * 64;
    const int64_t len = (int64_t)(a.lrpt[i + 1] - a.lrpt[i]) * 64;  // zero padded
    double s0 = 0.0, s1 = 0.0;
    int64_t j = tid;
    for (; j + kTB < len; j += 2 * kTB) {
      s0 += ld_stream(&a.lv[base + j]) * x[ld_stream(&a.lc[base + j])];
      s1 += ld_stream(&a.lv[base + j + kTB]) * x[ld_stream(&a.lc[base + j + kTB])];
    }
    if (j < len) s0 += ld_stream(&a.lv[base + j]) * x[ld_stream(&a.lc[base + j])];
    const double v = group_sum<kWave>(s0 + s1);
    if (lane == 0) red[tid / kWave] = v;
    __syncthreads();
    if (tid == 0) dasp_store(a, y, i, (red[0] + red[1]) + (red[2] + red[3]));
    return;
  }
  if (MF && bid < a.wg_long + a.wg_reg) {
    typedef int i4v __attribute__((ext_vector_type(4)));
    const int b = (int)((bid - a.wg_long) * (kTB / kWave) + tid / kWave);
    if (b >= a.blocknum) return;
    const int p0 = a.bptr[b], p1 = a.bptr[b + 1];
    const int g = lane & 15, ts = lane >> 4;
    const int off = (2 * ts + (g >> 3)) * 32 + (g & 7) * 4;  // slab, row
    dbl4m acc = {0.0, 0.0, 0.0, 0.0};
    for (int base = p0; base < p1; base += 4 * kWave) {  // eight slabs per step
      const int p = base + off;
      dbl4m v = {0.0, 0.0, 0.0, 0.0};
      double xv[4] = {0.0, 0.0, 0.0, 0.0};
      if (p < p1) {
        v = ld_stream(reinterpret_cast<const dbl4m*>(a.rv + p));
        const i4v c = ld_stream(reinterpret_cast<const i4v*>(a.rc + p));
        xv[0] = x[c.x];
        xv[1] = x[c.y];
        xv[2] = x[c.z];
        xv[3] = x[c.w];
      }
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.x, xv[0], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.y, xv[1], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.z, xv[2], acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f64_16x16x4f64(v.w, xv[3], acc, 0, 0, 0);
    }
    // row g (0..7): C[g][g] on lane 16(g%4) + g plus C[g+8][g+8] eight lanes up
    const double d = mfma_diag(acc, lane);
    const double rs = d + __shfl_down(d, 8, 64);
    const int gq = lane >> 3, q = lane & 7;  // lanes gq*8 + q: row gq's irregular tail
    const double reg = __shfl(rs, 16 * (gq & 3) + gq, 64);
    const int k = b * BS + gq;
    double tail = 0.0;
    if (k < a.nb)
      for (int t = a.irpt[k] + q, e = a.irpt[k + 1]; t < e; t += 8)
        tail += ld_stream(&a.iv[t]) * x[ld_stream(&a.ic[t])];
    tail += __shfl_xor(tail, 1, 64);
    tail += __shfl_xor(tail, 2, 64);
    tail += __shfl_xor(tail, 4, 64);
    if (q == 0 && k < a.nb) dasp_store(a, y, (int64_t)a.nl + k, reg + tail);
    return;
  }
  if (!MF && bid < a.wg_long + a.wg_reg) {
    const int b = (int)((bid - a.wg_long) * (kTB / kWave) + tid / kWave);
    if (b >= a.blocknum) return;
    const int p0 = a.bptr[b], p1 = a.bptr[b + 1];
    // two slabs per step: lane l reads entry l of the 64-entry window,
    // i.e. slab (l/32), row g = (l%32)/4, column t = l%4
    // two 64-entry windows per step, both windows' loads in flight (lanes
    // past the block's end re-read its first entry and add nothing)
    double acc0 = 0.0, acc1 = 0.0;
    for (int base = p0; base < p1; base += 2 * kWave) {
      double v[2], xv[2];
      int c[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int p = base + u * kWave + lane;
        const int pc = p < p1 ? p : p0;
        v[u] = p < p1 ? ld_stream(&a.rv[pc]) : 0.0;
        c[u] = ld_stream(&a.rc[pc]);
      }
#pragma unroll
      for (int u = 0; u < 2; ++u) xv[u] = x[c[u]];
      acc0 += v[0] * xv[0];
      acc1 += v[1] * xv[1];
    }
    double acc = acc0 + acc1;
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    acc += __shfl_xor(acc, 32, 64);
    // lanes g*8 + q (q = 0..7) add row g's irregular tail, stride 8
    const int g = lane >> 3, q = lane & 7;
    const double reg = __shfl(acc, g * 4, 64);
    const int k = b * BS + g;
    double tail = 0.0;
    if (k < a.nb)
      for (int t = a.irpt[k] + q, e = a.irpt[k + 1]; t < e; t += 8)
        tail += ld_stream(&a.iv[t]) * x[ld_stream(&a.ic[t])];
    tail += __shfl_xor(tail, 1, 64);
    tail += __shfl_xor(tail, 2, 64);
    tail += __shfl_xor(tail, 4, 64);
    if (q == 0 && k < a.nb) dasp_store(a, y, (int64_t)a.nl + k, reg + tail);
    return;
  }
  // short rows
  int64_t q = (bid - a.wg_long - a.wg_reg) * kTB + tid;
  const int64_t base = (int64_t)a.nl + a.nb;
  const int64_t k = base + q;
  double v = 0.0;
  if (q < a.s1) {
    v = a.sv[q] * x[a.sc[q]];
  } else if ((q -= a.s1) < 2 * (int64_t)a.c13) {
    const int64_t t = q / 16, w = q % 16;
    const int64_t o = a.s1 + t * 32 + (w % 8) * 4;
    if (w < 8) {
      v = a.sv[o] * x[a.sc[o]];
    } else {
      v = a.sv[o + 1] * x[a.sc[o + 1]] + a.sv[o + 2] * x[a.sc[o + 2]] +
          a.sv[o + 3] * x[a.sc[o + 3]];
    }
  } else if ((q -= 2 * (int64_t)a.c13) < a.s3 + a.s4) {
    const int64_t o = a.s1 + a.f13 + q * 4;
    const int n = q < a.s3 ? 3 : 4;
    for (int e = 0; e < n; ++e) v += a.sv[o + e] * x[a.sc[o + e]];
  } else if ((q -= a.s3 + a.s4) < a.s2) {
    const int64_t t = q / 16, j = q % 16;
    const int64_t o = a.s1 + a.f13 + a.f34 + t * 32 + (j % 8) * 4 + (j / 8) * 2;
    v = a.sv[o] * x[a.sc[o]] + a.sv[o + 1] * x[a.sc[o + 1]];
  } else if ((q -= a.s2) >= a.nz) {
    return;
  }
  dasp_store(a, y, k, v);
}


// Arguments of one SpMV over view `v` (x and y may be set later).
inline DaspArgs dasp_args(const bsp_dasp_view& v, const double* x, double* y, int permuted) {
  DaspArgs a{};
  a.x = x;
  a.y = y;
  a.order = v.order_rid;
  a.permuted = permuted;
  a.lv = v.long_val;
  a.lc = v.long_cid;
  a.lrpt = v.long_rpt;
  a.nl = v.row_long;
  a.rv = v.reg_val;
  a.rc = v.reg_cid;
  a.bptr = v.block_ptr;
  a.iv = v.irreg_val;
  a.ic = v.irreg_cid;
  a.irpt = v.irreg_rpt;
  a.nb = v.row_block;
  a.blocknum = v.blocknum;
  a.sv = v.short_val;
  a.sc = v.short_cid;
  a.s1 = v.short_row_1;
  a.c13 = v.common_13;
  a.s3 = v.short_row_3;
  a.s4 = v.short_row_4;
  a.s2 = v.short_row_2;
  a.nz = v.row_zero;
  a.f13 = v.fill0_nnz_short13;
  a.f34 = v.fill0_nnz_short34;
  const int64_t nshort = (int64_t)a.s1 + 2LL * a.c13 + a.s3 + a.s4 + a.s2 + a.nz;
  a.wg_long = a.nl;
  a.wg_reg = cdiv(a.blocknum, kTB / kWave);
  a.wg_short = cdiv(nshort, kTB);
  return a;
}

}  // namespace dasp
}  // namespace bsp

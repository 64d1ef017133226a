// Many CSR matrices of uneven size per step: FP64 SpMV and an 8-column FP16
// SpMM for each, every matrix's work cut into tiles of one persistent launch
// per operation. A per-matrix autotuner picks the format (CSR, 4x4 or 8x8
// BSR, DASP or the diagonal layout for the SpMV; CSR or the diagonal layout
// for the SpMM) and the tile window (rows per workgroup, hence lanes per
// row) by timing the candidates; a hipGraph holds a whole step. A plain
// schedule (one thread-per-row launch per matrix and operation) is kept as
// the comparison.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "csr_adaptive.hpp"
#include "dasp_kernel.hpp"
#include "dia.hpp"

namespace bsp {
namespace {

constexpr int kCap = 1024;  // window capacity of adaptive_rows<kCap>
constexpr int kCols = 8;    // SpMM columns
constexpr int kKindCsr = 0, kKindBsr4 = 1, kKindBsr8 = 2, kKindDasp = 3, kKindDia = 4;
// Diagonals a batch matrix may have to run in the diagonal layout (the
// 5- and 7-point stencils of a mixed batch have 5 and 7).
constexpr int kBatchDia = 8;
constexpr int kDaspPerTile = 4;  // workgroups of a matrix's DASP launch per tile
// A BSR (4x4 or 8x8) is a candidate only where it moves no more bytes than
// CSR: at least 11/16 of its block entries filled on average.
constexpr double kBsrFill = 11.0 / 16.0;
constexpr int kMaxGrid = 2048;  // bound on resident workgroups (8 per CU)
// Default launch size: this many times the resident workgroups. Short tile
// ranges let the hardware dispatcher even out what the weighted cuts miss
// (512-matrix mix: SpMV 0.418 -> 0.338 ms, SpMM 0.657 -> 0.533 ms against
// one resident wave of workgroups).
constexpr int kOversub = 16;
constexpr int kTileOverhead = 64;  // a tile's fixed cost in units of rows+entries
const int kWindows[] = {128, 256, 512, 1024};  // tile windows (rows+entries)

struct BTile {
  int32_t m, kind, r0, r1, p0, p1, pad0, pad1;
};

struct BMat {
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  const uint16_t* hval;
  const int32_t* brp;  // 4x4 BSR (a candidate or chosen)
  const int32_t* bci;
  const double* bval;
  const int32_t* brp8;  // 8x8 BSR (a candidate or chosen)
  const int32_t* bci8;
  const double* bval8;
  const dasp::DaspArgs* dasp;  // DASP layout (a candidate or chosen), x/y unset
  const double* dval;          // diagonal layout (a candidate or chosen): K x nrows
  const uint16_t* dhval;       // its FP16 copy, for the SpMM
  const uint32_t* dmask;       // row masks
  int64_t xoff, yoff;
  int32_t nrows, ncols;
  int32_t ndiag, pad;
  int32_t doff[kBatchDia];     // ascending diagonal offsets
};

__device__ __forceinline__ float half_to_f32(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

// Block rows [br0, br1) (at most four) of a 4x4 BSR, a wave per block row;
// lane = slot*4 + r, 16 blocks per step, partial sums of equal r added across
// slots with lane shuffles.
__device__ __forceinline__ void bsr4_rows(const int32_t* __restrict__ brp,
                                          const int32_t* __restrict__ bci,
                                          const double* __restrict__ bval,
                                          const double* __restrict__ x, double* __restrict__ y,
                                          int br0, int br1, int nrows, int ncols) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int br = br0 + wave;
  if (br >= br1) return;
  const int slot = lane >> 2, r = lane & 3;
  double acc = 0.0;
  for (int k = brp[br] + slot, e = brp[br + 1]; k < e; k += 16) {
    const int c0 = bci[k] * 4;
    const double* p = bval + (int64_t)k * 16 + r * 4;
    if (c0 + 4 <= ncols) {
      acc += ld_stream(&p[0]) * x[c0] + ld_stream(&p[1]) * x[c0 + 1] +
             ld_stream(&p[2]) * x[c0 + 2] + ld_stream(&p[3]) * x[c0 + 3];
    } else {
      for (int c = 0; c < ncols - c0; ++c) acc += ld_stream(&p[c]) * x[c0 + c];
    }
  }
#pragma unroll
  for (int o = 4; o < kWave; o <<= 1) acc += __shfl_xor(acc, o, 64);
  const int row = br * 4 + r;
  if (slot == 0 && row < nrows) y[row] = acc;
}

// Block rows [br0, br1) (at most four) of an 8x8 BSR, a wave per block row;
// lane = slot*8 + r, 8 blocks per step, partial sums of equal r added across
// slots with lane shuffles.
__device__ __forceinline__ void bsr8_rows(const int32_t* __restrict__ brp,
                                          const int32_t* __restrict__ bci,
                                          const double* __restrict__ bval,
                                          const double* __restrict__ x, double* __restrict__ y,
                                          int br0, int br1, int nrows, int ncols) {
  const int wave = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int br = br0 + wave;
  if (br >= br1) return;
  const int slot = lane >> 3, r = lane & 7;
  double acc = 0.0;
  for (int k = brp[br] + slot, e = brp[br + 1]; k < e; k += 8) {
    const int c0 = bci[k] * 8;
    const double* p = bval + (int64_t)k * 64 + r * 8;
    if (c0 + 8 <= ncols) {
      double s0 = 0.0, s1 = 0.0;
#pragma unroll
      for (int c = 0; c < 8; c += 2) {
        s0 += ld_stream(&p[c]) * x[c0 + c];
        s1 += ld_stream(&p[c + 1]) * x[c0 + c + 1];
      }
      acc += s0 + s1;
    } else {
      for (int c = 0; c < ncols - c0; ++c) acc += ld_stream(&p[c]) * x[c0 + c];
    }
  }
#pragma unroll
  for (int o = 8; o < kWave; o <<= 1) acc += __shfl_xor(acc, o, 64);
  const int row = br * 8 + r;
  if (slot == 0 && row < nrows) y[row] = acc;
}

// Rows [r0, r1) (at most kWG) of y = A x for a matrix in the diagonal layout,
// a thread per row, four diagonals at a time (registers: the batch kernel is
// held to 64 VGPRs): their value slots and x at the clamped columns loaded
// first, then added in ascending order where the row's mask bit is set,
// each product rounded before its add (the CPU loop's bits).
constexpr int kDiaStep = 4;
__device__ __forceinline__ void dia_rows_f64(const BMat* __restrict__ d, const double* x,
                                             double* y, int r0, int r1) {
#pragma clang fp contract(off)
  const int r = r0 + (int)threadIdx.x;
  if (r >= r1) return;
  const int n = d->nrows, K = d->ndiag, nc = d->ncols;
  const uint32_t m = d->dmask[r];
  double s = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < kBatchDia; k0 += kDiaStep) {
    if (k0 >= K) break;
    double a[kDiaStep], b[kDiaStep];
#pragma unroll
    for (int i = 0; i < kDiaStep; ++i)
      if (k0 + i < K) {
        a[i] = ld_stream(&d->dval[(int64_t)(k0 + i) * n + r]);
        b[i] = x[min(max(r + d->doff[k0 + i], 0), nc - 1)];
      }
#pragma unroll
    for (int i = 0; i < kDiaStep; ++i)
      if (k0 + i < K) {
        const double t = s + a[i] * b[i];
        s = ((m >> (k0 + i)) & 1u) ? t : s;
      }
  }
  y[r] = s;
}

// Rows [r0, r1) (at most kWG) of Y = A X for a matrix in the diagonal layout,
// a thread per row, four diagonals at a time: per diagonal a 2-byte value
// and the 16-byte X row at the clamped column (consecutive rows: coalesced),
// FP32 sums in ascending diagonal order over the slots the row holds.
__device__ __forceinline__ void dia_rows_spmm8(const BMat* __restrict__ d, const uint4* X,
                                               float* Y, int r0, int r1) {
  const int r = r0 + (int)threadIdx.x;
  if (r >= r1) return;
  const int n = d->nrows, K = d->ndiag, nc = d->ncols;
  const uint32_t m = d->dmask[r];
  float acc[kCols] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < kBatchDia; k0 += kDiaStep) {
    if (k0 >= K) break;
    uint16_t hv[kDiaStep];
    uint4 xq[kDiaStep];
#pragma unroll
    for (int i = 0; i < kDiaStep; ++i)
      if (k0 + i < K) {
        hv[i] = ld_stream(&d->dhval[(int64_t)(k0 + i) * n + r]);
        xq[i] = X[min(max(r + d->doff[k0 + i], 0), nc - 1)];
      }
#pragma unroll
    for (int i = 0; i < kDiaStep; ++i)
      if (k0 + i < K && ((m >> (k0 + i)) & 1u)) {
        const float v = half_to_f32(hv[i]);
        const uint32_t w[4] = {xq[i].x, xq[i].y, xq[i].z, xq[i].w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          acc[2 * c] += v * half_to_f32((uint16_t)(w[c] & 0xffff));
          acc[2 * c + 1] += v * half_to_f32((uint16_t)(w[c] >> 16));
        }
      }
  }
  float4* out = reinterpret_cast<float4*>(Y + (int64_t)r * kCols);
  out[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
  out[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
}

// Rows [r0, r1) with entries [p0, p1) of Y = A X (at most kCapM rows and
// 1.5 kCapM entries, or one row), X rows of 8 halves, Y rows of 8 floats.
// Entries are streamed coalesced, one lane per entry: a 16-byte load brings
// the entry's X row. The products go through LDS four columns at a time
// (column-major, odd stride: columns in distinct banks), and q lanes per
// (row, column) pair add a row's products (q = the largest power of two
// <= 8 with rows*4*q <= 128); half-width staging keeps the LDS per
// workgroup at 14 KB, so eight workgroups fit a CU. One long row takes the
// whole workgroup.
constexpr int kCapM = 512;
constexpr int kMaxM = kCapM + kCapM / 2;
constexpr int kItM = kMaxM / kWG;
constexpr int kHalf = kCols / 2;

__device__ __forceinline__ void spmm8_rows(const int32_t* __restrict__ rowptr,
                                           const int32_t* __restrict__ col,
                                           const uint16_t* __restrict__ hval,
                                           const uint4* __restrict__ X, float* __restrict__ Y,
                                           int r0, int p0, int r1, int p1) {
  __shared__ float prod[kHalf][kMaxM + 1];
  __shared__ int32_t rptr[kCapM + 1];
  __shared__ float red[kWG / kWave][kCols];
  const int tid = threadIdx.x;
  const int nrows = r1 - r0, nnz = p1 - p0;
  auto unpack = [](uint4 q, float v, float* acc) {
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[2 * k] += v * half_to_f32((uint16_t)(w[k] & 0xffff));
      acc[2 * k + 1] += v * half_to_f32((uint16_t)(w[k] >> 16));
    }
  };
  if (nrows == 1 && nnz > kCapM / 2) {
    float acc[kCols] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int i = p0 + tid; i < p1; i += kWG) unpack(X[col[i]], half_to_f32(hval[i]), acc);
#pragma unroll
    for (int k = 0; k < kCols; ++k) {
      float v = acc[k];
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      acc[k] = v;
    }
    if ((tid & (kWave - 1)) == 0)
#pragma unroll
      for (int k = 0; k < kCols; ++k) red[tid / kWave][k] = acc[k];
    __syncthreads();
    if (tid < kCols)
      Y[(int64_t)r0 * kCols + tid] = (red[0][tid] + red[1][tid]) + (red[2][tid] + red[3][tid]);
    return;
  }
  const int iters = (nnz + kWG - 1) / kWG;
  int ci[kItM];
  uint16_t hv[kItM];
  if (nnz > 0) {
#pragma unroll
    for (int k = 0; k < kItM; ++k)
      if (k < iters) {
        const int j = min(p0 + tid + k * kWG, p1 - 1);
        ci[k] = ld_stream(&col[j]);
        hv[k] = ld_stream(&hval[j]);
      }
  }
  for (int i = tid; i <= nrows; i += kWG) rptr[i] = rowptr[r0 + i] - p0;
  uint4 xq[kItM];
  if (nnz > 0) {
#pragma unroll
    for (int k = 0; k < kItM; ++k)
      if (k < iters) xq[k] = X[ci[k]];
  }
  int q = 1;
  while (q < 8 && nrows * kHalf * q * 2 <= kWG) q <<= 1;
  const int sub = tid & (q - 1);
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (h) __syncthreads();  // the first half's sums have read prod
    if (nnz > 0) {
#pragma unroll
      for (int k = 0; k < kItM; ++k) {
        const int i = tid + k * kWG;
        if (k < iters && i < nnz) {
          const float v = half_to_f32(hv[k]);
          const uint32_t w0 = h ? xq[k].z : xq[k].x, w1 = h ? xq[k].w : xq[k].y;
          prod[0][i] = v * half_to_f32((uint16_t)(w0 & 0xffff));
          prod[1][i] = v * half_to_f32((uint16_t)(w0 >> 16));
          prod[2][i] = v * half_to_f32((uint16_t)(w1 & 0xffff));
          prod[3][i] = v * half_to_f32((uint16_t)(w1 >> 16));
        }
      }
    }
    __syncthreads();
    for (int pr = tid / q; pr < nrows * kHalf; pr += kWG / q) {
      const int rr = pr >> 2, c = pr & 3;
      float sum = 0.f;
      for (int i = rptr[rr] + sub, e = rptr[rr + 1]; i < e; i += q) sum += prod[c][i];
      for (int o = q >> 1; o > 0; o >>= 1) sum += __shfl_xor(sum, o, 64);
      if (sub == 0) Y[(int64_t)(r0 + rr) * kCols + h * kHalf + c] = sum;
    }
  }
}

// Workgroups [v0, v1) of a matrix's fused DASP launch. Kept out of line so
// its register needs (the DASP paths are the batch kernel's largest) stay in
// this function instead of raising the pressure on the CSR and BSR paths.
__device__ __attribute__((noinline)) void dasp_tile(const dasp::DaspArgs* a, const double* x,
                                                    double* y, int v0, int v1) {
  for (int v = v0; v < v1; ++v) {
    dasp::dasp_block(*a, x, y, v);
    __syncthreads();  // the long-row part reuses its LDS
  }
}

// Persistent launches: workgroup w (XCD-remapped, so an XCD owns a contiguous
// run of tiles) runs tiles [wg_start[w], wg_start[w+1]) in order; the cuts
// balance rows+entries plus a fixed cost per tile. No counters, no atomics:
// every row of every matrix is written by exactly one tile.
// (The SpMV kernel is held to 64 VGPRs, eight waves per SIMD; unconstrained
// it takes 66 and seven.) DASP tiles run only in the DASP = true instance:
// their code would otherwise add call and scratch overhead to every launch
// (measured: the CSR-only step 0.33 -> 1.23 ms).
template <bool DASP>
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(8, 8))) void batch_spmv_f64(
    const BTile* __restrict__ tiles, const int32_t* __restrict__ wg_start,
    const BMat* __restrict__ mats, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  const int t1 = wg_start[w + 1];
  for (int t = wg_start[w]; t < t1; ++t) {
    const BTile T = tiles[t];
    const BMat* d = mats + T.m;
    if (T.kind == kKindCsr)
      adaptive_rows<kCap, false, false>(d->rowptr, d->col, d->val, x + d->xoff, y + d->yoff,
                                        T.r0, T.p0, T.r1, T.p1);
    else if (T.kind == kKindBsr4)
      bsr4_rows(d->brp, d->bci, d->bval, x + d->xoff, y + d->yoff, T.r0, T.r1, d->nrows,
                d->ncols);
    else if (T.kind == kKindBsr8)
      bsr8_rows(d->brp8, d->bci8, d->bval8, x + d->xoff, y + d->yoff, T.r0, T.r1, d->nrows,
                d->ncols);
    else if (T.kind == kKindDia)
      dia_rows_f64(d, x + d->xoff, y + d->yoff, T.r0, T.r1);
    else if constexpr (DASP)
      dasp_tile(d->dasp, x + d->xoff, y + d->yoff, T.r0, T.r1);
    __syncthreads();
  }
}

__global__ __launch_bounds__(kWG) void batch_spmm8_f16(const BTile* __restrict__ tiles,
                                                       const int32_t* __restrict__ wg_start,
                                                       const BMat* __restrict__ mats,
                                                       const uint4* __restrict__ X,
                                                       float* __restrict__ Y) {
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  int t = wg_start[w];
  const int t1 = wg_start[w + 1];
  for (; t < t1; ++t) {
    const BTile T = tiles[t];
    const BMat* m = mats + T.m;
    if (T.kind == kKindDia)
      dia_rows_spmm8(m, X + m->xoff, Y + m->yoff * kCols, T.r0, T.r1);
    else
      spmm8_rows(m->rowptr, m->col, m->hval, X + m->xoff, Y + m->yoff * kCols, T.r0, T.p0,
                 T.r1, T.p1);
    __syncthreads();  // LDS is reused by the next tile
  }
}

// The plain schedule: thread per row, one launch per matrix.
__global__ __launch_bounds__(kWG) void plain_spmv_f64(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, int n) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) s += val[j] * x[col[j]];
  y[r] = s;
}

__global__ __launch_bounds__(kWG) void plain_spmm8_f16(const int32_t* __restrict__ rowptr,
                                                       const int32_t* __restrict__ col,
                                                       const uint16_t* __restrict__ hval,
                                                       const uint4* __restrict__ X,
                                                       float* __restrict__ Y, int n) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float acc[kCols] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) {
    const float v = half_to_f32(hval[j]);
    const uint4 q = X[col[j]];
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      acc[2 * k] += v * half_to_f32((uint16_t)(w[k] & 0xffff));
      acc[2 * k + 1] += v * half_to_f32((uint16_t)(w[k] >> 16));
    }
  }
  float4* out = reinterpret_cast<float4*>(Y + (int64_t)r * kCols);
  out[0] = make_float4(acc[0], acc[1], acc[2], acc[3]);
  out[1] = make_float4(acc[4], acc[5], acc[6], acc[7]);
}

__global__ void f64_to_f16(const double* __restrict__ a, int64_t n, uint16_t* __restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = __builtin_bit_cast(uint16_t, (_Float16)a[i]);
}

// ---- host-side schedule ------------------------------------------------------

// Tile starts of one matrix for window c: the rule of csr_block_flags (a new
// tile where the merge-path window of u(r) = rowptr[r] + r changes, and around
// every row longer than c/2).
void csr_tiles(const std::vector<int32_t>& rp, int m, int c, std::vector<BTile>& out) {
  const int n = (int)rp.size() - 1;
  const int longrow = c / 2;
  int start = 0;
  for (int r = 1; r <= n; ++r) {
    bool cut = r == n;
    if (!cut) {
      const int64_t key = ((int64_t)rp[r] + r) / c, keyp = ((int64_t)rp[r - 1] + r - 1) / c;
      cut = key != keyp || rp[r + 1] - rp[r] > longrow || rp[r] - rp[r - 1] > longrow;
    }
    if (cut) {
      out.push_back(BTile{m, kKindCsr, start, r, rp[start], rp[r], 0, 0});
      start = r;
    }
  }
}

// Four block rows (a wave each) per tile; kind kKindBsr4 or kKindBsr8.
void bsr_tiles(const std::vector<int32_t>& brp, int m, int kind, std::vector<BTile>& out) {
  const int brows = (int)brp.size() - 1;
  for (int b = 0; b < brows; b += kWG / kWave)
    out.push_back(BTile{m, kind, b, std::min(brows, b + kWG / kWave), brp[b],
                        brp[std::min(brows, b + kWG / kWave)], 0, 0});
}

// Tiles of a DASP layout: its fused launch's workgroups, a long row alone,
// row blocks and short rows kDaspPerTile workgroups at a time; p0/p1 carry
// the entries a tile covers (the weight), counted from the long-row and
// row-block pointers (short rows: at most 4 entries per row).
void dasp_tiles(const dasp::DaspArgs& a, const std::vector<int32_t>& lrpt,
                const std::vector<int32_t>& bptr, int m, std::vector<BTile>& out) {
  for (int i = 0; i < (int)a.wg_long; ++i)
    out.push_back(BTile{m, kKindDasp, i, i + 1, 0, (lrpt[i + 1] - lrpt[i]) * 64, 0, 0});
  const int r0 = (int)a.wg_long, r1 = r0 + (int)a.wg_reg, v1 = r1 + (int)a.wg_short;
  constexpr int kBlocksPerWg = dasp::kTB / kWave;
  for (int v = r0; v < r1; v += kDaspPerTile) {
    const int e = std::min(r1, v + kDaspPerTile);
    const int b0 = (v - r0) * kBlocksPerWg, b1 = std::min(a.blocknum, (e - r0) * kBlocksPerWg);
    out.push_back(BTile{m, kKindDasp, v, e, 0, bptr[b1] - bptr[b0] + (b1 - b0) * 8, 0, 0});
  }
  for (int v = r1; v < v1; v += kDaspPerTile) {
    const int e = std::min(v1, v + kDaspPerTile);
    out.push_back(BTile{m, kKindDasp, v, e, 0, (e - v) * dasp::kTB * 4, 0, 0});
  }
}

// Tiles of a diagonal-layout matrix: kWG rows each (a thread per row); p1
// carries the slots a tile reads (rows x diagonals), its weight.
void dia_tiles(int n, int ndiag, int m, std::vector<BTile>& out) {
  for (int r = 0; r < n; r += kWG) {
    const int e = std::min(n, r + kWG);
    out.push_back(BTile{m, kKindDia, r, e, 0, (e - r) * ndiag, 0, 0});
  }
}

int64_t tile_weight(const BTile& t) {
  if (t.kind == kKindCsr) return (int64_t)(t.r1 - t.r0) + (t.p1 - t.p0) + kTileOverhead;
  if (t.kind == kKindDasp || t.kind == kKindDia) return (int64_t)(t.p1 - t.p0) + kTileOverhead;
  return (int64_t)(t.p1 - t.p0) * (t.kind == kKindBsr4 ? 16 : 64) + kTileOverhead;
}

// Contiguous per-workgroup tile ranges with balanced weight.
std::vector<int32_t> wg_cuts(const std::vector<BTile>& tiles, int grid) {
  std::vector<int64_t> pre(tiles.size() + 1, 0);
  for (size_t i = 0; i < tiles.size(); ++i) pre[i + 1] = pre[i] + tile_weight(tiles[i]);
  std::vector<int32_t> cut(grid + 1);
  const int64_t total = pre.back();
  for (int w = 0; w <= grid; ++w) {
    const int64_t target = total * w / grid;
    cut[w] = (int32_t)(std::lower_bound(pre.begin(), pre.end(), target) - pre.begin());
  }
  cut[0] = 0;
  cut[grid] = (int32_t)tiles.size();
  return cut;
}

struct Schedule {
  DevBuf<BTile> tiles;
  DevBuf<int32_t> cuts;
  int64_t ntiles = 0;
  int grid = 0;
  int resident = kMaxGrid;  // workgroups the device holds at once for the kernel
  int grid_cap = 0;         // workgroups per launch (0: kOversub * resident)
  bool has_dasp = false;    // some tiles are DASP workgroups
  bsp_status upload(const std::vector<BTile>& t, hipStream_t s) {
    ntiles = (int64_t)t.size();
    has_dasp = std::any_of(t.begin(), t.end(), [](const BTile& b) { return b.kind == kKindDasp; });
    const int64_t cap = grid_cap > 0 ? grid_cap : (int64_t)kOversub * resident;
    grid = (int)std::max<int64_t>(1, std::min<int64_t>(cap, ntiles));
    std::vector<int32_t> c = wg_cuts(t, grid);
    BSP_TRY(tiles.alloc(t.size()));
    BSP_TRY(cuts.alloc(c.size()));
    if (!t.empty())
      BSP_HIP(hipMemcpyAsync(tiles.get(), t.data(), sizeof(BTile) * t.size(), hipMemcpyHostToDevice, s));
    BSP_HIP(hipMemcpyAsync(cuts.get(), c.data(), sizeof(int32_t) * c.size(), hipMemcpyHostToDevice, s));
    BSP_HIP(hipStreamSynchronize(s));
    return BSP_OK;
  }
};

}  // namespace
}  // namespace bsp

using namespace bsp;

struct bsp_batch_s {
  std::vector<bsp_csr> A;
  std::vector<int64_t> row_off, col_off;
  std::vector<int32_t> fmt, win_spmv, win_spmm;
  std::vector<int32_t> fmt_spmm;   // BSP_BATCH_CSR or BSP_BATCH_DIA per matrix
  std::vector<bsp_bsr> bsr, bsr8;  // owned 4x4 / 8x8 BSR per matrix (null unless chosen)
  std::vector<bsp_dasp> dasp;      // owned DASP layout per matrix (null unless chosen)
  std::vector<bsp_dia> dia;        // owned diagonal layout per matrix (null unless chosen)
  std::vector<DevBuf<uint16_t>> dia_h;  // its FP16 values (the SpMM's)
  DevBuf<dasp::DaspArgs> dargs;    // their SpMV arguments (x, y unset)
  DevBuf<uint16_t> hval;     // FP16 copies of every matrix's values
  DevBuf<BMat> mats;
  Schedule spmv, spmm;
  hipStream_t cap = nullptr;  // capture stream
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // The SpMM of a step runs beside the SpMV: forked onto `side` after an
  // event on the caller's stream and joined back with another (default
  // priority, like every stream here).
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  ~bsp_batch_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    if (side) (void)hipStreamDestroy(side);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    for (bsp_bsr h : bsr)
      if (h) bsp_bsr_destroy(h);
    for (bsp_bsr h : bsr8)
      if (h) bsp_bsr_destroy(h);
    for (bsp_dasp h : dasp)
      if (h) bsp_dasp_destroy(h);
    for (bsp_dia h : dia)
      if (h) bsp_dia_destroy(h);
  }
};

namespace {

bsp_status upload_mats(bsp_batch_s* b, hipStream_t s) {
  const size_t M = b->A.size();
  std::vector<BMat> h(M);
  std::vector<dasp::DaspArgs> da(M);
  if (!b->dargs.get()) BSP_TRY(b->dargs.alloc(M));
  int64_t hoff = 0;
  for (size_t m = 0; m < M; ++m) {
    const bsp_csr& A = b->A[m];
    BMat d{};
    d.rowptr = A.rowptr;
    d.col = A.colidx;
    d.val = static_cast<const double*>(A.val);
    d.hval = b->hval.get() + hoff;
    d.xoff = b->col_off[m];
    d.yoff = b->row_off[m];
    d.nrows = (int32_t)A.nrows;
    d.ncols = (int32_t)A.ncols;
    if (b->bsr[m]) {
      bsp_bsr_view v;
      bsp_bsr_view_get(b->bsr[m], &v);
      d.brp = v.rowptr;
      d.bci = v.colidx;
      d.bval = v.val;
    }
    if (b->bsr8[m]) {
      bsp_bsr_view v;
      bsp_bsr_view_get(b->bsr8[m], &v);
      d.brp8 = v.rowptr;
      d.bci8 = v.colidx;
      d.bval8 = v.val;
    }
    if (b->dasp[m]) {
      bsp_dasp_view v;
      bsp_dasp_view_get(b->dasp[m], &v);
      da[m] = dasp::dasp_args(v, nullptr, nullptr, 0);
      d.dasp = b->dargs.get() + m;
    }
    if (b->dia[m]) {
      const bsp_dia_s* g = b->dia[m];
      d.dval = g->val.get();
      d.dhval = b->dia_h[m].get();
      d.dmask = g->mask.get();
      d.ndiag = g->ndiag;
      for (int k = 0; k < g->ndiag && k < kBatchDia; ++k) d.doff[k] = g->off.d[k];
    }
    h[m] = d;
    hoff += A.nnz;
  }
  BSP_TRY(b->mats.alloc(M));
  BSP_HIP(hipMemcpyAsync(b->mats.get(), h.data(), sizeof(BMat) * M, hipMemcpyHostToDevice, s));
  BSP_HIP(hipMemcpyAsync(b->dargs.get(), da.data(), sizeof(dasp::DaspArgs) * M,
                         hipMemcpyHostToDevice, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

void launch_spmv(const Schedule& S, const BMat* mats, const double* x, double* y, hipStream_t s) {
  if (S.ntiles == 0) return;
  if (S.has_dasp)
    batch_spmv_f64<true><<<S.grid, kWG, 0, s>>>(S.tiles.get(), S.cuts.get(), mats, x, y);
  else
    batch_spmv_f64<false><<<S.grid, kWG, 0, s>>>(S.tiles.get(), S.cuts.get(), mats, x, y);
}

void launch_spmm(const Schedule& S, const BMat* mats, const uint16_t* X, float* Y, hipStream_t s) {
  if (S.ntiles)
    batch_spmm8_f16<<<S.grid, kWG, 0, s>>>(S.tiles.get(), S.cuts.get(), mats,
                                           reinterpret_cast<const uint4*>(X), Y);
}

// Workgroups of `kernel` resident on the whole device at once (one full
// wave of the persistent grid), capped at kMaxGrid.
template <class K>
bsp_status resident_grid(K kernel, int* out) {
  int dev = 0, cus = 0, per = 0;
  BSP_HIP(hipGetDevice(&dev));
  BSP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  BSP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kWG, 0));
  *out = std::max(1, std::min(kMaxGrid, cus * std::max(per, 1)));
  return BSP_OK;
}

// Time of one persistent launch over `reps` copies of one matrix's tiles
// (enough copies to give every workgroup of the grid work): the matrix's
// throughput cost per copy inside a full launch.
bsp_status time_candidate(const std::vector<BTile>& one, bool spmm, int resident,
                          const BMat* mats, const double* x, double* y, const uint16_t* X, float* Y,
                          hipStream_t s, hipEvent_t e0, hipEvent_t e1, float* per_copy) {
  const int64_t reps = std::max<int64_t>(1, cdiv(2 * resident, (int64_t)one.size()));
  std::vector<BTile> t;
  t.reserve(one.size() * reps);
  for (int64_t r = 0; r < reps; ++r) t.insert(t.end(), one.begin(), one.end());
  Schedule S;
  S.resident = resident;
  BSP_TRY(S.upload(t, s));
  float best = 1e30f;
  for (int it = 0; it < 4; ++it) {
    BSP_HIP(hipEventRecord(e0, s));
    if (spmm)
      launch_spmm(S, mats, X, Y, s);
    else
      launch_spmv(S, mats, x, y, s);
    BSP_HIP(hipEventRecord(e1, s));
    BSP_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BSP_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (it > 0) best = std::min(best, ms);
  }
  BSP_HIP(hipGetLastError());
  *per_copy = best / (float)reps;
  return BSP_OK;
}

// Host copy of a BSR's block-row pointers.
bsp_status bsr_rowptr(bsp_bsr h, hipStream_t s, std::vector<int32_t>& brp) {
  bsp_bsr_view v;
  BSP_TRY(bsp_bsr_view_get(h, &v));
  brp.resize(v.brows + 1);
  BSP_HIP(hipMemcpyAsync(brp.data(), v.rowptr, sizeof(int32_t) * brp.size(),
                         hipMemcpyDeviceToHost, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

// DASP tiles of matrix m (host copies of its long-row and row-block pointers).
bsp_status dasp_matrix_tiles(bsp_dasp h, int m, hipStream_t s, std::vector<BTile>& out) {
  bsp_dasp_view v;
  BSP_TRY(bsp_dasp_view_get(h, &v));
  const dasp::DaspArgs a = dasp::dasp_args(v, nullptr, nullptr, 0);
  std::vector<int32_t> lrpt(v.row_long + 1), bptr(v.blocknum + 1);
  BSP_HIP(hipMemcpyAsync(lrpt.data(), v.long_rpt, sizeof(int32_t) * lrpt.size(),
                         hipMemcpyDeviceToHost, s));
  BSP_HIP(hipMemcpyAsync(bptr.data(), v.block_ptr, sizeof(int32_t) * bptr.size(),
                         hipMemcpyDeviceToHost, s));
  BSP_HIP(hipStreamSynchronize(s));
  dasp_tiles(a, lrpt, bptr, m, out);
  return BSP_OK;
}

// Tiles of every matrix for one operation: with uniform > 0 every matrix in
// CSR with that window; otherwise the per-matrix windows (and, for the SpMV,
// formats).
bsp_status op_tiles(const bsp_batch_s* b, const std::vector<std::vector<int32_t>>& rp, bool spmm,
                    int uniform, hipStream_t s, std::vector<BTile>& out) {
  out.clear();
  for (size_t m = 0; m < b->A.size(); ++m) {
    if (b->A[m].nrows == 0) continue;
    const int f = b->fmt[m];
    if (!spmm && uniform == 0 && (f == BSP_BATCH_BSR4 || f == BSP_BATCH_BSR8)) {
      std::vector<int32_t> brp;
      BSP_TRY(bsr_rowptr(f == BSP_BATCH_BSR4 ? b->bsr[m] : b->bsr8[m], s, brp));
      bsr_tiles(brp, (int)m, f == BSP_BATCH_BSR4 ? kKindBsr4 : kKindBsr8, out);
    } else if (!spmm && uniform == 0 && f == BSP_BATCH_DASP) {
      BSP_TRY(dasp_matrix_tiles(b->dasp[m], (int)m, s, out));
    } else if (uniform == 0 && (spmm ? b->fmt_spmm[m] : f) == BSP_BATCH_DIA) {
      dia_tiles((int)b->A[m].nrows, b->dia[m]->ndiag, (int)m, out);
    } else {
      const int c = uniform > 0 ? uniform : spmm ? b->win_spmm[m] : b->win_spmv[m];
      csr_tiles(rp[m], (int)m, c, out);
    }
  }
  return BSP_OK;
}

// Best of five launches (after one) of a whole operation's schedule.
bsp_status time_schedule(const Schedule& S, bool spmm, const BMat* mats, const double* x,
                         double* y, const uint16_t* X, float* Y, hipStream_t s, hipEvent_t e0,
                         hipEvent_t e1, float* ms_out) {
  float best = 1e30f;
  for (int it = 0; it < 6; ++it) {
    BSP_HIP(hipEventRecord(e0, s));
    if (spmm)
      launch_spmm(S, mats, X, Y, s);
    else
      launch_spmv(S, mats, x, y, s);
    BSP_HIP(hipEventRecord(e1, s));
    BSP_HIP(hipEventSynchronize(e1));
    float ms = 0;
    BSP_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (it > 0) best = std::min(best, ms);
  }
  BSP_HIP(hipGetLastError());
  *ms_out = best;
  return BSP_OK;
}

// The step itself decides, since the per-matrix timings are cache-warm
// proxies: each operation is timed over the whole batch with every uniform
// window and with the per-matrix choice, each at launch sizes of 1, 4 and 16
// resident waves and one workgroup per tile (or at the caller's grid).
// *uniform_out gets the fastest window (0: the per-matrix choice), *grid_out
// its launch size.
bsp_status pick_schedule(const bsp_batch_s* b, const std::vector<std::vector<int32_t>>& rp,
                         bool spmm, int resident, int grid_opt, const double* x, double* y,
                         const uint16_t* X, float* Y, hipStream_t s, hipEvent_t e0, hipEvent_t e1,
                         int* uniform_out, int* grid_out) {
  float best = 1e30f;
  *uniform_out = spmm ? kCapM : kCap;
  *grid_out = grid_opt;
  std::vector<int> cands;
  for (int c : kWindows)
    if (c <= (spmm ? kCapM : kCap)) cands.push_back(c);
  cands.push_back(0);
  for (int u : cands) {
    std::vector<BTile> t;
    BSP_TRY(op_tiles(b, rp, spmm, u, s, t));
    std::vector<int> grids;
    if (grid_opt > 0)
      grids = {grid_opt};
    else
      grids = {resident, 4 * resident, kOversub * resident, std::max<int>(1, (int)t.size())};
    for (int g : grids) {
      Schedule S;
      S.resident = resident;
      S.grid_cap = g;
      BSP_TRY(S.upload(t, s));
      float ms = 0;
      BSP_TRY(time_schedule(S, spmm, b->mats.get(), x, y, X, Y, s, e0, e1, &ms));
      if (ms < best) {
        best = ms;
        *uniform_out = u;
        *grid_out = g;
      }
    }
  }
  return BSP_OK;
}

}  // namespace

extern "C" {

bsp_status bsp_batch_create(const bsp_csr* mats, int64_t count, const bsp_batch_options* opt,
                            bsp_stream stream, bsp_batch* out) {
  BSP_REQUIRE(mats && out && count > 0 && count < (int64_t(1) << 31), "bad batch arguments");
  hipStream_t s = (hipStream_t)stream;
  const int tune = opt ? opt->tune : 0;
  const int grid = opt ? opt->grid : 0;
  BSP_REQUIRE(tune >= 0 && tune <= 4, "tune must be 0, 1, 2, 3 or 4");
  BSP_REQUIRE(grid >= 0, "grid must be >= 0");
  auto b = std::make_unique<bsp_batch_s>();
  const size_t M = (size_t)count;
  b->A.assign(mats, mats + M);
  b->row_off.assign(M + 1, 0);
  b->col_off.assign(M + 1, 0);
  int64_t total_nnz = 0, max_rows = 0, max_cols = 0;
  for (size_t m = 0; m < M; ++m) {
    const bsp_csr& A = b->A[m];
    BSP_TRY(check_csr_pattern(&A));
    BSP_REQUIRE(A.dtype == BSP_F64, "batch matrices need FP64 values");
    BSP_REQUIRE(A.nnz == 0 || A.val, "null values");
    b->row_off[m + 1] = b->row_off[m] + A.nrows;
    b->col_off[m + 1] = b->col_off[m] + A.ncols;
    total_nnz += A.nnz;
    max_rows = std::max(max_rows, A.nrows);
    max_cols = std::max(max_cols, A.ncols);
  }
  if (!sizes_fit(b->row_off[M], b->col_off[M], 0))
    return fail(BSP_ERR_TOO_LARGE, "batch total rows/columns must fit int32 (at most 2^31 - 2^20)");
  b->fmt.assign(M, BSP_BATCH_CSR);
  b->win_spmv.assign(M, kCap);
  b->win_spmm.assign(M, kCapM);
  b->bsr.assign(M, nullptr);
  b->bsr8.assign(M, nullptr);
  b->dasp.assign(M, nullptr);
  b->fmt_spmm.assign(M, BSP_BATCH_CSR);
  b->dia.assign(M, nullptr);
  b->dia_h.resize(M);
  BSP_HIP(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking));
  BSP_HIP(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming));
  BSP_HIP(hipEventCreateWithFlags(&b->join, hipEventDisableTiming));
  BSP_TRY(resident_grid(batch_spmv_f64<false>, &b->spmv.resident));
  BSP_TRY(resident_grid(batch_spmm8_f16, &b->spmm.resident));
  b->spmv.grid_cap = b->spmm.grid_cap = grid;

  // FP16 values (round to nearest even) for the SpMM
  BSP_TRY(b->hval.alloc(total_nnz));
  {
    int64_t off = 0;
    for (size_t m = 0; m < M; ++m) {
      const bsp_csr& A = b->A[m];
      if (A.nnz)
        f64_to_f16<<<cdiv(A.nnz, kWG), kWG, 0, s>>>(static_cast<const double*>(A.val), A.nnz,
                                                     b->hval.get() + off);
      off += A.nnz;
    }
    BSP_HIP(hipGetLastError());
  }
  // host copies of the row pointers (the schedule is cut on the host)
  std::vector<std::vector<int32_t>> rp(M);
  for (size_t m = 0; m < M; ++m) {
    rp[m].resize(b->A[m].nrows + 1);
    if (b->A[m].nrows)
      BSP_HIP(hipMemcpyAsync(rp[m].data(), b->A[m].rowptr, sizeof(int32_t) * rp[m].size(),
                             hipMemcpyDeviceToHost, s));
    else
      rp[m][0] = 0;
  }
  BSP_HIP(hipStreamSynchronize(s));

  if (tune == 1 || tune == 2) {
    // block sizes 4 and 8: a BSR is built only where its blocks are filled
    // enough (kBsrFill); block statistics first (key sort only), so most
    // matrices are screened out without building anything
    for (size_t m = 0; m < M; ++m) {
      const bsp_csr& A = b->A[m];
      if (A.nnz < 16 * 64) continue;
      for (const int bs : {4, 8}) {
        bsp_blockinfo bi;
        BSP_TRY(bsp_block_info(&A, bs, stream, &bi));
        if (bi.avg < kBsrFill * bs * bs) continue;
        bsp_bsr h = nullptr;
        BSP_TRY(bsp_bsr_build(&A, bs, bs, stream, &h));
        (bs == 4 ? b->bsr : b->bsr8)[m] = h;
      }
    }
  }
  // DASP (threshold 0.75, block_longest 256, the reference's values): a
  // candidate of every matrix of at least 1024 entries (tune = 1), or the
  // format of every matrix with rows (tune = 3, no timing)
  if (tune == 1 || tune == 3)
    for (size_t m = 0; m < M; ++m) {
      const bsp_csr& A = b->A[m];
      if (A.nrows == 0 || (tune == 1 && A.nnz < 16 * 64)) continue;
      BSP_TRY(bsp_dasp_build(&A, 0.75, 256, stream, &b->dasp[m]));
      if (tune == 3) b->fmt[m] = BSP_BATCH_DASP;
    }
  // The diagonal layout: a candidate (tune = 1), or the format of both
  // operations (tune = 4, no timing), of every matrix with at most kBatchDia
  // diagonals and ascending columns (the builder refuses the others).
  if (tune == 1 || tune == 4)
    for (size_t m = 0; m < M; ++m) {
      const bsp_csr& A = b->A[m];
      if (A.nrows == 0 || A.nnz == 0 || A.nnz > (int64_t)kBatchDia * A.nrows) continue;
      bsp_dia h = nullptr;
      const bsp_status st = bsp_dia_build(&A, stream, &h);
      if (st == BSP_ERR_UNSUPPORTED || st == BSP_ERR_UNSORTED) continue;
      BSP_TRY(st);
      if (h->ndiag > kBatchDia) {
        bsp_dia_destroy(h);
        continue;
      }
      b->dia[m] = h;
      const int64_t slots = (int64_t)h->ndiag * A.nrows;
      BSP_TRY(b->dia_h[m].alloc(slots));
      f64_to_f16<<<cdiv(slots, kWG), kWG, 0, s>>>(h->val.get(), slots, b->dia_h[m].get());
      BSP_HIP(hipGetLastError());
      if (tune == 4) b->fmt[m] = b->fmt_spmm[m] = BSP_BATCH_DIA;
    }
  if (tune == 2)  // every eligible matrix in BSR (8x8 where eligible), no timing
    for (size_t m = 0; m < M; ++m) {
      if (b->bsr8[m]) {
        b->fmt[m] = BSP_BATCH_BSR8;
        if (b->bsr[m]) bsp_bsr_destroy(b->bsr[m]);
        b->bsr[m] = nullptr;
      } else if (b->bsr[m]) {
        b->fmt[m] = BSP_BATCH_BSR4;
      }
    }
  if (tune == 1) {
    BSP_TRY(upload_mats(b.get(), s));
    DevBuf<double> x, y;
    DevBuf<uint16_t> X;
    DevBuf<float> Y;
    BSP_TRY(x.alloc(b->col_off[M]));
    BSP_TRY(y.alloc(b->row_off[M]));
    BSP_TRY(X.alloc(b->col_off[M] * kCols));
    BSP_TRY(Y.alloc(b->row_off[M] * kCols));
    BSP_HIP(hipMemsetAsync(x.get(), 0, sizeof(double) * std::max<int64_t>(1, b->col_off[M]), s));
    BSP_HIP(hipMemsetAsync(X.get(), 0, sizeof(uint16_t) * std::max<int64_t>(1, b->col_off[M] * kCols), s));
    EventPair ev;
    BSP_TRY(ev.create());
    hipEvent_t e0 = ev.e0, e1 = ev.e1;
    bsp_status st = BSP_OK;
    for (size_t m = 0; m < M && st == BSP_OK; ++m) {
      if (b->A[m].nrows == 0) continue;
      // The timing is cache-warm (the copies share one matrix), so it ranks
      // per-tile cost only: a candidate replaces the default window only
      // when it is at least kMargin faster than the default.
      constexpr float kMargin = 0.9f;
      float best_v = 1e30f, best_m = 1e30f, def_v = 1e30f, def_m = 1e30f;
      int win_v = kCap, win_m = kCapM;
      for (int c : kWindows) {
        std::vector<BTile> one;
        csr_tiles(rp[m], (int)m, c, one);
        float t = 0;
        st = time_candidate(one, false, b->spmv.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st != BSP_OK) break;
        if (c == kCap) def_v = t;
        if (t < best_v) best_v = t, win_v = c;
        if (c > kCapM) continue;
        st = time_candidate(one, true, b->spmm.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st != BSP_OK) break;
        if (c == kCapM) def_m = t;
        if (t < best_m) best_m = t, win_m = c;
      }
      b->fmt[m] = BSP_BATCH_CSR;
      b->win_spmv[m] = best_v < kMargin * def_v ? win_v : kCap;
      b->win_spmm[m] = best_m < kMargin * def_m ? win_m : kCapM;
      best_v = b->win_spmv[m] == kCap ? def_v : best_v;
      best_m = b->win_spmm[m] == kCapM ? def_m : best_m;
      // each BSR candidate must beat the best so far by the margin
      for (const int f : {BSP_BATCH_BSR4, BSP_BATCH_BSR8}) {
        const bsp_bsr h = f == BSP_BATCH_BSR4 ? b->bsr[m] : b->bsr8[m];
        if (st != BSP_OK || !h) continue;
        std::vector<int32_t> brp;
        st = bsr_rowptr(h, s, brp);
        if (st != BSP_OK) break;
        std::vector<BTile> one;
        bsr_tiles(brp, (int)m, f == BSP_BATCH_BSR4 ? kKindBsr4 : kKindBsr8, one);
        float t = 0;
        st = time_candidate(one, false, b->spmv.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st == BSP_OK && t < kMargin * best_v) best_v = t, b->fmt[m] = f;
      }
      if (st == BSP_OK && b->dasp[m]) {
        std::vector<BTile> one;
        st = dasp_matrix_tiles(b->dasp[m], (int)m, s, one);
        float t = 0;
        if (st == BSP_OK)
          st = time_candidate(one, false, b->spmv.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st == BSP_OK && t < kMargin * best_v) best_v = t, b->fmt[m] = BSP_BATCH_DASP;
      }
      // the diagonal layout, for each operation against its best so far
      if (st == BSP_OK && b->dia[m]) {
        std::vector<BTile> one;
        dia_tiles((int)b->A[m].nrows, b->dia[m]->ndiag, (int)m, one);
        float t = 0;
        st = time_candidate(one, false, b->spmv.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st == BSP_OK && t < kMargin * best_v) best_v = t, b->fmt[m] = BSP_BATCH_DIA;
        if (st == BSP_OK)
          st = time_candidate(one, true, b->spmm.resident, b->mats.get(), x.get(), y.get(), X.get(), Y.get(), s, e0, e1, &t);
        if (st == BSP_OK && t < kMargin * best_m) best_m = t, b->fmt_spmm[m] = BSP_BATCH_DIA;
      }
      // only the chosen format's layout is kept (decided after DASP's turn)
      if (st == BSP_OK && b->fmt[m] != BSP_BATCH_BSR4 && b->bsr[m]) {
        bsp_bsr_destroy(b->bsr[m]);
        b->bsr[m] = nullptr;
      }
      if (st == BSP_OK && b->fmt[m] != BSP_BATCH_BSR8 && b->bsr8[m]) {
        bsp_bsr_destroy(b->bsr8[m]);
        b->bsr8[m] = nullptr;
      }
      if (st == BSP_OK && b->fmt[m] != BSP_BATCH_DASP && b->dasp[m]) {
        bsp_dasp_destroy(b->dasp[m]);
        b->dasp[m] = nullptr;
      }
      if (st == BSP_OK && b->fmt[m] != BSP_BATCH_DIA && b->fmt_spmm[m] != BSP_BATCH_DIA &&
          b->dia[m]) {
        bsp_dia_destroy(b->dia[m]);
        b->dia[m] = nullptr;
        b->dia_h[m].reset();
      }
    }
    if (st == BSP_OK) st = upload_mats(b.get(), s);
    int uni_v = 0, uni_m = 0, grid_v = grid, grid_m = grid;
    if (st == BSP_OK)
      st = pick_schedule(b.get(), rp, false, b->spmv.resident, grid, x.get(), y.get(), X.get(),
                         Y.get(), s, e0, e1, &uni_v, &grid_v);
    if (st == BSP_OK)
      st = pick_schedule(b.get(), rp, true, b->spmm.resident, grid, x.get(), y.get(), X.get(),
                         Y.get(), s, e0, e1, &uni_m, &grid_m);
    BSP_TRY(st);
    b->spmv.grid_cap = grid_v;
    b->spmm.grid_cap = grid_m;
    if (uni_v > 0) {  // a uniform SpMV window won: every matrix in CSR with it
      for (size_t m = 0; m < M; ++m) {
        b->fmt[m] = BSP_BATCH_CSR;
        b->win_spmv[m] = uni_v;
        if (b->bsr[m]) bsp_bsr_destroy(b->bsr[m]);
        if (b->bsr8[m]) bsp_bsr_destroy(b->bsr8[m]);
        if (b->dasp[m]) bsp_dasp_destroy(b->dasp[m]);
        b->bsr[m] = b->bsr8[m] = nullptr;
        b->dasp[m] = nullptr;
      }
    }
    if (uni_m > 0) {  // a uniform SpMM window won: every matrix's SpMM in CSR with it
      std::fill(b->win_spmm.begin(), b->win_spmm.end(), uni_m);
      std::fill(b->fmt_spmm.begin(), b->fmt_spmm.end(), BSP_BATCH_CSR);
    }
    for (size_t m = 0; m < M; ++m)  // diagonal layouts neither operation kept
      if (b->dia[m] && b->fmt[m] != BSP_BATCH_DIA && b->fmt_spmm[m] != BSP_BATCH_DIA) {
        bsp_dia_destroy(b->dia[m]);
        b->dia[m] = nullptr;
        b->dia_h[m].reset();
      }
  }
  BSP_TRY(upload_mats(b.get(), s));

  std::vector<BTile> tv, tm;
  BSP_TRY(op_tiles(b.get(), rp, false, 0, s, tv));
  BSP_TRY(op_tiles(b.get(), rp, true, 0, s, tm));
  BSP_TRY(b->spmv.upload(tv, s));
  BSP_TRY(b->spmm.upload(tm, s));
  *out = b.release();
  return BSP_OK;
}

bsp_status bsp_batch_layout(bsp_batch b, int64_t* row_off, int64_t* col_off) {
  BSP_REQUIRE(b, "null batch");
  if (row_off) std::copy(b->row_off.begin(), b->row_off.end(), row_off);
  if (col_off) std::copy(b->col_off.begin(), b->col_off.end(), col_off);
  return BSP_OK;
}

bsp_status bsp_batch_choice(bsp_batch b, int32_t* format, int32_t* spmv_window,
                            int32_t* spmm_window) {
  BSP_REQUIRE(b, "null batch");
  if (format) std::copy(b->fmt.begin(), b->fmt.end(), format);
  if (spmv_window) std::copy(b->win_spmv.begin(), b->win_spmv.end(), spmv_window);
  if (spmm_window) std::copy(b->win_spmm.begin(), b->win_spmm.end(), spmm_window);
  return BSP_OK;
}

bsp_status bsp_batch_spmm_format(bsp_batch b, int32_t* format) {
  BSP_REQUIRE(b && format, "null argument");
  std::copy(b->fmt_spmm.begin(), b->fmt_spmm.end(), format);
  return BSP_OK;
}

bsp_status bsp_batch_info(bsp_batch b, int64_t* spmv_tiles, int64_t* spmm_tiles, int64_t* grid) {
  BSP_REQUIRE(b, "null batch");
  if (spmv_tiles) *spmv_tiles = b->spmv.ntiles;
  if (spmm_tiles) *spmm_tiles = b->spmm.ntiles;
  if (grid) *grid = std::max(b->spmv.grid, b->spmm.grid);
  return BSP_OK;
}

bsp_status bsp_batch_run(bsp_batch b, int schedule, const double* x, double* y, const uint16_t* X,
                         float* Y, bsp_stream stream) {
  BSP_REQUIRE(b, "null batch");
  BSP_REQUIRE((x == nullptr) == (y == nullptr) && (X == nullptr) == (Y == nullptr),
              "x/y and X/Y come in pairs");
  hipStream_t s = (hipStream_t)stream;
  if (schedule == BSP_BATCH_PERSISTENT) {
    if (x && X) {  // independent operations: the SpMM on the side stream
      BSP_HIP(hipEventRecord(b->fork, s));
      BSP_HIP(hipStreamWaitEvent(b->side, b->fork, 0));
      launch_spmm(b->spmm, b->mats.get(), X, Y, b->side);
      launch_spmv(b->spmv, b->mats.get(), x, y, s);
      BSP_HIP(hipEventRecord(b->join, b->side));
      BSP_HIP(hipStreamWaitEvent(s, b->join, 0));
    } else {
      if (x) launch_spmv(b->spmv, b->mats.get(), x, y, s);
      if (X) launch_spmm(b->spmm, b->mats.get(), X, Y, s);
    }
  } else if (schedule == BSP_BATCH_PLAIN) {
    int64_t hoff = 0;
    for (size_t m = 0; m < b->A.size(); ++m) {
      const bsp_csr& A = b->A[m];
      const int n = (int)A.nrows;
      if (n > 0) {
        if (x)
          plain_spmv_f64<<<cdiv(n, kWG), kWG, 0, s>>>(A.rowptr, A.colidx,
                                                      static_cast<const double*>(A.val),
                                                      x + b->col_off[m], y + b->row_off[m], n);
        if (X)
          plain_spmm8_f16<<<cdiv(n, kWG), kWG, 0, s>>>(
              A.rowptr, A.colidx, b->hval.get() + hoff,
              reinterpret_cast<const uint4*>(X) + b->col_off[m], Y + b->row_off[m] * kCols, n);
      }
      hoff += A.nnz;
    }
  } else {
    return fail(BSP_ERR_INVALID, "unknown batch schedule");
  }
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_batch_graph_capture(bsp_batch b, int schedule, const double* x, double* y,
                                   const uint16_t* X, float* Y) {
  BSP_REQUIRE(b, "null batch");
  if (b->exec) {
    BSP_HIP(hipGraphExecDestroy(b->exec));
    b->exec = nullptr;
  }
  if (b->graph) {
    BSP_HIP(hipGraphDestroy(b->graph));
    b->graph = nullptr;
  }
  if (!b->cap) BSP_HIP(hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking));
  BSP_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
  bsp_status st = bsp_batch_run(b, schedule, x, y, X, Y, b->cap);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->cap, &g);
  if (st != BSP_OK || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (st != BSP_OK) return st;
    return fail(BSP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  }
  b->graph = g;
  BSP_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  return BSP_OK;
}

bsp_status bsp_batch_graph_launch(bsp_batch b, bsp_stream stream) {
  BSP_REQUIRE(b && b->exec, "no captured graph");
  BSP_HIP(hipGraphLaunch(b->exec, (hipStream_t)stream));
  return BSP_OK;
}

bsp_status bsp_batch_destroy(bsp_batch b) {
  delete b;
  return BSP_OK;
}

}  // extern "C"

// Block CSR (tm x tk) and DBSR: GPU builders and FP64 SpMV kernels.
#include <cmath>
#include <memory>
#include <vector>

#include "bsp_hip.hpp"

struct bsp_bsr_s {
  bsp_bsr_view v{};
  bsp::DevBuf<int32_t> rowptr, colidx, rpw, riw;
  bsp::DevBuf<double> val;
  bsp::DevBuf<uint16_t> map;
  // balanced 4x4 kernels: block rows whose chunk count is not one (empty
  // ones, and those split over several 64-block chunks), and per-chunk
  // partial sums (4 per chunk) that the combine pass adds in chunk order
  bsp::DevBuf<int32_t> multi;
  int64_t nmulti = 0;
  bsp::DevBuf<double> part;
  bsp::DevBuf<int4> ctab;  // per chunk {block row, first block, blocks, single}
};

struct bsp_dbsr_s {
  bsp_dbsr_view v{};
  bsp::DevBuf<int32_t> recorder, blcptr, colidx, prefix;
  bsp::DevBuf<double> val;
  bsp::DevBuf<uint16_t> map;
  int64_t narrow = 0;  // 4-row groups at most 8 blocks wide (a prefix: rows sorted by length)
};

namespace bsp {
namespace {

constexpr int kTB = 256;

__global__ void row_ids(const int32_t* __restrict__ rowptr, int64_t nrows, int32_t* __restrict__ rid) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) rid[j] = (int32_t)r;
}

__global__ void bsr_keys(const int32_t* __restrict__ rid, const int32_t* __restrict__ col,
                         int64_t nnz, int tm, int tk, int64_t bcols, int64_t* __restrict__ key,
                         int32_t* __restrict__ idx) {
  int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  key[j] = (int64_t)(rid[j] / tm) * bcols + col[j] / tk;
  idx[j] = (int32_t)j;
}

__global__ void key_flags(const int64_t* __restrict__ skey, int64_t n, int32_t* __restrict__ flag) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  flag[i] = (i == 0 || skey[i] != skey[i - 1]) ? 1 : 0;
}

__global__ void bsr_blocks(const int64_t* __restrict__ skey, const int32_t* __restrict__ flag,
                           const int32_t* __restrict__ pos, int64_t n, int64_t bcols,
                           int32_t* __restrict__ colidx, int32_t* __restrict__ bstart,
                           int32_t* __restrict__ rowcnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int b = pos[i];
  colidx[b] = (int32_t)(skey[i] % bcols);
  bstart[b] = (int32_t)i;
  atomicAdd(&rowcnt[skey[i] / bcols], 1);
}

__global__ void bsr_values(const int32_t* __restrict__ sidx, const int32_t* __restrict__ flag,
                           const int32_t* __restrict__ pos, int64_t n,
                           const int32_t* __restrict__ rid, const int32_t* __restrict__ col,
                           const double* __restrict__ val, int tm, int tk,
                           double* __restrict__ bval) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = sidx[i];
  const int64_t b = pos[i] + flag[i] - 1;
  bval[b * tm * tk + (rid[j] % tm) * tk + col[j] % tk] = val[j];
}

__global__ void bsr_map4(const int32_t* __restrict__ bstart, int64_t nblocks, int64_t n,
                         const int32_t* __restrict__ sidx, const int32_t* __restrict__ rid,
                         const int32_t* __restrict__ col, uint16_t* __restrict__ map) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nblocks) return;
  if (b == nblocks) {
    map[b] = 0;
    return;
  }
  const int e = b + 1 < nblocks ? bstart[b + 1] : (int)n;
  unsigned bits = 0;
  for (int i = bstart[b]; i < e; ++i) {
    const int j = sidx[i];
    bits |= 1u << ((rid[j] % 4) * 4 + col[j] % 4);
  }
  map[b] = (uint16_t)bits;
}

__global__ void warp_counts(const int32_t* __restrict__ rowptr, int64_t brows, int cap,
                            int32_t* __restrict__ cnt) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < brows) cnt[r] = (rowptr[r + 1] - rowptr[r] + cap - 1) / cap;
  if (r == brows) cnt[r] = 0;
}

__global__ void warp_rows(const int32_t* __restrict__ rpw, int64_t brows, int32_t* __restrict__ riw) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= brows) return;
  for (int w = rpw[r]; w < rpw[r + 1]; ++w) riw[w] = (int32_t)r;
}

// sum over block rows of (len - avg)^2, fixed assignment and tree order.
__global__ void row_len_sq_dev(const int32_t* __restrict__ rowptr, int64_t brows, double avg,
                               double* __restrict__ out) {
  __shared__ double red[kTB];
  double s = 0.0;
  for (int64_t r = threadIdx.x; r < brows; r += kTB) {
    const double d = (double)(rowptr[r + 1] - rowptr[r]) - avg;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = kTB / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// ---- SpMV kernels -----------------------------------------------------------

// Wave per block row; lane = slot*tm + r (slots = 64/tm blocks per step).
// Partials of equal r are added in slot order through LDS. TK > 0: the
// block width known at compile time, a block row's TK values and x entries
// loaded together before the sum (same order of additions as TK = 0, which
// takes any width).
template <int TK>
__global__ __launch_bounds__(kTB) void bsr_valu_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const double* __restrict__ x, double* __restrict__ y,
    int64_t brows, int64_t nrows, int64_t ncols, int tm, int tk, int64_t first) {
  __shared__ double part[kTB];
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t br = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int slots = kWave / tm;
  const int slot = lane / tm, r = lane % tm;
  double acc = 0.0;
  if (br < brows && slot < slots) {
    for (int k = rowptr[br] + slot, e = rowptr[br + 1]; k < e; k += slots) {
      const int64_t c0 = (int64_t)colidx[k] * tk;
      const double* p = bval + (int64_t)k * tm * tk + r * tk;
      const int cmax = (int)min<int64_t>(tk, ncols - c0);
      if constexpr (TK > 0) {
        double pv[TK], xv[TK];
#pragma unroll
        for (int c = 0; c < TK; ++c) {
          pv[c] = c < cmax ? ld_stream(&p[c]) : 0.0;
          xv[c] = c < cmax ? x[c0 + c] : 0.0;
        }
#pragma unroll
        for (int c = 0; c < TK; ++c)
          if (c < cmax) acc += pv[c] * xv[c];
      } else {
        for (int c = 0; c < cmax; ++c) acc += ld_stream(&p[c]) * x[c0 + c];
      }
    }
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  if (br < brows && lane < tm) {
    const int base = threadIdx.x - lane;
    double s = 0.0;
    for (int q = 0; q < slots; ++q) s += part[base + q * tm + lane];
    const int64_t row = br * tm + lane;
    if (row < nrows) y[row] = s;
  }
}

// 4x4 blocks on v_mfma_f64_16x16x4f64. Operand layouts (wave64, measured on
// gfx950 with power-of-two encoded operands):
//   A (16x4):  lane l holds A[l%16][l/16]
//   B (4x16):  lane l holds B[l/16][l%16]
//   C (16x16): lane l holds C[l/16 + 4v][l%16] in acc[v], v = 0..3.
// Rows 4a..4a+3 of A are block a of the current group of four blocks of the
// block row; column j of B is the x segment of block j/4. The diagonal 4x4
// tile a then holds A_a*x_a in every column: its row r sits in acc[a] of
// lanes 16r + 4a + t. Lane 16r + 4a (t = 0) contributes acc[a]; the four
// lanes 16r + {0, 4, 8, 12} add up to y row r of the block row.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef double dbl2v __attribute__((ext_vector_type(2)));

//
// Two products share one 16-byte operand load per lane: K index 2j + h
// (j = kk >> 1, h = kk & 1) takes block a + 4j of a group of eight and, in
// product m = 0, 1, its column 2h + m, so a lane's two A values (columns 2h,
// 2h + 1 of row r) and its two x values are adjacent, and a wave's A load is
// the group's eight blocks, 1 KB contiguous. The diagonal tile a then holds
// blocks a and a + 4 summed. X16: x is 16-byte aligned (else its pairs are
// two 8-byte loads).
template <bool X16>
__global__ __launch_bounds__(kTB) void bsr4_mfma_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const double* __restrict__ x, double* __restrict__ y,
    int64_t brows, int64_t nrows, int64_t ncols, int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t br = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (br >= brows) return;  // whole wave exits together
  const int a = (lane & 15) >> 2, r = lane & 3;  // A row 4a + r / B column slot a
  const int h = (lane >> 4) & 1, j = lane >> 5;  // k index 2j + h
  const int s0 = rowptr[br], e0 = rowptr[br + 1];
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  // block column indices 64 at a time (one coalesced load), then four
  // groups of eight blocks per step with all their operand loads in flight
  for (int base = s0; base < e0; base += kWave) {
    const int nb = min(kWave, e0 - base);
    const int cid = lane < nb ? colidx[base + lane] : 0;
    for (int k0 = 0; k0 < nb; k0 += 32) {
      dbl2v av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kb = k0 + 8 * u + a + 4 * j;
        const int64_t c0 = (int64_t)__shfl(cid, min(kb, kWave - 1), 64) * 4 + 2 * h;
        const bool ok = kb < nb;
        av[u] = ok ? ld_stream(reinterpret_cast<const dbl2v*>(
                         bval + (int64_t)(base + kb) * 16 + r * 4 + 2 * h))
                   : dbl2v{0.0, 0.0};
        if (ok && c0 + 1 < ncols) {
          if (X16)
            bv[u] = *reinterpret_cast<const dbl2v*>(x + c0);
          else
            bv[u] = dbl2v{x[c0], x[c0 + 1]};
        } else {
          bv[u] = dbl2v{ok && c0 < ncols ? x[c0] : 0.0, 0.0};
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (k0 + 8 * u < nb) {
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, acc, 0, 0, 0);
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].y, acc, 0, 0, 0);
        }
    }
  }
  // lane 16r + 4a + t keeps acc[a]; sum over a (lane bits 2..3)
  double t = a == 0 ? acc[0] : a == 1 ? acc[1] : a == 2 ? acc[2] : acc[3];
  t += __shfl_xor(t, 4, 64);
  t += __shfl_xor(t, 8, 64);
  if ((lane & 15) == 0) {
    const int64_t row = br * 4 + (lane >> 4);
    if (row < nrows) y[row] = t;
  }
}

// ---- balanced 4x4 kernels (src/amgT.cu:160-209, 464-521) -------------------
// Work unit: chunk w of the block-row balance arrays, blocks
// [rowptr[br] + 64k, min(rowptr[br] + 64(k+1), rowptr[br+1])) of block row
// br = rowidx_by_warp[w], k = w - rowptr_by_warp[br], described by one
// 16-byte entry of the chunk table {br, first block, blocks, single} built
// with the layout. A block row of one chunk ("single") gets its y rows
// written directly; otherwise the chunk's four row sums go to
// part[4w..4w+3] and bsr4_bal_combine adds them in chunk order (no atomics:
// the same bits every run).
__global__ void bal_chunk_table(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ rpw,
                                const int32_t* __restrict__ riw, int64_t nwarps,
                                int4* __restrict__ ctab) {
  const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (w >= nwarps) return;
  const int br = riw[w];
  const int w0 = rpw[br], w1 = rpw[br + 1];
  const int s = rowptr[br] + 64 * (int)(w - w0);
  ctab[w] = make_int4(br, s, min(64, rowptr[br + 1] - s), w1 - w0 == 1);
}

__device__ __forceinline__ void bal_store(const int4& d, int64_t w, int r, double t,
                                          int64_t nrows, double* __restrict__ y,
                                          double* __restrict__ part) {
  if (d.w) {
    const int64_t row = (int64_t)d.x * 4 + r;
    if (row < nrows) y[row] = t;
  } else {
    part[w * 4 + r] = t;
  }
}

// MFMA: the chunk's blocks eight at a time, two v_mfma_f64_16x16x4f64 per
// eight blocks sharing one 16-byte operand load per lane. A row 4a + r (lane
// bits 0..3) is row r of block slot a; K index 2j + h (lane bits 4..5)
// takes block a + 4j of the step and, in product m = 0, 1, its column
// 2h + m. So a lane's two A values A_b[r][2h], A_b[r][2h+1] are adjacent (a
// wave's A load is the step's eight blocks, 1 KB contiguous) and so are its
// two x values x[4 col_b + 2h], x[4 col_b + 2h + 1]; the diagonal 4x4 tiles
// of the accumulated result hold the rows of blocks a and a + 4 summed
// (layouts otherwise as in bsr4_mfma_f64). 32 blocks' loads in flight per
// step. X16: x is 16-byte aligned (else its pairs are two 8-byte loads).
template <bool X16>
__global__ __launch_bounds__(kTB) void bsr4_bal_mfma_f64(
    const int4* __restrict__ ctab, int64_t nwarps, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ part, int64_t nrows, int64_t ncols, int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (w >= nwarps) return;  // whole wave exits together
  const int4 d = ctab[w];
  const int a = (lane & 15) >> 2, r = lane & 3, h = (lane >> 4) & 1, j = lane >> 5;
  const int cid = lane < d.z ? colidx[d.y + lane] : 0;
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < d.z; k0 += 32) {
    dbl2v av[4], bv[4];
    // A operands first: they do not wait for the column indices
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kb = k0 + 8 * u + a + 4 * j;
      av[u] = kb < d.z ? ld_stream(reinterpret_cast<const dbl2v*>(
                             bval + (int64_t)(d.y + kb) * 16 + r * 4 + 2 * h))
                       : dbl2v{0.0, 0.0};
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int kb = k0 + 8 * u + a + 4 * j;
      const int64_t c0 = (int64_t)__shfl(cid, min(kb, kWave - 1), 64) * 4 + 2 * h;
      if (kb < d.z && c0 + 1 < ncols) {
        if (X16)
          bv[u] = *reinterpret_cast<const dbl2v*>(x + c0);
        else
          bv[u] = dbl2v{x[c0], x[c0 + 1]};
      } else {
        bv[u] = dbl2v{kb < d.z && c0 < ncols ? x[c0] : 0.0, 0.0};
      }
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k0 + 8 * u < d.z) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].y, acc, 0, 0, 0);
      }
  }
  double t = a == 0 ? acc[0] : a == 1 ? acc[1] : a == 2 ? acc[2] : acc[3];
  t += __shfl_xor(t, 4, 64);
  t += __shfl_xor(t, 8, 64);
  if ((lane & 15) == 0) bal_store(d, w, lane >> 4, t, nrows, y, part);
}

// VALU with the occupancy bitmap, four chunks per wave: 16 lanes per chunk
// in four groups of four, group g takes blocks g, g+4, g+8, ... of the
// chunk, lane r of a group row r of its block; only entries whose map bit
// r*4+c is set are multiplied (a non-finite x entry in a padded column does
// not reach y). Row sums over the groups by lane shuffles. X16: x is 16-byte
// aligned, and a block's four x values are two 16-byte loads whenever the
// block column lies inside x (else, and for an unaligned x, four masked
// 8-byte loads).
template <bool X16>
__global__ __launch_bounds__(kTB) void bsr4_bal_valu_f64(
    const int4* __restrict__ ctab, int64_t nwarps, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const uint16_t* __restrict__ map,
    const double* __restrict__ x, double* __restrict__ y, double* __restrict__ part,
    int64_t nrows, int64_t ncols, int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t q = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (4 * q >= nwarps) return;
  const int slot = lane >> 4, g = (lane >> 2) & 3, r = lane & 3;
  const int64_t w = 4 * q + slot;
  const int4 d = w < nwarps ? ctab[w] : make_int4(0, 0, 0, 1);
  double t = 0.0;
  for (int i0 = 0; i0 < d.z; i0 += 16) {  // 16 blocks of the chunk per step
    unsigned bits[4];
    int cb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = i0 + g + 4 * i;
      const bool ok = k < d.z;
      bits[i] = ok ? ((unsigned)map[d.y + k] >> (4 * r)) & 15u : 0u;
      cb[i] = ok ? colidx[d.y + k] : 0;
    }
    dbl4 v[4];
    double xv[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int64_t b = (int64_t)(d.y + i0 + g + 4 * i) * 16 + r * 4;
      v[i] = i0 + g + 4 * i < d.z ? ld_stream(reinterpret_cast<const dbl4*>(bval + b))
                                  : dbl4{0.0, 0.0, 0.0, 0.0};
      const int64_t c0 = (int64_t)cb[i] * 4;
      if (X16 && bits[i] && c0 + 3 < ncols) {
        const dbl2v lo = *reinterpret_cast<const dbl2v*>(x + c0);
        const dbl2v hi = *reinterpret_cast<const dbl2v*>(x + c0 + 2);
        xv[i][0] = lo.x;
        xv[i][1] = lo.y;
        xv[i][2] = hi.x;
        xv[i][3] = hi.y;
      } else {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          xv[i][c] = ((bits[i] >> c) & 1u) && c0 + c < ncols ? x[c0 + c] : 0.0;
      }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      if (bits[i] & 1u) t += v[i].x * xv[i][0];
      if (bits[i] & 2u) t += v[i].y * xv[i][1];
      if (bits[i] & 4u) t += v[i].z * xv[i][2];
      if (bits[i] & 8u) t += v[i].w * xv[i][3];
    }
  }
  t += __shfl_xor(t, 4, 64);
  t += __shfl_xor(t, 8, 64);
  if ((lane & 12) == 0 && w < nwarps) bal_store(d, w, r, t, nrows, y, part);
}

// y rows of the block rows listed in `multi`: the chunk partials in chunk
// order (zero for a block row without blocks).
__global__ void bsr4_bal_combine(const int32_t* __restrict__ multi, int64_t nmulti,
                                 const int32_t* __restrict__ rpw, const double* __restrict__ part,
                                 int64_t nrows, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nmulti * 4) return;
  const int64_t br = multi[i >> 2];
  const int r = (int)(i & 3);
  double s = 0.0;
  for (int w = rpw[br]; w < rpw[br + 1]; ++w) s += part[(int64_t)w * 4 + r];
  const int64_t row = br * 4 + r;
  if (row < nrows) y[row] = s;
}

__global__ void bal_multi_flags(const int32_t* __restrict__ rpw, int64_t brows,
                                int32_t* __restrict__ flag) {
  const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (br > brows) return;
  flag[br] = br < brows && rpw[br + 1] - rpw[br] != 1;
}

__global__ void bal_multi_scatter(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos,
                                  int64_t brows, int32_t* __restrict__ multi) {
  const int64_t br = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (br < brows && flag[br]) multi[pos[br]] = (int32_t)br;
}

// 8x8 blocks on the same v_mfma_f64_16x16x4f64, two blocks per product:
// A rows 8a..8a+7 are rows of block a (a = 0, 1) at four of its columns
// (columns 2k + h at K index k in product h = 0, 1: two products per pair),
// B column j is the matching x segment of block j/8. The two 8x8 diagonal
// tiles of the 16x16 result then hold A_a*x_a in every column: result row R,
// column R sits in acc[R/4] of lane 16(R%4) + R (C layout above), and y row
// r of the block row is the sum of result rows r and r+8, found in lanes L
// and L+8, L = 16(r%4) + r. A lane's two A values (columns 2k, 2k + 1) and
// its two x values are adjacent: one 16-byte load each (X16: x 16-byte
// aligned, else two 8-byte loads).
template <bool X16>
__global__ __launch_bounds__(kTB) void bsr8_mfma_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const double* __restrict__ x, double* __restrict__ y,
    int64_t brows, int64_t nrows, int64_t ncols, int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t br = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (br >= brows) return;  // whole wave exits together
  const int i = lane & 15;   // A row / B column
  const int kk = lane >> 4;  // k index
  const int a = i >> 3, r = i & 7;
  const int s0 = rowptr[br], e0 = rowptr[br + 1];
  dbl4 acc = {0.0, 0.0, 0.0, 0.0};
  // block column indices 64 at a time, then four pairs (eight blocks, eight
  // products) per step with all their operand loads in flight
  for (int base = s0; base < e0; base += kWave) {
    const int nb = min(kWave, e0 - base);
    const int cid = lane < nb ? colidx[base + lane] : 0;
    for (int k0 = 0; k0 < nb; k0 += 8) {
      dbl2v av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int kb = k0 + 2 * u + a;
        const int c8 = __shfl(cid, min(kb, kWave - 1), 64);
        const bool ok = kb < nb;
        const int64_t c = (int64_t)c8 * 8 + 2 * kk;  // columns 2kk, 2kk + 1 of the block
        av[u] = ok ? ld_stream(reinterpret_cast<const dbl2v*>(bval + (int64_t)(base + kb) * 64 +
                                                              r * 8 + 2 * kk))
                   : dbl2v{0.0, 0.0};
        if (ok && c + 1 < ncols) {
          if (X16)
            bv[u] = *reinterpret_cast<const dbl2v*>(x + c);
          else
            bv[u] = dbl2v{x[c], x[c + 1]};
        } else {
          bv[u] = dbl2v{ok && c < ncols ? x[c] : 0.0, 0.0};
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].x, bv[u].x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[u].y, bv[u].y, acc, 0, 0, 0);
      }
    }
  }
  // lane (J = lane%16, g = lane/16) holds result row g + 4v, column J in
  // acc[v]; the diagonal entry (row == column) is v = (J - g)/4
  const int J = lane & 15, g = lane >> 4, d = J - g;
  double v = 0.0;
  if (d >= 0 && (d & 3) == 0) v = (d >> 2) == 0 ? acc[0] : (d >> 2) == 1 ? acc[1] : (d >> 2) == 2 ? acc[2] : acc[3];
  const double t = v + __shfl_down(v, 8, 64);
  if (J < 8 && g == (J & 3)) {
    const int64_t row = br * 8 + J;
    if (row < nrows) y[row] = t;
  }
}

// ---- DBSR -------------------------------------------------------------------

__global__ void row_lengths(const int32_t* __restrict__ rowptr, int64_t nrows,
                            int32_t* __restrict__ len, int32_t* __restrict__ idx) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  len[r] = rowptr[r + 1] - rowptr[r];
  idx[r] = (int32_t)r;
}

__global__ void dbsr_widths(const int32_t* __restrict__ slen, int64_t nrows, int64_t brows,
                            int32_t* __restrict__ w) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g > brows) return;
  if (g == brows) {
    w[g] = 0;
    return;
  }
  int m = 0;
  for (int64_t i = 4 * g; i < 4 * g + 4 && i < nrows; ++i) m = max(m, (slen[i] + 3) / 4);
  w[g] = m;
}

// wave per (sorted) row, lanes striding the row's entries
__global__ __launch_bounds__(kTB) void dbsr_place(const int32_t* __restrict__ rowptr,
                                                  const int32_t* __restrict__ col,
                                                  const double* __restrict__ val,
                                                  const int32_t* __restrict__ rec,
                                                  const int32_t* __restrict__ blcptr, int64_t nrows,
                                                  int32_t* __restrict__ ocol,
                                                  double* __restrict__ oval, int64_t first) {
  const int64_t i = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= nrows) return;
  const int o = rec[i];
  const int64_t b0 = blcptr[i / 4];
  const int rr = (int)(i % 4) * 4;
  const int j0 = rowptr[o], e = rowptr[o + 1];
  for (int z = lane; z < e - j0; z += kWave) {
    const int64_t d = (b0 + z / 4) * 16 + rr + (z % 4);
    ocol[d] = col[j0 + z];
    oval[d] = val[j0 + z];
  }
}

__global__ void dbsr_map(const int32_t* __restrict__ slen, const int32_t* __restrict__ blcptr,
                         int64_t nrows, int64_t brows, uint16_t* __restrict__ map) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= brows) return;
  for (int b = blcptr[g]; b < blcptr[g + 1]; ++b) {
    const int z4 = 4 * (b - blcptr[g]);
    unsigned bits = 0;
    for (int q = 0; q < 4; ++q) {
      const int64_t i = 4 * g + q;
      if (i >= nrows) break;
      const int cnt = min(4, max(0, slen[i] - z4));
      bits |= ((1u << cnt) - 1u) << (4 * q);
    }
    map[b] = (uint16_t)bits;
  }
}

// balanceDbsr (include/dbsr.h:94-169) with WARP_CAPACITY 64, both passes as
// written: pass 2 starts from the fill pass 1 ended with, and a block row
// that overflows the open warp in pass 2 opens the next warp without adding
// its length to the fill. Each pass is a greedy packing whose only state
// between block rows is the fill of the open warp (0..64), so it runs as a
// parallel scan over chunks of block rows: dbsr_chunk_tables records, for
// each of the 65 fills a chunk can be entered with, the fill it leaves and
// the warps it opens; dbsr_chunk_walk composes those tables in order (one
// wave, 65 table entries per chunk); dbsr_chunk_fill replays pass 2 over each
// chunk from its entry fill and adds the per-warp block-row counts.
constexpr int kDbCap = 64;
constexpr int kDbStates = kDbCap + 1;
constexpr int kDbChunk = 1024;
constexpr int kDbTB = 128;

// pass 1 step: the warps it closes or opens go to wc
__device__ inline void db_pass1(int& now, int len, int& wc) {
  if (now + len <= kDbCap) {
    now += len;
    return;
  }
  if (now != 0) {
    ++wc;
    now = 0;
  }
  if (len <= kDbCap)
    now = len;
  else
    wc += (len + kDbCap - 1) / kDbCap;
}

// pass 2 step: returns how many warps the block row advances the open warp by
// (0: it joins the open warp)
__device__ inline int db_pass2(int& now, int len) {
  if (now + len <= kDbCap) {
    now += len;
    return 0;
  }
  now = 0;
  return len <= kDbCap ? 1 : (len + kDbCap - 1) / kDbCap;
}

__device__ inline int db_chunk_lens(const int32_t* __restrict__ blcptr, int64_t brows, int* lens) {
  const int64_t b0 = (int64_t)blockIdx.x * kDbChunk;
  const int m = (int)min((int64_t)kDbChunk, brows - b0);
  for (int i = threadIdx.x; i < m; i += blockDim.x) lens[i] = blcptr[b0 + i + 1] - blcptr[b0 + i];
  __syncthreads();
  return m;
}

// tab[c][k][s], k = 0..3: pass-1 exit fill, pass-1 warps, pass-2 exit fill,
// pass-2 warp advance, for chunk c entered with fill s
__global__ __launch_bounds__(kDbTB) void dbsr_chunk_tables(const int32_t* __restrict__ blcptr,
                                                           int64_t brows, int32_t* __restrict__ tab) {
  __shared__ int lens[kDbChunk];
  const int m = db_chunk_lens(blcptr, brows, lens);
  const int s = threadIdx.x;
  if (s >= kDbStates) return;
  int n1 = s, wc = 0, n2 = s, adv = 0;
  for (int i = 0; i < m; ++i) {
    const int len = lens[i];
    db_pass1(n1, len, wc);
    adv += db_pass2(n2, len);
  }
  int32_t* t = tab + (int64_t)blockIdx.x * 4 * kDbStates;
  t[s] = n1;
  t[kDbStates + s] = wc;
  t[2 * kDbStates + s] = n2;
  t[3 * kDbStates + s] = adv;
}

// the table entry of state st (0..64) from one wave holding entries 0..63 in
// lanes and entry 64 in every lane
__device__ inline int db_pick(int lo, int hi, int st) { return st < kWave ? __shfl(lo, st) : hi; }

// one wave: out[0] = pass-1 warps, out[1] = pass-2 warps; entry[c] = pass-2
// fill and first warp of chunk c
__global__ __launch_bounds__(kWave) void dbsr_chunk_walk(const int32_t* __restrict__ tab,
                                                         int64_t nchunks, int64_t brows,
                                                         int32_t* __restrict__ out,
                                                         int2* __restrict__ entry) {
  // The table loads do not depend on the running fill, so eight chunks'
  // entries are loaded together before the eight dependent picks.
  constexpr int G = 8;
  const int lane = threadIdx.x;
  int st = 0, wc = 0;
  for (int64_t c0 = 0; c0 < nchunks; c0 += G) {
    int el[G], eh[G], wl[G], wh[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int32_t* t = tab + min(c0 + q, nchunks - 1) * 4 * kDbStates;
      el[q] = t[lane];
      eh[q] = t[kDbStates - 1];
      wl[q] = t[kDbStates + lane];
      wh[q] = t[2 * kDbStates - 1];
    }
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (c0 + q < nchunks) {
        wc += db_pick(wl[q], wh[q], st);
        st = db_pick(el[q], eh[q], st);
      }
  }
  if (st) ++wc;
  int wv = 0;
  for (int64_t c0 = 0; c0 < nchunks; c0 += G) {
    int el[G], eh[G], al[G], ah[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
      const int32_t* t = tab + min(c0 + q, nchunks - 1) * 4 * kDbStates;
      el[q] = t[2 * kDbStates + lane];
      eh[q] = t[3 * kDbStates - 1];
      al[q] = t[3 * kDbStates + lane];
      ah[q] = t[4 * kDbStates - 1];
    }
#pragma unroll
    for (int q = 0; q < G; ++q)
      if (c0 + q < nchunks) {
        if (lane == 0) entry[c0 + q] = make_int2(st, wv);
        wv += db_pick(al[q], ah[q], st);
        st = db_pick(el[q], eh[q], st);
      }
  }
  if (lane == 0) {
    out[0] = wc;
    out[1] = brows > 0 ? wv + 1 : 0;
  }
}

// replays pass 2 over chunk c from entry[c]; cnt[w] += block rows of warp w
// (atomics: the first and last warp of a chunk can be shared with neighbours)
__global__ __launch_bounds__(kDbTB) void dbsr_chunk_fill(const int32_t* __restrict__ blcptr,
                                                         int64_t brows, const int2* __restrict__ entry,
                                                         int32_t* __restrict__ cnt) {
  __shared__ int lens[kDbChunk];
  const int m = db_chunk_lens(blcptr, brows, lens);
  if (threadIdx.x != 0) return;
  const int2 e = entry[blockIdx.x];
  int now = e.x, w = e.y, run = 0;
  for (int i = 0; i < m; ++i) {
    const int adv = db_pass2(now, lens[i]);
    if (adv == 0) {
      ++run;
      continue;
    }
    if (run) atomicAdd(cnt + w, run);
    for (int j = 1; j < adv; ++j) atomicAdd(cnt + w + j, 1);
    w += adv;
    run = 1;
  }
  if (run) atomicAdd(cnt + w, run);
}

// Lane = slot*4 + r; each lane reads its row's four slots of a block with
// one 32-byte value load and one 16-byte index load and gathers x through
// the per-slot column indices. Rows are sorted by length, so the groups at
// most 8 blocks wide come first (`narrow` of them, counted when the handle is
// made): there a wave serves two groups, each on half the wave with 8 blocks
// per step (a pair with a wider group runs its two groups one after the
// other); every later group has a wave of its own, 16 blocks per step.
typedef double dbl4 __attribute__((ext_vector_type(4)));
typedef int int4v __attribute__((ext_vector_type(4)));

// Partial sum of lane l (l < 4*SL) for group g; blocks added in step order.
template <int SL>
__device__ __forceinline__ double dbsr_group(const int32_t* __restrict__ blcptr,
                                             const int32_t* __restrict__ colidx,
                                             const double* __restrict__ bval,
                                             const double* __restrict__ x, int64_t g, int l) {
  const int slot = l >> 2, r = l & 3;
  double acc = 0.0;
  auto add = [&](const dbl4& v, const int4v& c) {
    acc += v.x * x[c.x];
    acc += v.y * x[c.y];
    acc += v.z * x[c.z];
    acc += v.w * x[c.w];
  };
  int k = blcptr[g] + slot;
  const int e = blcptr[g + 1];
  for (; k + SL < e; k += 2 * SL) {  // two steps' loads together
    const int64_t o0 = (int64_t)k * 16 + r * 4, o1 = o0 + SL * 16;
    const dbl4 v0 = __builtin_nontemporal_load(reinterpret_cast<const dbl4*>(bval + o0));
    const int4v c0 = __builtin_nontemporal_load(reinterpret_cast<const int4v*>(colidx + o0));
    const dbl4 v1 = __builtin_nontemporal_load(reinterpret_cast<const dbl4*>(bval + o1));
    const int4v c1 = __builtin_nontemporal_load(reinterpret_cast<const int4v*>(colidx + o1));
    add(v0, c0);
    add(v1, c1);
  }
  if (k < e) {
    const int64_t o = (int64_t)k * 16 + r * 4;
    add(__builtin_nontemporal_load(reinterpret_cast<const dbl4*>(bval + o)),
        __builtin_nontemporal_load(reinterpret_cast<const int4v*>(colidx + o)));
  }
  return acc;
}

__global__ __launch_bounds__(kTB) void dbsr_spmv_f64(
    const int32_t* __restrict__ blcptr, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const int32_t* __restrict__ rec, const double* __restrict__ x,
    double* __restrict__ y, int64_t brows, int64_t nrows, int64_t narrow, int sorted_order,
    int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t npair = (narrow + 1) / 2;
  auto store = [&](int64_t g, int r, double v) {
    const int64_t i = 4 * g + r;
    if (i < nrows) y[sorted_order ? i : rec[i]] = v;
  };
  if (w >= npair) {  // a wide group, one wave
    const int64_t g = narrow + (w - npair);
    if (g >= brows) return;
    double acc = dbsr_group<16>(blcptr, colidx, bval, x, g, lane);
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 8, 64);
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    if (lane < 4) store(g, lane, acc);
    return;
  }
  const int64_t g0 = 2 * w;
  const bool pair = g0 + 1 < narrow && blcptr[g0 + 1] - blcptr[g0] <= 8 &&
                    blcptr[g0 + 2] - blcptr[g0 + 1] <= 8;
  if (pair) {
    const int h = lane >> 5, l = lane & 31;
    double acc = dbsr_group<8>(blcptr, colidx, bval, x, g0 + h, l);
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 8, 64);
    acc += __shfl_xor(acc, 16, 64);
    if (l < 4) store(g0 + h, l, acc);
    return;
  }
  for (int64_t g = g0; g < g0 + 2 && g < narrow; ++g) {
    double acc = dbsr_group<16>(blcptr, colidx, bval, x, g, lane);
    acc += __shfl_xor(acc, 4, 64);
    acc += __shfl_xor(acc, 8, 64);
    acc += __shfl_xor(acc, 16, 64);
    acc += __shfl_xor(acc, 32, 64);
    if (lane < 4) store(g, lane, acc);
  }
}

// DBSR on v_mfma_f64_16x16x4f64 (the reference's DBSR kernels use
// mma.m8n8k4 with a diagonal pick, src/amgT.cu:324-387). Lane (i, t) =
// (l%16, l/16) owns row i%4 of one block of the step: one 32-byte load of
// its four values and one 16-byte load of their columns, and in product u
// it supplies slot u (value and its column's x) as A[i][t] and B[t][i], so
// the result's diagonal C[i][i] sums the row over the four blocks that lanes
// (i, 0..3) own (C layout as in bsr4_mfma_f64: C[R][R] sits in acc[R/4] of
// lane 16(R%4) + R). A wide group takes a wave, sixteen blocks per step
// (lane (i, t) on block 4t + i/4), and its row r is the sum of C[4a+r][4a+r]
// over a; two narrow groups (at most 8 blocks wide, the prefix of the
// length-sorted groups) share a wave, A rows 8h..8h+7 for group h and lane
// (i, t) on block 2t + (i/4)%2 of its group, so each group's row r is
// C[8h+r][8h+r] + C[8h+4+r][8h+4+r].
__device__ __forceinline__ void dbsr_mfma_step(const double* __restrict__ bval,
                                               const int32_t* __restrict__ colidx,
                                               const double* __restrict__ x, int blk, bool ok,
                                               int r, dbl4& acc) {
  typedef int i4v __attribute__((ext_vector_type(4)));
  dbl4 av = {0.0, 0.0, 0.0, 0.0};
  double xv[4] = {0.0, 0.0, 0.0, 0.0};
  if (ok) {
    const int64_t o = (int64_t)blk * 16 + r * 4;
    av = ld_stream(reinterpret_cast<const dbl4*>(bval + o));
    const i4v c = ld_stream(reinterpret_cast<const i4v*>(colidx + o));
    xv[0] = x[c.x];
    xv[1] = x[c.y];
    xv[2] = x[c.z];
    xv[3] = x[c.w];
  }
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.x, xv[0], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.y, xv[1], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.z, xv[2], acc, 0, 0, 0);
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av.w, xv[3], acc, 0, 0, 0);
}

__device__ __forceinline__ double dbsr_mfma_diag(const dbl4& acc, int lane) {
  const int R = lane & 15;
  if ((lane >> 4) != (R & 3)) return 0.0;
  return (R >> 2) == 0 ? acc[0] : (R >> 2) == 1 ? acc[1] : (R >> 2) == 2 ? acc[2] : acc[3];
}

__global__ __launch_bounds__(kTB) void dbsr_spmv_f64_mfma(
    const int32_t* __restrict__ blcptr, const int32_t* __restrict__ colidx,
    const double* __restrict__ bval, const int32_t* __restrict__ rec, const double* __restrict__ x,
    double* __restrict__ y, int64_t brows, int64_t nrows, int64_t narrow, int sorted_order,
    int64_t first) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t w = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int64_t npair = (narrow + 1) / 2;
  const int i = lane & 15, t = lane >> 4, r = lane & 3, R = lane & 15;
  auto store = [&](int64_t g, int rr, double v) {
    const int64_t row = 4 * g + rr;
    if (row < nrows) y[sorted_order ? row : rec[row]] = v;
  };
  auto wide = [&](int64_t g) {  // 16 blocks per step
    const int s0 = blcptr[g], e0 = blcptr[g + 1];
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    for (int k0 = s0; k0 < e0; k0 += 16) {
      const int blk = k0 + 4 * t + (i >> 2);
      dbsr_mfma_step(bval, colidx, x, blk, blk < e0, r, acc);
    }
    double d = dbsr_mfma_diag(acc, lane);
    d += __shfl_xor(d, 4, 64);
    d += __shfl_xor(d, 8, 64);
    if (R == t) store(g, t, d);
  };
  if (w >= npair) {  // a wide group, one wave
    const int64_t g = narrow + (w - npair);
    if (g < brows) wide(g);
    return;
  }
  const int64_t g0 = 2 * w;
  const bool pair = g0 + 1 < narrow && blcptr[g0 + 1] - blcptr[g0] <= 8 &&
                    blcptr[g0 + 2] - blcptr[g0 + 1] <= 8;
  if (pair) {
    const int h = i >> 3;
    const int s0 = blcptr[g0 + h], e0 = blcptr[g0 + h + 1];
    const int blk = s0 + 2 * t + ((i >> 2) & 1);
    dbl4 acc = {0.0, 0.0, 0.0, 0.0};
    dbsr_mfma_step(bval, colidx, x, blk, blk < e0, r, acc);
    double d = dbsr_mfma_diag(acc, lane);
    d += __shfl_xor(d, 4, 64);
    if (R == t || R == 8 + t) store(g0 + (R >> 3), t, d);
    return;
  }
  for (int64_t g = g0; g < g0 + 2 && g < narrow; ++g) wide(g);
}

// f[g] = 1 when 4-row group g is at most 8 blocks wide
__global__ void dbsr_narrow_flags(const int32_t* __restrict__ blcptr, int64_t brows,
                                  int64_t* __restrict__ f) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < brows) f[g] = blcptr[g + 1] - blcptr[g] <= 8;
}

// Rows of a CSR in recorder order: new row i is old row rec[i] (wave per row).
__global__ void csr_permute_rows(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
                                 const double* __restrict__ val, const int32_t* __restrict__ rec,
                                 const int32_t* __restrict__ nrp, int64_t n,
                                 int32_t* __restrict__ ocol, double* __restrict__ oval,
                                 int64_t first) {
  const int64_t i = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  if (i >= n) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int o = rec[i];
  const int s = rowptr[o], len = rowptr[o + 1] - s, d = nrp[i];
  for (int k = lane; k < len; k += kWave) {
    ocol[d + k] = col[s + k];
    oval[d + k] = val[s + k];
  }
}

}  // namespace

// Groups of a DBSR handle that take the two-groups-per-wave path.
static bsp_status dbsr_count_narrow(bsp_dbsr_s* h, hipStream_t s) {
  const int64_t brows = h->v.brows;
  h->narrow = 0;
  if (brows == 0) return BSP_OK;
  DevBuf<int64_t> f, sum;
  BSP_TRY(f.alloc(brows));
  BSP_TRY(sum.alloc(1));
  dbsr_narrow_flags<<<cdiv(brows, kTB), kTB, 0, s>>>(h->blcptr.get(), brows, f.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(reduce_sum_i64(f.get(), sum.get(), brows, s));
  BSP_TRY(fetch(sum.get(), &h->narrow, s));
  return BSP_OK;
}

// ---- builders -----------------------------------------------------------------

static int bits_for(int64_t v) {
  int b = 1;
  while (b < 63 && (int64_t(1) << b) <= v) ++b;
  return b;
}

bsp_status bsr_count_blocks(const bsp_csr& A, int tm, int tk, hipStream_t s, int64_t* nblocks) {
  const int64_t n = A.nrows, nnz = A.nnz;
  *nblocks = 0;
  if (n == 0 || nnz == 0) return BSP_OK;
  const int64_t brows = cdiv(n, tm), bcols = cdiv(A.ncols, tk);
  DevBuf<int32_t> rid, idx, sidx, flag, pos;
  DevBuf<int64_t> key, skey;
  BSP_TRY(rid.alloc(nnz));
  BSP_TRY(idx.alloc(nnz));
  BSP_TRY(sidx.alloc(nnz));
  BSP_TRY(flag.alloc(nnz));
  BSP_TRY(pos.alloc(nnz));
  BSP_TRY(key.alloc(nnz));
  BSP_TRY(skey.alloc(nnz));
  row_ids<<<cdiv(n, kTB), kTB, 0, s>>>(A.rowptr, n, rid.get());
  bsr_keys<<<cdiv(nnz, kTB), kTB, 0, s>>>(rid.get(), A.colidx, nnz, tm, tk, bcols, key.get(),
                                          idx.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(sort_pairs_i64key(key.get(), skey.get(), idx.get(), sidx.get(), nnz,
                            bits_for(brows * bcols), s));
  key_flags<<<cdiv(nnz, kTB), kTB, 0, s>>>(skey.get(), nnz, flag.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(exclusive_scan_i32(flag.get(), pos.get(), nnz, s));
  int32_t lp = 0, lf = 0;
  BSP_TRY(fetch(pos.get() + nnz - 1, &lp, s));
  BSP_TRY(fetch(flag.get() + nnz - 1, &lf, s));
  *nblocks = (int64_t)lp + lf;
  return BSP_OK;
}

static bsp_status build_bsr(const bsp_csr& A, int tm, int tk, hipStream_t s, bsp_bsr_s* h) {
  const int64_t n = A.nrows, nnz = A.nnz;
  bsp_bsr_view& v = h->v;
  v.nrows = n;
  v.ncols = A.ncols;
  v.nnz = nnz;
  v.tm = tm;
  v.tk = tk;
  v.brows = cdiv(n, tm);
  v.bcols = cdiv(A.ncols, tk);
  const auto* aval = static_cast<const double*>(A.val);
  DevBuf<int32_t> rid, idx, sidx, flag, pos, bstart;
  DevBuf<int64_t> key, skey;
  BSP_TRY(rid.alloc(nnz));
  BSP_TRY(idx.alloc(nnz));
  BSP_TRY(sidx.alloc(nnz));
  BSP_TRY(flag.alloc(nnz));
  BSP_TRY(pos.alloc(nnz));
  BSP_TRY(key.alloc(nnz));
  BSP_TRY(skey.alloc(nnz));
  int64_t nblocks = 0;
  if (nnz > 0) {
    row_ids<<<cdiv(n, kTB), kTB, 0, s>>>(A.rowptr, n, rid.get());
    bsr_keys<<<cdiv(nnz, kTB), kTB, 0, s>>>(rid.get(), A.colidx, nnz, tm, tk, v.bcols, key.get(),
                                            idx.get());
    BSP_HIP(hipGetLastError());
    BSP_TRY(sort_pairs_i64key(key.get(), skey.get(), idx.get(), sidx.get(), nnz,
                              bits_for(v.brows * v.bcols), s));
    key_flags<<<cdiv(nnz, kTB), kTB, 0, s>>>(skey.get(), nnz, flag.get());
    BSP_HIP(hipGetLastError());
    BSP_TRY(exclusive_scan_i32(flag.get(), pos.get(), nnz, s));
    int32_t lp = 0, lf = 0;
    BSP_TRY(fetch(pos.get() + nnz - 1, &lp, s));
    BSP_TRY(fetch(flag.get() + nnz - 1, &lf, s));
    nblocks = (int64_t)lp + lf;
  }
  v.nblocks = nblocks;
  DevBuf<int32_t> rowcnt;
  BSP_TRY(rowcnt.alloc(v.brows + 1));
  BSP_TRY(h->rowptr.alloc(v.brows + 1));
  BSP_TRY(h->colidx.alloc(nblocks));
  BSP_TRY(h->val.alloc(nblocks * tm * tk));
  BSP_TRY(bstart.alloc(nblocks + 1));
  BSP_HIP(hipMemsetAsync(rowcnt.get(), 0, sizeof(int32_t) * (v.brows + 1), s));
  BSP_HIP(hipMemsetAsync(h->val.get(), 0, sizeof(double) * (nblocks ? nblocks * tm * tk : 1), s));
  if (nnz > 0) {
    bsr_blocks<<<cdiv(nnz, kTB), kTB, 0, s>>>(skey.get(), flag.get(), pos.get(), nnz, v.bcols,
                                              h->colidx.get(), bstart.get(), rowcnt.get());
    bsr_values<<<cdiv(nnz, kTB), kTB, 0, s>>>(sidx.get(), flag.get(), pos.get(), nnz, rid.get(),
                                              A.colidx, aval, tm, tk, h->val.get());
    BSP_HIP(hipGetLastError());
  }
  BSP_TRY(exclusive_scan_i32(rowcnt.get(), h->rowptr.get(), v.brows + 1, s));
  v.rowptr = h->rowptr.get();
  v.colidx = h->colidx.get();
  v.val = h->val.get();
  v.map = nullptr;
  v.avg_nnz = nblocks ? (double)nnz / (double)nblocks : 0.0;
  if (tm == 4 && tk == 4) {
    BSP_TRY(h->map.alloc(nblocks + 1));
    bsr_map4<<<cdiv(nblocks + 1, kTB), kTB, 0, s>>>(bstart.get(), nblocks, nnz, sidx.get(),
                                                    rid.get(), A.colidx, h->map.get());
    BSP_HIP(hipGetLastError());
    v.map = h->map.get();
    DevBuf<int32_t> wc;
    BSP_TRY(wc.alloc(v.brows + 1));
    BSP_TRY(h->rpw.alloc(v.brows + 1));
    warp_counts<<<cdiv(v.brows + 1, kTB), kTB, 0, s>>>(h->rowptr.get(), v.brows, 64, wc.get());
    BSP_HIP(hipGetLastError());
    BSP_TRY(exclusive_scan_i32(wc.get(), h->rpw.get(), v.brows + 1, s));
    int32_t wn = 0;
    BSP_TRY(fetch(h->rpw.get() + v.brows, &wn, s));
    v.warpnum = wn;
    BSP_TRY(h->riw.alloc(wn));
    warp_rows<<<cdiv(v.brows, kTB), kTB, 0, s>>>(h->rpw.get(), v.brows, h->riw.get());
    BSP_HIP(hipGetLastError());
    v.rowptr_by_warp = h->rpw.get();
    v.rowidx_by_warp = h->riw.get();
    // balanced kernels' workspace: the list of block rows not covered by
    // exactly one chunk, and four partial sums per chunk
    {
      DevBuf<int32_t> mf, mp;
      BSP_TRY(mf.alloc(v.brows + 1));
      BSP_TRY(mp.alloc(v.brows + 1));
      bal_multi_flags<<<cdiv(v.brows + 1, kTB), kTB, 0, s>>>(h->rpw.get(), v.brows, mf.get());
      BSP_HIP(hipGetLastError());
      BSP_TRY(exclusive_scan_i32(mf.get(), mp.get(), v.brows + 1, s));
      int32_t nm = 0;
      BSP_TRY(fetch(mp.get() + v.brows, &nm, s));
      h->nmulti = nm;
      BSP_TRY(h->multi.alloc(nm));
      if (v.brows > 0)
        bal_multi_scatter<<<cdiv(v.brows, kTB), kTB, 0, s>>>(mf.get(), mp.get(), v.brows,
                                                             h->multi.get());
      BSP_HIP(hipGetLastError());
      BSP_TRY(h->part.alloc((int64_t)wn * 4));
      BSP_TRY(h->ctab.alloc(wn));
      if (wn > 0)
        bal_chunk_table<<<cdiv(wn, kTB), kTB, 0, s>>>(h->rowptr.get(), h->rpw.get(),
                                                      h->riw.get(), wn, h->ctab.get());
      BSP_HIP(hipGetLastError());
      BSP_HIP(hipStreamSynchronize(s));
    }
    DevBuf<double> sq;
    BSP_TRY(sq.alloc(1));
    const double avg_len = v.brows ? (double)nblocks / (double)v.brows : 0.0;
    row_len_sq_dev<<<1, kTB, 0, s>>>(h->rowptr.get(), v.brows, avg_len, sq.get());
    BSP_HIP(hipGetLastError());
    double acc = 0.0;
    BSP_TRY(fetch(sq.get(), &acc, s));
    v.stand = v.brows ? std::sqrt(acc / (double)v.brows) : 0.0;
  }
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

// balanceDbsr (dbsr.h:94-169) on a device block-row pointer: prefix
// (warps + 1 entries, allocated here), the warp count, and whether the
// reference's second pass would run past its warp list.
static bsp_status dbsr_balance(const int32_t* blcptr, int64_t brows, hipStream_t s,
                               DevBuf<int32_t>& prefix, int32_t* warpnum, int32_t* overflow) {
  const int64_t nchunks = cdiv(brows, kDbChunk);
  DevBuf<int32_t> counts, tab;
  DevBuf<int2> entry;
  BSP_TRY(counts.alloc(2));
  BSP_TRY(tab.alloc(nchunks * 4 * kDbStates));
  BSP_TRY(entry.alloc(nchunks));
  if (nchunks > 0) dbsr_chunk_tables<<<nchunks, kDbTB, 0, s>>>(blcptr, brows, tab.get());
  dbsr_chunk_walk<<<1, kWave, 0, s>>>(tab.get(), nchunks, brows, counts.get(), entry.get());
  BSP_HIP(hipGetLastError());
  int32_t c[2] = {0, 0};
  BSP_HIP(hipMemcpyAsync(c, counts.get(), sizeof(c), hipMemcpyDeviceToHost, s));
  BSP_HIP(hipStreamSynchronize(s));
  *overflow = c[1] > c[0] ? 1 : 0;
  const int32_t warps = c[1] > c[0] ? c[1] : c[0];
  *warpnum = warps;
  DevBuf<int32_t> cnt;
  BSP_TRY(cnt.alloc(warps + 1));
  BSP_TRY(prefix.alloc(warps + 1));
  BSP_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int32_t) * (warps + 1), s));
  if (nchunks > 0) dbsr_chunk_fill<<<nchunks, kDbTB, 0, s>>>(blcptr, brows, entry.get(), cnt.get());
  BSP_HIP(hipGetLastError());
  // prefix[i] = block rows of warps 0..i-1, i = 0..warps (cnt[warps] is 0)
  BSP_TRY(exclusive_scan_i32(cnt.get(), prefix.get(), (int64_t)warps + 1, s));
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

// sort_rows: csrTodbsrWithRecorder (rows sorted by (length, index) first);
// otherwise csrTodbsr on the rows as given (recorder = identity). Both end
// with balanceDbsr.
static bsp_status build_dbsr(const bsp_csr& A, bool sort_rows, hipStream_t s, bsp_dbsr_s* h) {
  const int64_t n = A.nrows;
  bsp_dbsr_view& v = h->v;
  v.nrows = n;
  v.ncols = A.ncols;
  v.nnz = A.nnz;
  v.brows = cdiv(n, 4);
  DevBuf<int32_t> len, idx, slen, w;
  BSP_TRY(len.alloc(n));
  BSP_TRY(idx.alloc(n));
  BSP_TRY(slen.alloc(n));
  BSP_TRY(h->recorder.alloc(n));
  if (n > 0) {
    row_lengths<<<cdiv(n, kTB), kTB, 0, s>>>(A.rowptr, n, len.get(), idx.get());
    BSP_HIP(hipGetLastError());
    if (sort_rows) {
      // stable radix sort by length keeps equal lengths in index order: the
      // (length, index) order of dbsr.h:213-216
      BSP_TRY(sort_pairs_i32(len.get(), slen.get(), idx.get(), h->recorder.get(), n, false, 32, s));
    } else {
      BSP_HIP(hipMemcpyAsync(slen.get(), len.get(), sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
      BSP_HIP(hipMemcpyAsync(h->recorder.get(), idx.get(), sizeof(int32_t) * n,
                             hipMemcpyDeviceToDevice, s));
    }
  }
  BSP_TRY(w.alloc(v.brows + 1));
  BSP_TRY(h->blcptr.alloc(v.brows + 1));
  dbsr_widths<<<cdiv(v.brows + 1, kTB), kTB, 0, s>>>(slen.get(), n, v.brows, w.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(exclusive_scan_i32(w.get(), h->blcptr.get(), v.brows + 1, s));
  int32_t nb = 0;
  BSP_TRY(fetch(h->blcptr.get() + v.brows, &nb, s));
  v.nblocks = nb;
  BSP_TRY(h->colidx.alloc((int64_t)nb * 16));
  BSP_TRY(h->val.alloc((int64_t)nb * 16));
  BSP_TRY(h->map.alloc(nb));
  BSP_HIP(hipMemsetAsync(h->colidx.get(), 0, sizeof(int32_t) * (nb ? nb * 16 : 1), s));
  BSP_HIP(hipMemsetAsync(h->val.get(), 0, sizeof(double) * (nb ? nb * 16 : 1), s));
  if (n > 0) {
    in_slices(n, kWave, [&](int64_t f, int64_t c) {
      dbsr_place<<<cdiv(c * kWave, kTB), kTB, 0, s>>>(
          A.rowptr, A.colidx, static_cast<const double*>(A.val), h->recorder.get(),
          h->blcptr.get(), n, h->colidx.get(), h->val.get(), f);
    });
    dbsr_map<<<cdiv(v.brows, kTB), kTB, 0, s>>>(slen.get(), h->blcptr.get(), n, v.brows,
                                                h->map.get());
    BSP_HIP(hipGetLastError());
  }
  BSP_TRY(dbsr_balance(h->blcptr.get(), v.brows, s, h->prefix, &v.warpnum, &v.overflow));
  v.recorder = h->recorder.get();
  v.blcptr = h->blcptr.get();
  v.colidx = h->colidx.get();
  v.val = h->val.get();
  v.map = h->map.get();
  v.prefix = h->prefix.get();
  BSP_TRY(dbsr_count_narrow(h, s));
  return BSP_OK;
}

bsp_status dbsr_build_ordered(const bsp_csr& A, bool sort_rows, hipStream_t s, bsp_dbsr* out) {
  auto h = std::make_unique<bsp_dbsr_s>();
  BSP_TRY(build_dbsr(A, sort_rows, s, h.get()));
  *out = h.release();
  return BSP_OK;
}

bsp_status dbsr_balance_host(const int32_t* blcptr, int64_t brows, int32_t** prefix,
                             int32_t* warpnum) {
  DevBuf<int32_t> bp, pre;
  BSP_TRY(bp.alloc(brows + 1));
  BSP_HIP(hipMemcpy(bp.get(), blcptr, sizeof(int32_t) * (brows + 1), hipMemcpyHostToDevice));
  int32_t wn = 0, overflow = 0;
  BSP_TRY(dbsr_balance(bp.get(), brows, nullptr, pre, &wn, &overflow));
  int32_t* h = (int32_t*)std::malloc(sizeof(int32_t) * ((size_t)wn + 1));
  if (!h) return fail(BSP_ERR_ALLOC, "host allocation failed");
  if (hipMemcpy(h, pre.get(), sizeof(int32_t) * ((size_t)wn + 1), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    std::free(h);
    return fail(BSP_ERR_HIP, "prefix download failed");
  }
  *prefix = h;
  *warpnum = wn;
  return BSP_OK;
}

bsp_status csr_sort_rows(const bsp_csr& A, hipStream_t s, int32_t* nrp, int32_t* ncol,
                         double* nval) {
  const int64_t n = A.nrows;
  if (n == 0) {
    BSP_HIP(hipMemsetAsync(nrp, 0, sizeof(int32_t), s));
    BSP_HIP(hipStreamSynchronize(s));
    return BSP_OK;
  }
  DevBuf<int32_t> len, idx, slen, rec;
  BSP_TRY(len.alloc(n));
  BSP_TRY(idx.alloc(n));
  BSP_TRY(slen.alloc(n + 1));
  BSP_TRY(rec.alloc(n));
  row_lengths<<<cdiv(n, kTB), kTB, 0, s>>>(A.rowptr, n, len.get(), idx.get());
  BSP_HIP(hipGetLastError());
  // stable by length: the (length, index) order of dbsr.h:176-179
  BSP_TRY(sort_pairs_i32(len.get(), slen.get(), idx.get(), rec.get(), n, false, 32, s));
  BSP_HIP(hipMemsetAsync(slen.get() + n, 0, sizeof(int32_t), s));
  BSP_TRY(exclusive_scan_i32(slen.get(), nrp, n + 1, s));
  in_slices(n, kWave, [&](int64_t f, int64_t c) {
    csr_permute_rows<<<cdiv(c * kWave, kTB), kTB, 0, s>>>(
        A.rowptr, A.colidx, static_cast<const double*>(A.val), rec.get(), nrp, n, ncol, nval, f);
  });
  BSP_HIP(hipGetLastError());
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

}  // namespace bsp

using namespace bsp;

namespace {
bsp_status check_csr_f64(const bsp_csr* A) {
  BSP_REQUIRE(A, "null CSR descriptor");
  BSP_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, "negative CSR size");
  if (!sizes_fit(A->nrows, A->ncols, A->nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(A->nrows == 0 || A->rowptr, "null rowptr");
  BSP_REQUIRE(A->nnz == 0 || (A->colidx && A->val), "null colidx/val");
  BSP_REQUIRE(A->dtype == BSP_F64, "FP64 values required");
  return BSP_OK;
}
}  // namespace

extern "C" {

bsp_status bsp_bsr_build(const bsp_csr* A, int tm, int tk, bsp_stream s, bsp_bsr* out) {
  BSP_TRY(check_csr_f64(A));
  BSP_REQUIRE(out, "null output handle");
  BSP_REQUIRE(tm >= 1 && tk >= 1 && tm <= 64 && tk <= 64, "block size must be in 1..64");
  auto h = std::make_unique<bsp_bsr_s>();
  BSP_TRY(build_bsr(*A, tm, tk, (hipStream_t)s, h.get()));
  if (h->v.nblocks * (int64_t)tm * tk >= (int64_t(1) << 40))
    return fail(BSP_ERR_TOO_LARGE, "BSR too large");
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_bsr_view_get(bsp_bsr h, bsp_bsr_view* v) {
  BSP_REQUIRE(h && v, "null handle");
  *v = h->v;
  return BSP_OK;
}

bsp_status bsp_bsr_spmv(bsp_bsr h, int kernel, const double* x, double* y, bsp_stream s) {
  BSP_REQUIRE(h, "null handle");
  const bsp_bsr_view& v = h->v;
  BSP_REQUIRE((v.ncols == 0 || x) && (v.nrows == 0 || y), "null x/y");
  if (v.brows == 0) return BSP_OK;
  if (kernel == BSP_BSR_AUTO) kernel = BSP_BSR_VALU;
  hipStream_t st = (hipStream_t)s;
  if (kernel == BSP_BSR_VALU) {
    auto go = [&](auto kernel) {
      in_slices(v.brows, kWave, [&](int64_t f, int64_t c) {
        kernel<<<cdiv(c * kWave, kTB), kTB, 0, st>>>(v.rowptr, v.colidx, v.val, x, y, v.brows,
                                                     v.nrows, v.ncols, v.tm, v.tk, f);
      });
    };
    switch (v.tk) {  // the widths the benchmarks use, unrolled; any other at run time
      case 3: go(bsr_valu_f64<3>); break;
      case 4: go(bsr_valu_f64<4>); break;
      case 8: go(bsr_valu_f64<8>); break;
      case 16: go(bsr_valu_f64<16>); break;
      default: go(bsr_valu_f64<0>); break;
    }
  } else if (kernel == BSP_BSR_MFMA) {
    const bool b4 = v.tm == 4 && v.tk == 4, b8 = v.tm == 8 && v.tk == 8;
    if (!b4 && !b8) return fail(BSP_ERR_UNSUPPORTED, "MFMA BSR kernel needs 4x4 or 8x8 blocks");
    in_slices(v.brows, kWave, [&](int64_t f, int64_t c) {
      if (b4 && ((uintptr_t)x & 15) == 0)
        bsr4_mfma_f64<true><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(v.rowptr, v.colidx, v.val, x, y,
                                                                  v.brows, v.nrows, v.ncols, f);
      else if (b4)
        bsr4_mfma_f64<false><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(v.rowptr, v.colidx, v.val, x,
                                                                   y, v.brows, v.nrows, v.ncols, f);
      else if (((uintptr_t)x & 15) == 0)
        bsr8_mfma_f64<true><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(v.rowptr, v.colidx, v.val, x, y,
                                                                  v.brows, v.nrows, v.ncols, f);
      else
        bsr8_mfma_f64<false><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(v.rowptr, v.colidx, v.val, x,
                                                                   y, v.brows, v.nrows, v.ncols, f);
    });
  } else if (kernel == BSP_BSR_MFMA_BALANCED || kernel == BSP_BSR_VALU_BALANCED) {
    if (v.tm != 4 || v.tk != 4)
      return fail(BSP_ERR_UNSUPPORTED, "balanced BSR kernels need 4x4 blocks");
    if (kernel == BSP_BSR_MFMA_BALANCED)  // a wave per chunk
      in_slices(v.warpnum, kWave, [&](int64_t f, int64_t c) {
        if (((uintptr_t)x & 15) == 0)
          bsr4_bal_mfma_f64<true><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
              h->ctab.get(), v.warpnum, v.colidx, v.val, x, y, h->part.get(), v.nrows, v.ncols, f);
        else
          bsr4_bal_mfma_f64<false><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
              h->ctab.get(), v.warpnum, v.colidx, v.val, x, y, h->part.get(), v.nrows, v.ncols, f);
      });
    else  // a wave per four chunks
      in_slices(cdiv(v.warpnum, 4), kWave, [&](int64_t f, int64_t c) {
        if (((uintptr_t)x & 15) == 0)
          bsr4_bal_valu_f64<true><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
              h->ctab.get(), v.warpnum, v.colidx, v.val, v.map, x, y, h->part.get(), v.nrows,
              v.ncols, f);
        else
          bsr4_bal_valu_f64<false><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
              h->ctab.get(), v.warpnum, v.colidx, v.val, v.map, x, y, h->part.get(), v.nrows,
              v.ncols, f);
      });
    if (h->nmulti > 0)
      bsr4_bal_combine<<<cdiv(h->nmulti * 4, kTB), kTB, 0, st>>>(
          h->multi.get(), h->nmulti, v.rowptr_by_warp, h->part.get(), v.nrows, y);
  } else {
    return fail(BSP_ERR_INVALID, "unknown BSR kernel");
  }
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_bsr_destroy(bsp_bsr h) {
  delete h;
  return BSP_OK;
}

bsp_status bsp_dbsr_build(const bsp_csr* A, bsp_stream s, bsp_dbsr* out) {
  BSP_TRY(check_csr_f64(A));
  BSP_REQUIRE(out, "null output handle");
  auto h = std::make_unique<bsp_dbsr_s>();
  BSP_TRY(build_dbsr(*A, true, (hipStream_t)s, h.get()));
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_dbsr_view_get(bsp_dbsr h, bsp_dbsr_view* v) {
  BSP_REQUIRE(h && v, "null handle");
  *v = h->v;
  return BSP_OK;
}

bsp_status bsp_dbsr_spmv(bsp_dbsr h, const double* x, double* y, int sorted_order, bsp_stream s) {
  return bsp_dbsr_spmv_ex(h, BSP_DBSR_VALU, x, y, sorted_order, s);
}

bsp_status bsp_dbsr_spmv_ex(bsp_dbsr h, int kernel, const double* x, double* y, int sorted_order,
                            bsp_stream s) {
  BSP_REQUIRE(h, "null handle");
  BSP_REQUIRE(kernel == BSP_DBSR_VALU || kernel == BSP_DBSR_MFMA, "unknown DBSR kernel");
  const bsp_dbsr_view& v = h->v;
  BSP_REQUIRE((v.ncols == 0 || x) && (v.nrows == 0 || y), "null x/y");
  if (v.brows == 0) return BSP_OK;
  const int64_t waves = (h->narrow + 1) / 2 + (v.brows - h->narrow);
  if (kernel == BSP_DBSR_MFMA) {
    in_slices(waves, kWave, [&](int64_t f, int64_t c) {
      dbsr_spmv_f64_mfma<<<cdiv(c * kWave, kTB), kTB, 0, (hipStream_t)s>>>(
          v.blcptr, v.colidx, v.val, v.recorder, x, y, v.brows, v.nrows, h->narrow, sorted_order,
          f);
    });
    BSP_HIP(hipGetLastError());
    return BSP_OK;
  }
  in_slices(waves, kWave, [&](int64_t f, int64_t c) {
    dbsr_spmv_f64<<<cdiv(c * kWave, kTB), kTB, 0, (hipStream_t)s>>>(
        v.blcptr, v.colidx, v.val, v.recorder, x, y, v.brows, v.nrows, h->narrow, sorted_order, f);
  });
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_dbsr_destroy(bsp_dbsr h) {
  delete h;
  return BSP_OK;
}

bsp_status bsp_dbsr_from_host(const bsp_dbsrMat* m, bsp_stream s, bsp_dbsr* out) {
  BSP_REQUIRE(m && out, "null argument");
  BSP_REQUIRE(m->row >= 0 && m->blc_row == (m->row + 3) / 4 && m->blocknnz >= 0 &&
                  m->blocknnz % 16 == 0 && m->warpnum >= 0,
              "inconsistent dbsrMat sizes");
  BSP_REQUIRE(m->blcPtr &&
                  (m->blocknnz == 0 || (m->colIdx && m->blcVal && m->blcMap)),
              "null dbsrMat array");
  // The kernel gathers x through every stored column index (padding too)
  // and writes y through recorderMap, so both are checked on the host
  // before anything is uploaded.
  BSP_REQUIRE(m->col >= 0 && m->nnz >= 0, "negative dbsrMat size");
  if (!sizes_fit(m->row, m->col, m->nnz) || !sizes_fit(0, 0, m->blocknnz))
    return fail(BSP_ERR_TOO_LARGE, "dbsrMat sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(m->blcPtr[0] == 0, "blcPtr[0] must be 0");
  for (int64_t g = 0; g < m->blc_row; ++g)
    if (m->blcPtr[g + 1] < m->blcPtr[g])
      return fail(BSP_ERR_INVALID, "blcPtr decreases at block row " + std::to_string(g));
  BSP_REQUIRE(m->blcPtr[m->blc_row] == m->blocknnz / 16, "blcPtr[blc_row] must equal blocknnz/16");
  for (int64_t k = 0; k < m->blocknnz; ++k)
    if (m->colIdx[k] < 0 || m->colIdx[k] >= m->col)
      return fail(BSP_ERR_INVALID, "colIdx out of range at slot " + std::to_string(k));
  if (m->recorderMap) {  // null: rows in their own order (csrTodbsr leaves it unset)
    std::vector<char> seen(m->row, 0);
    for (int64_t i = 0; i < m->row; ++i) {
      const int32_t r = m->recorderMap[i];
      if (r < 0 || r >= m->row || seen[r])
        return fail(BSP_ERR_INVALID, "recorderMap is not a permutation of the rows");
      seen[r] = 1;
    }
  }
  hipStream_t st = (hipStream_t)s;
  auto h = std::make_unique<bsp_dbsr_s>();
  bsp_dbsr_view& v = h->v;
  v.nrows = m->row;
  v.ncols = m->col;
  v.nnz = m->nnz;
  v.brows = m->blc_row;
  v.nblocks = m->blocknnz / 16;
  v.warpnum = m->warpnum;
  BSP_TRY(h->recorder.alloc(v.nrows));
  BSP_TRY(h->blcptr.alloc(v.brows + 1));
  BSP_TRY(h->colidx.alloc(v.nblocks * 16));
  BSP_TRY(h->val.alloc(v.nblocks * 16));
  BSP_TRY(h->map.alloc(v.nblocks));
  BSP_TRY(h->prefix.alloc(v.warpnum + 1));
  if (v.nrows && m->recorderMap) {
    BSP_HIP(hipMemcpyAsync(h->recorder.get(), m->recorderMap, sizeof(int32_t) * v.nrows,
                           hipMemcpyHostToDevice, st));
  } else if (v.nrows) {
    std::vector<int32_t> id(v.nrows);
    for (int64_t i = 0; i < v.nrows; ++i) id[i] = (int32_t)i;
    BSP_HIP(hipMemcpy(h->recorder.get(), id.data(), sizeof(int32_t) * v.nrows,
                      hipMemcpyHostToDevice));
  }
  BSP_HIP(hipMemcpyAsync(h->blcptr.get(), m->blcPtr, sizeof(int32_t) * (v.brows + 1),
                         hipMemcpyHostToDevice, st));
  if (v.nblocks) {
    BSP_HIP(hipMemcpyAsync(h->colidx.get(), m->colIdx, sizeof(int32_t) * v.nblocks * 16,
                           hipMemcpyHostToDevice, st));
    BSP_HIP(hipMemcpyAsync(h->val.get(), m->blcVal, sizeof(double) * v.nblocks * 16,
                           hipMemcpyHostToDevice, st));
    BSP_HIP(hipMemcpyAsync(h->map.get(), m->blcMap, sizeof(uint16_t) * v.nblocks,
                           hipMemcpyHostToDevice, st));
  }
  if (m->prefix)
    BSP_HIP(hipMemcpyAsync(h->prefix.get(), m->prefix, sizeof(int32_t) * (v.warpnum + 1),
                           hipMemcpyHostToDevice, st));
  BSP_HIP(hipStreamSynchronize(st));
  v.recorder = h->recorder.get();
  v.blcptr = h->blcptr.get();
  v.colidx = h->colidx.get();
  v.val = h->val.get();
  v.map = h->map.get();
  v.prefix = h->prefix.get();
  BSP_TRY(dbsr_count_narrow(h.get(), st));
  *out = h.release();
  return BSP_OK;
}

}  // extern "C"

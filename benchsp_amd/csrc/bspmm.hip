// Block-sparse SpMM on FP16/BF16 MFMA with FP32 accumulation (the SMaT
// kernel of src/smat.cu:22-91 re-designed for gfx950), with its backward.
//
// Format (built on the GPU from a 16x16 BSR with FP64 values):
//   scale[b]  = 2^e with max|A_b| in [2^(e-1), 2^e): a power of two computed
//               in FP64, so A_b / scale[b] is exact before the one rounding
//               to FP16/BF16 and the values fill (-1, 1);
//   hv[b]     = A_b / scale[b], 16x16 row-major, FP16 or BF16;
//   hvt[b]    = the same block transposed (for A^T products);
//   a transposed block index (blocks of each block column).
//
// MFMA operand layouts (v_mfma_f32_16x16x16{f16,bf16_1k}, x32 for K = 32):
//   A (16xK): lane l holds A[l%16][E*(l/16) .. +E-1]   (E = K/4)
//   B (Kx16): lane l holds B[E*(l/16) .. +E-1][l%16]
//   D (16x16 FP32): lane l holds D[4*(l/16)+v][l%16] in acc[v]
// Both operands are read k-contiguous: A from hv (row-major), B from a
// per-call panel copy of the dense operand that stores every 16-row slab in
// MFMA fragment order, so a wave's operand load is 512 contiguous bytes.
//
//   forward  C[16 rows of br, :] = sum_b scale_b * hv_b * B[16 rows of bc_b, :]
//   dB       dB[16 rows of bc, :] = sum_b scale_b * hvt_b * dC[16 rows of br_b, :]
//   dvals    dA_b = dC[16 rows of br, :] * B[16 rows of bc, :]^T   (K = N)
// One workgroup per block row (block column for dB); each wave owns a set of
// 16-column tiles and adds the row's blocks in index order, so results are
// deterministic.
#include <cmath>
#include <memory>

#include "bsp_hip.hpp"

struct bsp_bspmm_s {
  int64_t M = 0, K = 0, brows = 0, bcols = 0, nblocks = 0;
  bsp_dtype ht = BSP_F16;
  const int32_t* browptr = nullptr;  // borrowed from the BSR
  const int32_t* bcol = nullptr;
  bsp::DevBuf<float> scale;
  bsp::DevBuf<uint16_t> hv, hvt;
  bsp::DevBuf<int32_t> tcolptr, tperm, tbrow;
  bsp::DevBuf<uint16_t> work;  // dense operand in panel layout
  int64_t work_cap = 0;
};

namespace bsp {
namespace {

constexpr int kTB = 256;
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned u2 __attribute__((ext_vector_type(2)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef short s4 __attribute__((ext_vector_type(4)));
typedef __bf16 b8 __attribute__((ext_vector_type(8)));

template <bool BF>
__device__ __forceinline__ f4 mma16(u2 a, u2 b, f4 c) {
  if constexpr (BF)
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(s4, a),
                                                     __builtin_bit_cast(s4, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(h4, a),
                                                 __builtin_bit_cast(h4, b), c, 0, 0, 0);
}

template <bool BF>
__device__ __forceinline__ f4 mma32(u4 a, u4 b, f4 c) {
  if constexpr (BF)
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(b8, a),
                                                   __builtin_bit_cast(b8, b), c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(h8, a),
                                                  __builtin_bit_cast(h8, b), c, 0, 0, 0);
}

template <bool BF>
__device__ __forceinline__ uint16_t to_half(double v) {
  if constexpr (BF)
    return __builtin_bit_cast(uint16_t, (__bf16)v);
  else
    return __builtin_bit_cast(uint16_t, (_Float16)v);
}

// wave per block: FP64 max|a|, power-of-two scale, one rounding to half
template <bool BF>
__global__ __launch_bounds__(kTB) void block_scale_convert(const double* __restrict__ bval,
                                                           int64_t nblocks, float* __restrict__ scale,
                                                           uint16_t* __restrict__ hv,
                                                           uint16_t* __restrict__ hvt,
                                                           int64_t first) {
  const int64_t b = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (b >= nblocks) return;
  const double* a = bval + b * 256;
  double v[4], m = 0.0;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    v[q] = a[lane + 64 * q];
    m = fmax(m, fabs(v[q]));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  int e = 0;
  if (m > 0.0) frexp(m, &e);
  const double s = ldexp(1.0, e);
  if (lane == 0) scale[b] = (float)s;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int idx = lane + 64 * q, i = idx / 16, j = idx % 16;
    const uint16_t h = to_half<BF>(v[q] / s);
    hv[b * 256 + idx] = h;
    hvt[b * 256 + j * 16 + i] = h;
  }
}

__global__ void block_rows_of(const int32_t* __restrict__ browptr, int64_t brows,
                              int32_t* __restrict__ brow) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= brows) return;
  for (int b = browptr[r]; b < browptr[r + 1]; ++b) brow[b] = (int32_t)r;
}

__global__ void tkeys(const int32_t* __restrict__ bcol, const int32_t* __restrict__ brow,
                      int64_t nblocks, int64_t brows, int64_t* __restrict__ key,
                      int32_t* __restrict__ idx, int32_t* __restrict__ cnt) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  key[b] = (int64_t)bcol[b] * brows + brow[b];
  idx[b] = (int32_t)b;
  atomicAdd(&cnt[bcol[b]], 1);
}

__global__ void gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ perm,
                           int64_t n, int32_t* __restrict__ dst) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) dst[k] = src[perm[k]];
}

// Panel layout of a dense R x N operand (row-major, N = 16*NT): for every
// 16-row slab o and 16-column tile t the MFMA B-operand fragment in lane
// order, P[((o*NT + t)*64 + l)*4 + e] = src[16o + 4(l/16) + e][16t + l%16]
// (rows >= R read as zero), so a wave's fragment load is 512 contiguous bytes.
template <int NT>
__global__ __launch_bounds__(kTB) void panelize(const uint16_t* __restrict__ src, int64_t R,
                                                uint16_t* __restrict__ dst, int64_t first) {
  constexpr int N = NT * 16;
  __shared__ uint16_t t[16][N + 2];
  const int64_t o = first + blockIdx.x;
  for (int idx = threadIdx.x; idx < 16 * N / 8; idx += kTB) {
    const int row = idx / (N / 8), c8 = (idx % (N / 8)) * 8;
    const int64_t r = o * 16 + row;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r < R) v = *reinterpret_cast<const uint4*>(src + r * N + c8);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      t[row][c8 + 2 * j] = (uint16_t)(w[j] & 0xffff);
      t[row][c8 + 2 * j + 1] = (uint16_t)(w[j] >> 16);
    }
  }
  __syncthreads();
  for (int idx = threadIdx.x; idx < NT * 32; idx += kTB) {
    uint32_t w[4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int f = 2 * idx + h, tt = f / 64, l = f % 64, i = l % 16, kg = l / 16;
      w[2 * h] = (uint32_t)t[4 * kg][tt * 16 + i] | ((uint32_t)t[4 * kg + 1][tt * 16 + i] << 16);
      w[2 * h + 1] = (uint32_t)t[4 * kg + 2][tt * 16 + i] | ((uint32_t)t[4 * kg + 3][tt * 16 + i] << 16);
    }
    *reinterpret_cast<uint4*>(dst + (o * NT * 64 + 2 * idx) * 4) = make_uint4(w[0], w[1], w[2], w[3]);
  }
}

// forward (TRANS = false) and dB (TRANS = true). One workgroup per block
// row (block column for dB) with NW waves; wave w owns column tiles
// [w*TPW, (w+1)*TPW) across all of the row's blocks, so its accumulators are
// final (no cross-wave reduction) and blocks are added in index order. The
// row's blocks pass through LDS in chunks of 8*NW: their indices and scales
// first, then their A fragments in lane order (each block read from HBM once
// per workgroup rather than once per wave); each wave then streams its panel
// fragments, U blocks in flight.
template <int NT>
struct RowsShape {
  static constexpr int NW = NT < 8 ? NT : 8;  // waves
  static constexpr int TPW = NT / NW;         // 16-column tiles per wave
  static constexpr int CH = 8 * NW;           // blocks per LDS chunk
};

template <int NT, bool BF, bool TRANS>
__global__ __launch_bounds__(512) void bspmm_rows(const int32_t* __restrict__ ptr,
                                                  const int32_t* __restrict__ perm,
                                                  const int32_t* __restrict__ other,
                                                  const uint16_t* __restrict__ blocks,
                                                  const float* __restrict__ scale,
                                                  const uint16_t* __restrict__ panel,
                                                  float* __restrict__ out, int64_t out_rows,
                                                  int64_t nouter, int64_t first) {
  constexpr int NW = RowsShape<NT>::NW, TPW = RowsShape<NT>::TPW, CH = RowsShape<NT>::CH;
  constexpr int U = 16 / TPW;
  __shared__ u2 sa[CH][kWave];
  __shared__ int32_t sb[CH], so[CH];
  __shared__ float ss[CH];
  const int64_t r = first + blockIdx.x;
  if (r >= nouter) return;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  const int i = lane & 15, kg = lane >> 4;
  const uint16_t* pw = panel + ((int64_t)w * TPW * 64 + lane) * 4;
  f4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  const int k0 = ptr[r], k1 = ptr[r + 1];
  for (int c0 = k0; c0 < k1; c0 += CH) {
    const int cn = min(CH, k1 - c0);
    if (tid < cn) {
      const int b = TRANS ? perm[c0 + tid] : c0 + tid;
      sb[tid] = b;
      so[tid] = other[c0 + tid];
      ss[tid] = scale[b];
    }
    __syncthreads();
    // unconditional loads (slots past cn repeat the last block and are never
    // read), so all CH/NW of a wave's loads are in flight together
    u2 stage[CH / NW];
#pragma unroll
    for (int q = 0; q < CH / NW; ++q)
      stage[q] = *reinterpret_cast<const u2*>(blocks + (int64_t)sb[min(w + q * NW, cn - 1)] * 256 +
                                              i * 16 + 4 * kg);
#pragma unroll
    for (int q = 0; q < CH / NW; ++q) sa[w + q * NW][lane] = stage[q];
    __syncthreads();
    for (int j = 0; j < cn; j += U) {
      u2 bo[U][TPW];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int64_t o = so[min(j + u, cn - 1)];
#pragma unroll
        for (int t = 0; t < TPW; ++t)
          bo[u][t] = *reinterpret_cast<const u2*>(pw + (o * NT + t) * 256);
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (j + u < cn) {
          const u2 a = sa[j + u][lane];
          const float sc = ss[j + u];
#pragma unroll
          for (int t = 0; t < TPW; ++t) acc[t] += sc * mma16<BF>(a, bo[u][t], f4{0.f, 0.f, 0.f, 0.f});
        }
    }
    __syncthreads();  // the next chunk overwrites the LDS arrays
  }
  constexpr int N = NT * 16;
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = r * 16 + 4 * kg + v;
      if (row < out_rows) out[row * N + (w * TPW + t) * 16 + i] = acc[t][v];
    }
}

// dvals: dA_b = dC[16 rows of br, :] * B[16 rows of bc, :]^T (K = N), one
// workgroup of kDvWaves waves per block row, each wave every kDvWaves-th block.
constexpr int kDvWaves = 4;

template <int NT, bool BF>
__global__ __launch_bounds__(kDvWaves * 64) void bspmm_dvals(const int32_t* __restrict__ browptr,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint16_t* __restrict__ dC,
                                                             int64_t M,
                                                             const uint16_t* __restrict__ B,
                                                             int64_t K, float* __restrict__ dvals,
                                                             int64_t brows, int64_t first) {
  constexpr int N = NT * 16, S = N / 32;
  const int64_t br = first + blockIdx.x;
  if (br >= brows) return;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int i = lane & 15, kg = lane >> 4;
  u4 a[S];
  const int64_t arow = br * 16 + i;
  // clamped unconditional loads, rows past M (K) zeroed afterwards: no branch
  // splits a batch of loads
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const u4 v = *reinterpret_cast<const u4*>(dC + min(arow, M - 1) * N + 32 * s + 8 * kg);
    a[s] = arow < M ? v : u4{0, 0, 0, 0};
  }
  // two blocks per step (b and b + kDvWaves), their B loads in flight together
  const int e = browptr[br + 1];
  for (int b = browptr[br] + w; b < e; b += 2 * kDvWaves) {
    const int b2 = b + kDvWaves;
    const bool two = b2 < e;
    const int64_t brow[2] = {(int64_t)bcol[b] * 16 + i, (int64_t)bcol[two ? b2 : b] * 16 + i};
    u4 bo[2][S];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const u4 v = *reinterpret_cast<const u4*>(B + min(brow[h], K - 1) * N + 32 * s + 8 * kg);
        bo[h][s] = brow[h] < K ? v : u4{0, 0, 0, 0};
      }
    f4 acc[2] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}};
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
      for (int s = 0; s < S; ++s) acc[h] = mma32<BF>(a[s], bo[h][s], acc[h]);
#pragma unroll
    for (int v = 0; v < 4; ++v) dvals[(int64_t)b * 256 + (4 * kg + v) * 16 + i] = acc[0][v];
    if (two)
#pragma unroll
      for (int v = 0; v < 4; ++v) dvals[(int64_t)b2 * 256 + (4 * kg + v) * 16 + i] = acc[1][v];
  }
}

template <bool BF, bool TRANS>
static bsp_status launch_rows(bsp_bspmm_s* h, int64_t N, const uint16_t* panel, float* out,
                              int64_t out_rows, hipStream_t s) {
  const int64_t nouter = TRANS ? h->bcols : h->brows;
  const int32_t* ptr = TRANS ? h->tcolptr.get() : h->browptr;
  const int32_t* perm = TRANS ? h->tperm.get() : nullptr;
  const int32_t* other = TRANS ? h->tbrow.get() : h->bcol;
  const uint16_t* blocks = TRANS ? h->hvt.get() : h->hv.get();
  if (nouter == 0) return BSP_OK;
#define BSPMM_CASE(NT)                                                                           \
  case NT:                                                                                       \
    in_slices(nouter, RowsShape<NT>::NW * kWave, [&](int64_t f, int64_t c) {                     \
      bspmm_rows<NT, BF, TRANS><<<c, RowsShape<NT>::NW * kWave, 0, s>>>(                         \
          ptr, perm, other, blocks, h->scale.get(), panel, out, out_rows, nouter, f);            \
    });                                                                                          \
    break;
  switch (N / 16) {
    BSPMM_CASE(2)
    BSPMM_CASE(4)
    BSPMM_CASE(8)
    BSPMM_CASE(16)
    default:
      return fail(BSP_ERR_UNSUPPORTED, "block SpMM needs N in {32, 64, 128, 256}");
  }
#undef BSPMM_CASE
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

template <bool BF>
static bsp_status launch_dvals(bsp_bspmm_s* h, int64_t N, const uint16_t* dC, const uint16_t* B,
                               float* dvals, hipStream_t s) {
  if (h->brows == 0) return BSP_OK;
#define DV_CASE(NT)                                                                             \
  case NT:                                                                                      \
    in_slices(h->brows, kDvWaves * 64, [&](int64_t f, int64_t c) {                             \
      bspmm_dvals<NT, BF><<<c, kDvWaves * 64, 0, s>>>(h->browptr, h->bcol, dC, h->M, B, h->K,  \
                                                      dvals, h->brows, f);                      \
    });                                                                                         \
    break;
  switch (N / 16) {
    DV_CASE(2)
    DV_CASE(4)
    DV_CASE(8)
    DV_CASE(16)
    default:
      return fail(BSP_ERR_UNSUPPORTED, "block SpMM needs N in {32, 64, 128, 256}");
  }
#undef DV_CASE
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

// The dense operand (R x N) in panel layout, nslab 16-row slabs, in h->work.
static bsp_status paneled(bsp_bspmm_s* h, const uint16_t* src, int64_t R, int64_t nslab,
                          int64_t N, hipStream_t s) {
  if (nslab * 16 * N > h->work_cap) {
    BSP_HIP(hipStreamSynchronize(s));
    BSP_TRY(h->work.alloc(nslab * 16 * N));
    h->work_cap = nslab * 16 * N;
  }
  if (nslab == 0) return BSP_OK;
  auto go = [&](auto kernel) {
    in_slices(nslab, kTB, [&](int64_t f, int64_t c) {
      kernel<<<c, kTB, 0, s>>>(src, R, h->work.get(), f);
    });
  };
  switch (N / 16) {
    case 2: go(panelize<2>); break;
    case 4: go(panelize<4>); break;
    case 8: go(panelize<8>); break;
    case 16: go(panelize<16>); break;
    default: return fail(BSP_ERR_UNSUPPORTED, "block SpMM needs N in {32, 64, 128, 256}");
  }
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

}  // namespace
}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_bspmm_build(bsp_bsr A, bsp_dtype half_type, bsp_stream s, bsp_bspmm* out) {
  BSP_REQUIRE(A && out, "null argument");
  BSP_REQUIRE(half_type == BSP_F16 || half_type == BSP_BF16, "half type must be FP16 or BF16");
  bsp_bsr_view v;
  BSP_TRY(bsp_bsr_view_get(A, &v));
  BSP_REQUIRE(v.tm == 16 && v.tk == 16, "block SpMM needs 16x16 blocks");
  hipStream_t st = (hipStream_t)s;
  auto h = std::make_unique<bsp_bspmm_s>();
  h->M = v.nrows;
  h->K = v.ncols;
  h->brows = v.brows;
  h->bcols = v.bcols;
  h->nblocks = v.nblocks;
  h->ht = half_type;
  h->browptr = v.rowptr;
  h->bcol = v.colidx;
  const int64_t nb = v.nblocks;
  BSP_TRY(h->scale.alloc(nb));
  BSP_TRY(h->hv.alloc(nb * 256));
  BSP_TRY(h->hvt.alloc(nb * 256));
  if (nb > 0) {
    in_slices(nb, kWave, [&](int64_t f, int64_t c) {
      if (half_type == BSP_BF16)
        block_scale_convert<true><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
            v.val, nb, h->scale.get(), h->hv.get(), h->hvt.get(), f);
      else
        block_scale_convert<false><<<cdiv(c * kWave, kTB), kTB, 0, st>>>(
            v.val, nb, h->scale.get(), h->hv.get(), h->hvt.get(), f);
    });
    BSP_HIP(hipGetLastError());
  }
  // transposed block index: blocks ordered by (block column, block row)
  DevBuf<int32_t> brow, idx, cnt;
  DevBuf<int64_t> key, skey;
  BSP_TRY(brow.alloc(nb));
  BSP_TRY(idx.alloc(nb));
  BSP_TRY(key.alloc(nb));
  BSP_TRY(skey.alloc(nb));
  BSP_TRY(cnt.alloc(v.bcols + 1));
  BSP_TRY(h->tperm.alloc(nb));
  BSP_TRY(h->tbrow.alloc(nb));
  BSP_TRY(h->tcolptr.alloc(v.bcols + 1));
  BSP_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int32_t) * (v.bcols + 1), st));
  if (nb > 0) {
    block_rows_of<<<cdiv(v.brows, kTB), kTB, 0, st>>>(v.rowptr, v.brows, brow.get());
    tkeys<<<cdiv(nb, kTB), kTB, 0, st>>>(v.colidx, brow.get(), nb, v.brows, key.get(), idx.get(),
                                         cnt.get());
    BSP_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 62 && (int64_t(1) << bits) <= v.bcols * v.brows) ++bits;
    BSP_TRY(sort_pairs_i64key(key.get(), skey.get(), idx.get(), h->tperm.get(), nb, bits, st));
    gather_i32<<<cdiv(nb, kTB), kTB, 0, st>>>(brow.get(), h->tperm.get(), nb, h->tbrow.get());
    BSP_HIP(hipGetLastError());
  }
  BSP_TRY(exclusive_scan_i32(cnt.get(), h->tcolptr.get(), v.bcols + 1, st));
  BSP_HIP(hipStreamSynchronize(st));
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_bspmm_view_get(bsp_bspmm h, const float** scale, const uint16_t** hv,
                              const uint16_t** hvt, int64_t* nblocks) {
  BSP_REQUIRE(h, "null handle");
  if (scale) *scale = h->scale.get();
  if (hv) *hv = h->hv.get();
  if (hvt) *hvt = h->hvt.get();
  if (nblocks) *nblocks = h->nblocks;
  return BSP_OK;
}

bsp_status bsp_bspmm_fwd(bsp_bspmm h, const void* B, int64_t N, float* Cout, bsp_stream s) {
  BSP_REQUIRE(h && B && Cout, "null argument");
  hipStream_t st = (hipStream_t)s;
  BSP_REQUIRE(N == 32 || N == 64 || N == 128 || N == 256, "block SpMM needs N in {32, 64, 128, 256}");
  // B in panel layout: each block column's 16 x N slab in MFMA operand order
  BSP_TRY(paneled(h, static_cast<const uint16_t*>(B), h->K, h->bcols, N, st));
  if (h->ht == BSP_BF16) return launch_rows<true, false>(h, N, h->work.get(), Cout, h->M, st);
  return launch_rows<false, false>(h, N, h->work.get(), Cout, h->M, st);
}

bsp_status bsp_bspmm_bwd(bsp_bspmm h, const void* B, const void* dC, int64_t N, float* dB,
                         float* dvals, bsp_stream s) {
  BSP_REQUIRE(h && B && dC, "null argument");
  hipStream_t st = (hipStream_t)s;
  const bool bf = h->ht == BSP_BF16;
  if (dvals) {
    if (bf)
      BSP_TRY(launch_dvals<true>(h, N, (const uint16_t*)dC, (const uint16_t*)B, dvals, st));
    else
      BSP_TRY(launch_dvals<false>(h, N, (const uint16_t*)dC, (const uint16_t*)B, dvals, st));
  }
  if (dB) {
    BSP_REQUIRE(N == 32 || N == 64 || N == 128 || N == 256,
                "block SpMM needs N in {32, 64, 128, 256}");
    BSP_TRY(paneled(h, static_cast<const uint16_t*>(dC), h->M, h->brows, N, st));
    if (bf)
      BSP_TRY((launch_rows<true, true>(h, N, h->work.get(), dB, h->K, st)));
    else
      BSP_TRY((launch_rows<false, true>(h, N, h->work.get(), dB, h->K, st)));
  }
  return BSP_OK;
}

bsp_status bsp_bspmm_destroy(bsp_bspmm h) {
  delete h;
  return BSP_OK;
}

}  // extern "C"

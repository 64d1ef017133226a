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
// Every operand load of a wave is one contiguous run of 1 KB (16 bytes per
// lane) or 512 bytes: the number of cache lines a wave instruction touches,
// not where the lines live, sets the cost of these gather-heavy products
// (a layout whose wave loads touch 16 or 32 half-used lines ran at half the
// speed). A is read from hv (row-major blocks); the dense operand goes
// through a per-call panel copy in MFMA fragment order.
//
//   forward  C[16 rows of br, :] = sum_b scale_b * hv_b * B[16 rows of bc_b, :]
//   dB       dB[16 rows of bc, :] = sum_b scale_b * hvt_b * dC[16 rows of br_b, :]
//   dvals    dA_b = dC[16 rows of br, :] * B[16 rows of bc, :]^T   (K = N)
// The two products run a workgroup per block row (block column for dB) whose
// waves take every WB-th block of the row; their partial tiles are added in
// wave order through LDS, so results are deterministic.
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
  bsp::DevBuf<int32_t> tcolptr, tperm, tbrow;  // blocks by (block column, block row)
  bsp::DevBuf<uint16_t> work;   // pair panel: B (forward) or dC (dB)
  bsp::DevBuf<uint16_t> work2;  // row panel of B (dvals)
  int64_t work_cap = 0, work2_cap = 0;
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

// Pair panel of a dense R x N operand (row-major, N = 16*NT, NT even): for
// every 16-row slab o and pair of 16-column tiles (2p, 2p+1), each lane's
// two B-operand fragments side by side,
//   P[((o*NT/2 + p)*64 + l)*8 + 4h + e] = src[16o + 4(l/16) + e][16(2p+h) + l%16]
// (rows >= R read as zero), so one wave load of 16 bytes per lane is 1 KB
// contiguous and feeds two MFMAs.
template <int NT>
__device__ __forceinline__ void pair_panel_slab(const uint16_t* __restrict__ src, int64_t R,
                                                uint16_t* __restrict__ dst, int64_t o) {
  constexpr int N = NT * 16;
  __shared__ uint16_t t[16][N + 2];
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
  for (int idx = threadIdx.x; idx < NT / 2 * kWave; idx += kTB) {
    const int p = idx / kWave, l = idx % kWave, i = l % 16, kg = l / 16;
    uint32_t w[4];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = (2 * p + h) * 16 + i;
      w[2 * h] = (uint32_t)t[4 * kg][c] | ((uint32_t)t[4 * kg + 1][c] << 16);
      w[2 * h + 1] = (uint32_t)t[4 * kg + 2][c] | ((uint32_t)t[4 * kg + 3][c] << 16);
    }
    *reinterpret_cast<uint4*>(dst + (o * (NT / 2) * kWave + idx) * 8) =
        make_uint4(w[0], w[1], w[2], w[3]);
  }
}

template <int NT>
__global__ __launch_bounds__(kTB) void panelize(const uint16_t* __restrict__ src, int64_t R,
                                                uint16_t* __restrict__ dst, int64_t first) {
  pair_panel_slab<NT>(src, R, dst, first + blockIdx.x);
}

// Row panel of a dense R x N operand, the A operand of the dvals product
// (its rows against the block's K = N): for every 16-row slab o and K step s
// (S = N/32 of them), lane l's 8 values
//   Q[((o*S + s)*64 + l)*8 + e] = src[16o + l%16][32s + 8(l/16) + e]
// (rows >= R read as zero): a wave's load is again 1 KB contiguous, where the
// row-major operand gives 16 half-used lines per load.
template <int NT>
__device__ __forceinline__ void row_panel_slab(const uint16_t* __restrict__ src, int64_t R,
                                               uint16_t* __restrict__ dst, int64_t o) {
  constexpr int N = NT * 16, S = N / 32;
  for (int idx = threadIdx.x; idx < S * kWave; idx += kTB) {
    const int s = idx / kWave, l = idx % kWave;
    const int64_t r = o * 16 + l % 16;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r < R) v = *reinterpret_cast<const uint4*>(src + r * N + 32 * s + 8 * (l / 16));
    *reinterpret_cast<uint4*>(dst + (o * S * kWave + idx) * 8) = v;
  }
}

// The backward's two panels in one launch: workgroups [0, bslabs) write the
// row panel of B (dvals), the next cslabs the pair panel of dC (dB).
template <int NT>
__global__ __launch_bounds__(kTB) void bwd_panels(const uint16_t* __restrict__ B, int64_t K,
                                                  int64_t bslabs, uint16_t* __restrict__ Bq,
                                                  const uint16_t* __restrict__ dC, int64_t M,
                                                  uint16_t* __restrict__ dCp, int64_t first) {
  const int64_t g = first + blockIdx.x;
  if (g < bslabs)
    row_panel_slab<NT>(B, K, Bq, g);
  else
    pair_panel_slab<NT>(dC, M, dCp, g - bslabs);
}

// forward (TRANS = false) and dB (TRANS = true). One workgroup per block
// row (block column for dB, whose blocks `perm` lists) of WC*WB waves: wave
// (wc, wb) owns column tiles [wc*TPW, (wc+1)*TPW) of the output and blocks
// wb, wb + WB, ... of the row. Each block costs the wave one 8-byte A load,
// TPW/2 16-byte panel loads and TPW MFMAs, whose results are scaled into the
// FP32 sums. Block, slab and scale of 64 blocks at a time are loaded one per
// lane and broadcast with readlane. The WB partial tiles are added through
// LDS in wave order. At most 128 registers (four waves per SIMD): one
// workgroup per block row then keeps every block row of a 16384-row matrix
// resident at once.
template <int NT>
struct RowsShape {
  static constexpr int TPW = NT < 8 ? NT : 8;  // 16-column tiles per wave
  static constexpr int WC = NT / TPW;          // column groups
  static constexpr int WB = 4;                 // block strides
  static constexpr int NW = WC * WB;           // waves
};

template <int NT, bool BF, bool TRANS>
__global__ __launch_bounds__(RowsShape<NT>::NW * kWave, 4) void bspmm_rows(
    const int32_t* __restrict__ ptr, const int32_t* __restrict__ perm,
    const int32_t* __restrict__ other, const uint16_t* __restrict__ blocks,
    const float* __restrict__ scale, const uint16_t* __restrict__ panel, float* __restrict__ out,
    int64_t out_rows, int64_t nouter, int64_t first) {
  constexpr int TPW = RowsShape<NT>::TPW, WC = RowsShape<NT>::WC, WB = RowsShape<NT>::WB;
  constexpr int NP = TPW / 2;  // panel pairs per wave
  __shared__ f4 red[WB - 1][WC][TPW][kWave];
  const int64_t r = first + blockIdx.x;
  if (r >= nouter) return;
  const int tid = threadIdx.x, w = tid / kWave, lane = tid & (kWave - 1);
  const int wc = w % WC, wb = w / WC;
  const int i = lane & 15, kg = lane >> 4;
  const uint16_t* pw = panel + ((int64_t)wc * NP * kWave + lane) * 8;
  f4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f4{0.f, 0.f, 0.f, 0.f};
  const int k0 = ptr[r], nb = ptr[r + 1] - k0;
  const int cnt = nb > wb ? (nb - wb + WB - 1) / WB : 0;  // this wave's blocks
  for (int base = 0; base < cnt; base += kWave) {
    const int cn = min(kWave, cnt - base);
    int myb = 0, myo = 0;
    float mys = 0.f;
    if (lane < cn) {
      const int k = k0 + wb + WB * (base + lane);
      myb = TRANS ? perm[k] : k;
      myo = other[k];
      mys = scale[myb];
    }
    for (int m = 0; m < cn; ++m) {
      const int b = __builtin_amdgcn_readlane(myb, m);
      const int64_t o = __builtin_amdgcn_readlane(myo, m);
      const float sc =
          __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, mys), m));
      const u2 a = *reinterpret_cast<const u2*>(blocks + (int64_t)b * 256 + i * 16 + 4 * kg);
      u4 bo[NP];
#pragma unroll
      for (int p = 0; p < NP; ++p)
        bo[p] = *reinterpret_cast<const u4*>(pw + (o * (NT / 2) + p) * (kWave * 8));
      const f4 z = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int p = 0; p < NP; ++p) {
        acc[2 * p] += sc * mma16<BF>(a, u2{bo[p].x, bo[p].y}, z);
        acc[2 * p + 1] += sc * mma16<BF>(a, u2{bo[p].z, bo[p].w}, z);
      }
    }
  }
  if (wb > 0)
#pragma unroll
    for (int t = 0; t < TPW; ++t) red[wb - 1][wc][t][lane] = acc[t];
  __syncthreads();
  if (wb != 0) return;
#pragma unroll
  for (int q = 0; q < WB - 1; ++q)
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] += red[q][wc][t][lane];
  constexpr int N = NT * 16;
#pragma unroll
  for (int t = 0; t < TPW; ++t)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = r * 16 + 4 * kg + v;
      if (row < out_rows) out[row * N + (wc * TPW + t) * 16 + i] = acc[t][v];
    }
}

// dvals: dA_b = dC[16 rows of br, :] * B[16 rows of bc, :]^T (K = N) on the
// K = 32 MFMA with B's rows (from the row panel) as the A operand and dC's
// rows as the B operand, so lane (i, kg) ends with dA_b[i][4kg .. 4kg+3] and
// a block's 256 results leave in one contiguous 1 KB store. One workgroup of
// kDvWaves waves per block row, wave w taking blocks w, w + kDvWaves, ...;
// block columns of 64 blocks at a time loaded one per lane.
constexpr int kDvWaves = 4;

template <int NT, bool BF>
__global__ __launch_bounds__(kDvWaves * 64) void bspmm_dvals(const int32_t* __restrict__ browptr,
                                                             const int32_t* __restrict__ bcol,
                                                             const uint16_t* __restrict__ dC,
                                                             int64_t M,
                                                             const uint16_t* __restrict__ Bq,
                                                             float* __restrict__ dvals,
                                                             int64_t brows, int64_t first) {
  constexpr int N = NT * 16, S = N / 32;
  const int64_t br = first + blockIdx.x;
  if (br >= brows) return;
  const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
  const int i = lane & 15, kg = lane >> 4;
  u4 a[S];
  const int64_t arow = br * 16 + i;
  // clamped unconditional loads, rows past M zeroed afterwards: no branch
  // splits the batch of loads
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const u4 v = *reinterpret_cast<const u4*>(dC + min(arow, M - 1) * N + 32 * s + 8 * kg);
    a[s] = arow < M ? v : u4{0, 0, 0, 0};
  }
  const int k0 = browptr[br], nb = browptr[br + 1] - k0;
  const int cnt = nb > w ? (nb - w + kDvWaves - 1) / kDvWaves : 0;  // this wave's blocks
  for (int base = 0; base < cnt; base += kWave) {
    const int cn = min(kWave, cnt - base);
    int myo = 0;
    if (lane < cn) myo = bcol[k0 + w + kDvWaves * (base + lane)];
    for (int m = 0; m < cn; ++m) {
      const int64_t o = __builtin_amdgcn_readlane(myo, m);
      u4 bo[S];
#pragma unroll
      for (int s = 0; s < S; ++s)
        bo[s] = *reinterpret_cast<const u4*>(Bq + ((o * S + s) * kWave + lane) * 8);
      f4 acc = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int s = 0; s < S; ++s) acc = mma32<BF>(bo[s], a[s], acc);
      const int64_t b = k0 + w + kDvWaves * (base + m);
      *reinterpret_cast<f4*>(dvals + b * 256 + i * 16 + 4 * kg) = acc;
    }
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
static bsp_status launch_dvals(bsp_bspmm_s* h, int64_t N, const uint16_t* dC, const uint16_t* Bq,
                               float* dvals, hipStream_t s) {
  if (h->brows == 0) return BSP_OK;
#define DV_CASE(NT)                                                                             \
  case NT:                                                                                      \
    in_slices(h->brows, kDvWaves * 64, [&](int64_t f, int64_t c) {                             \
      bspmm_dvals<NT, BF><<<c, kDvWaves * 64, 0, s>>>(h->browptr, h->bcol, dC, h->M, Bq,       \
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

// Grows a panel buffer (the first call of a product, outside any capture).
static bsp_status ensure(DevBuf<uint16_t>& buf, int64_t& cap, int64_t need, hipStream_t s) {
  if (need <= cap) return BSP_OK;
  BSP_HIP(hipStreamSynchronize(s));
  BSP_TRY(buf.alloc(need));
  cap = need;
  return BSP_OK;
}

#define BSPMM_NT_SWITCH(N, CASE)                                                    \
  switch ((N) / 16) {                                                               \
    CASE(2)                                                                         \
    CASE(4)                                                                         \
    CASE(8)                                                                         \
    CASE(16)                                                                        \
    default:                                                                        \
      return fail(BSP_ERR_UNSUPPORTED, "block SpMM needs N in {32, 64, 128, 256}"); \
  }

// B (K x N) in pair-panel layout in h->work, for the forward.
static bsp_status fwd_panel(bsp_bspmm_s* h, const uint16_t* B, int64_t N, hipStream_t s) {
  BSP_TRY(ensure(h->work, h->work_cap, h->bcols * 16 * N, s));
  if (h->bcols == 0) return BSP_OK;
#define PANEL_CASE(NT)                                                                  \
  case NT:                                                                              \
    in_slices(h->bcols, kTB, [&](int64_t f, int64_t c) {                                \
      panelize<NT><<<c, kTB, 0, s>>>(B, h->K, h->work.get(), f);                        \
    });                                                                                 \
    break;
  BSPMM_NT_SWITCH(N, PANEL_CASE)
#undef PANEL_CASE
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

// The backward's panels in one launch: B's row panel in h->work2 (when dvals
// is wanted) and dC's pair panel in h->work (when dB is).
static bsp_status bwd_panel(bsp_bspmm_s* h, const uint16_t* B, const uint16_t* dC, int64_t N,
                            bool want_dvals, bool want_dB, hipStream_t s) {
  const int64_t bslabs = want_dvals ? h->bcols : 0, cslabs = want_dB ? h->brows : 0;
  if (want_dvals) BSP_TRY(ensure(h->work2, h->work2_cap, h->bcols * 16 * N, s));
  if (want_dB) BSP_TRY(ensure(h->work, h->work_cap, h->brows * 16 * N, s));
  if (bslabs + cslabs == 0) return BSP_OK;
#define BWD_PANEL_CASE(NT)                                                              \
  case NT:                                                                              \
    in_slices(bslabs + cslabs, kTB, [&](int64_t f, int64_t c) {                         \
      bwd_panels<NT><<<c, kTB, 0, s>>>(B, h->K, bslabs, h->work2.get(), dC, h->M,       \
                                       h->work.get(), f);                               \
    });                                                                                 \
    break;
  BSPMM_NT_SWITCH(N, BWD_PANEL_CASE)
#undef BWD_PANEL_CASE
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

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

bsp_status bsp_bspmm_fwd(bsp_bspmm h, const void* B, int64_t N, float* Cout, bsp_stream s) {
  BSP_REQUIRE(h && B && Cout, "null argument");
  BSP_REQUIRE(aligned16(B), "B must be 16-byte aligned");
  hipStream_t st = (hipStream_t)s;
  BSP_REQUIRE(N == 32 || N == 64 || N == 128 || N == 256, "block SpMM needs N in {32, 64, 128, 256}");
  // B in pair-panel layout: each block column's 16 x N slab in MFMA operand order
  BSP_TRY(fwd_panel(h, static_cast<const uint16_t*>(B), N, st));
  if (h->ht == BSP_BF16) return launch_rows<true, false>(h, N, h->work.get(), Cout, h->M, st);
  return launch_rows<false, false>(h, N, h->work.get(), Cout, h->M, st);
}

bsp_status bsp_bspmm_bwd(bsp_bspmm h, const void* B, const void* dC, int64_t N, float* dB,
                         float* dvals, bsp_stream s) {
  BSP_REQUIRE(h && B && dC, "null argument");
  BSP_REQUIRE(aligned16(B) && aligned16(dC) && (!dvals || aligned16(dvals)),
              "B, dC and dvals must be 16-byte aligned");
  hipStream_t st = (hipStream_t)s;
  const bool bf = h->ht == BSP_BF16;
  BSP_REQUIRE(N == 32 || N == 64 || N == 128 || N == 256, "block SpMM needs N in {32, 64, 128, 256}");
  // B's rows in row-panel layout (the A operand of each block's dvals
  // product) and dC in pair-panel layout (the dB product's B operand)
  BSP_TRY(bwd_panel(h, static_cast<const uint16_t*>(B), static_cast<const uint16_t*>(dC), N,
                    dvals != nullptr, dB != nullptr, st));
  if (dvals) {
    if (bf)
      BSP_TRY(launch_dvals<true>(h, N, (const uint16_t*)dC, h->work2.get(), dvals, st));
    else
      BSP_TRY(launch_dvals<false>(h, N, (const uint16_t*)dC, h->work2.get(), dvals, st));
  }
  if (dB) {
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

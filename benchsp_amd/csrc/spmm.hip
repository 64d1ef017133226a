// CSR SpMM with dense row-major operands, its transpose-free transpose product
// and SDDMM on the pattern: the forward and backward of Y = A*X.
//
// Work list: every row is cut into chunks of at most kChunk entries; a wave
// owns one chunk. Rows with one chunk write their output row directly; longer
// rows (graph hubs) write per-chunk partials that a second pass adds up in
// chunk order, so results are deterministic and free of atomics.
//
// A^T * dY reuses the same kernel over the pattern's CSC (column pointers,
// row indices and the CSR position of every entry, built once with the
// plan); values are read through that permutation, never copied.
#include <memory>

#include <type_traits>

#include "bsp_hip.hpp"

struct bsp_spmm_plan_s {
  bsp_csr A{};
  // forward (CSR) and transposed (CSC) traversals
  struct Side {
    const int32_t* ptr = nullptr;
    const int32_t* idx = nullptr;
    const int32_t* perm = nullptr;  // CSC only: CSR position of each entry
    int64_t nrows = 0;
    bsp::DevBuf<int4> items;        // {row, begin, end, partial slot or -1}
    int64_t nitems = 0;
    bsp::DevBuf<int4> combine;      // {row, first slot, slot count, 0}
    int64_t ncombine = 0;
    int64_t nslots = 0;
  } fwd, bwd;
  bsp::DevBuf<int32_t> colptr, rowidx, perm;
  bsp::DevBuf<float> partial;       // max slots * F floats, grown on demand
  int64_t partial_cap = 0;
  // fused backward: CSC position of every CSR entry and the CSC-ordered
  // SDDMM results (nnz each, built on its first call)
  bsp::DevBuf<int32_t> iperm;
  bsp::DevBuf<float> dvt;
};

namespace bsp {
namespace {

constexpr int kTB = 256;
constexpr int kChunk = 512;

__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

template <class T>
__device__ __forceinline__ float to_f32(T v);
template <>
__device__ __forceinline__ float to_f32<float>(float v) { return v; }
template <>
__device__ __forceinline__ float to_f32<uint16_t>(uint16_t v) { return bf16_to_f32(v); }

// Four consecutive features of a row of X starting at p, as floats (one
// 16-byte FP32 or 8-byte BF16 load).
template <class T>
__device__ __forceinline__ void load_feats4(const T* p, float (&f)[4]);
template <>
__device__ __forceinline__ void load_feats4<float>(const float* p, float (&f)[4]) {
  const float4 v = *reinterpret_cast<const float4*>(p);
  f[0] = v.x; f[1] = v.y; f[2] = v.z; f[3] = v.w;
}
template <>
__device__ __forceinline__ void load_feats4<uint16_t>(const uint16_t* p, float (&f)[4]) {
  const uint2 v = *reinterpret_cast<const uint2*>(p);
  f[0] = __uint_as_float(v.x << 16); f[1] = __uint_as_float(v.x & 0xffff0000u);
  f[2] = __uint_as_float(v.y << 16); f[3] = __uint_as_float(v.y & 0xffff0000u);
}

// v read from another lane of its row of 16 through a DPP lane pattern.
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}

// Sum of p over each aligned group of G lanes (G = 16, 32 or 64), in every
// lane of the group. Within a row of 16 lanes: quad swaps [1,0,3,2] and
// [2,3,0,1], then the half-row and row mirrors, as DPP operands of the adds
// (no LDS); wider groups add their rows with lane shuffles.
template <int G>
__device__ __forceinline__ float group_sum(float p) {
  p += dpp_f32<0xB1>(p);
  p += dpp_f32<0x4E>(p);
  p += dpp_f32<0x141>(p);
  p += dpp_f32<0x140>(p);
#pragma unroll
  for (int o = 16; o < G; o <<= 1) p += __shfl_xor(p, o, 64);
  return p;
}

// Y[row, :] (or a partial) = sum over the item's entries of a * X[col, :];
// wave per work item. G = F/4 lanes share an entry, each owning four
// features (one 16-byte FP32 or 8-byte BF16 load), so NZ = 64/G lane groups
// sit side by side in a wave; a step covers NZ*U entries, group `sub` taking
// the U consecutive ones from t0 + sub*U with all U rows in flight. The
// item's (col, a) pairs are loaded 64 at a time, one per lane, and handed
// out with lane shuffles. Each group sums its entries in order; the NZ group
// sums are added with a fixed shuffle tree.
//
// SDD (transposed traversal only): the same pass also writes the SDDMM of
// the backward in traversal order, dvals[j] = dot(X[row_j, :], D[item, :])
// with D the item's own row of the second operand, so every gathered row
// serves both products. After the group sums lane q < U of a group holds
// the dot product of the group's entry q, and one store writes the step's
// results contiguously. (Storing each straight to its CSR position instead,
// a random 4-byte write per entry, cost 140 us FP32 / 176 us BF16 on a
// 10 M-entry graph; non-temporal, 165 / 234 us.)
template <class TV, class TX, int F, bool GATHER, bool SDD = false>
__global__ __launch_bounds__(kTB) void spmm_items(const int32_t* __restrict__ idx,
                                                  const TV* __restrict__ val,
                                                  const int32_t* __restrict__ perm,
                                                  const TX* __restrict__ X, float* __restrict__ Y,
                                                  const int4* __restrict__ items, int64_t nitems,
                                                  float* __restrict__ partial,
                                                  const TX* __restrict__ D = nullptr,
                                                  float* __restrict__ dvals = nullptr) {
  constexpr int G = F / 4, NZ = kWave / G, U = 4;
  static_assert(!SDD || GATHER, "the fused SDDMM runs on the transposed traversal");
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (w >= nitems) return;
  const int4 it = items[w];
  const int q = lane % G, sub = lane / G;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  float dr[4] = {0.f, 0.f, 0.f, 0.f};
  if constexpr (SDD) load_feats4<TX>(D + (int64_t)it.x * F + q * 4, dr);
  for (int base = it.y; base < it.z; base += kWave) {
    const int j = base + lane;
    int c = 0, pj = 0;
    float a = 0.f;
    if (j < it.z) {
      c = ld_stream(&idx[j]);
      if constexpr (GATHER) pj = ld_stream(&perm[j]);
      a = to_f32<TV>(GATHER ? val[pj] : ld_stream(&val[j]));
    }
    const int n = min(kWave, it.z - base);
    for (int t0 = 0; t0 < n; t0 += NZ * U) {
      const int j0 = t0 + sub * U;
      float xv[U][4], av[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int e = min(j0 + u, kWave - 1);
        av[u] = __shfl(a, e, 64);
        load_feats4<TX>(X + (int64_t)__shfl(c, e, 64) * F + q * 4, xv[u]);
      }
      float d = 0.f;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (j0 + u < n)
#pragma unroll
          for (int v = 0; v < 4; ++v) acc[v] += av[u] * xv[u][v];
        if constexpr (SDD) {
          const float p = group_sum<G>(dr[0] * xv[u][0] + dr[1] * xv[u][1] + dr[2] * xv[u][2] +
                                       dr[3] * xv[u][3]);
          d = q == u ? p : d;
        }
      }
      if constexpr (SDD) {
        if (q < U && j0 + q < n) dvals[base + j0 + q] = d;
      }
    }
  }
#pragma unroll
  for (int o = G; o < kWave; o <<= 1)
#pragma unroll
    for (int v = 0; v < 4; ++v) acc[v] += __shfl_xor(acc[v], o, 64);
  if (sub == 0) {
    float* dst = it.w < 0 ? Y + (int64_t)it.x * F : partial + (int64_t)it.w * F;
    *reinterpret_cast<float4*>(dst + q * 4) = make_float4(acc[0], acc[1], acc[2], acc[3]);
  }
}

// dst[j] = src[iperm[j]]: results in CSC order back to CSR order, a thread
// per entry (streamed index, gathered value, coalesced store).
__global__ __launch_bounds__(kTB) void to_csr_order(const int32_t* __restrict__ iperm,
                                                    const float* __restrict__ src, int64_t nnz,
                                                    float* __restrict__ dst) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nnz) dst[j] = src[ld_stream(&iperm[j])];
}

// iperm[perm[k]] = k: the inverse of the CSC permutation.
__global__ __launch_bounds__(kTB) void invert_perm(const int32_t* __restrict__ perm, int64_t nnz,
                                                   int32_t* __restrict__ iperm) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < nnz) iperm[perm[k]] = (int32_t)k;
}

// Y[row, :] = sum of its chunk partials; a workgroup per row. Thread t owns
// feature t % F of chunk group t / F (256 / F groups): group g adds chunks
// g, g + G, g + 2G, ... in four interleaved running sums, and the groups'
// sums are added in group order through LDS. Fixed order throughout: the
// same bits every run. (A wave per row with four loads in flight took
// 15 us on a graph whose largest row has 188 chunks.)
__global__ __launch_bounds__(kTB) void spmm_combine(const int4* __restrict__ comb, int64_t ncomb,
                                                    const float* __restrict__ partial, int64_t F,
                                                    float* __restrict__ Y) {
  __shared__ float red[kTB];
  const int64_t w = blockIdx.x;
  if (w >= ncomb) return;
  const int4 c = comb[w];
  const int G = kTB / (int)F, f = threadIdx.x % (int)F, g = threadIdx.x / (int)F;
  const float* p = partial + (int64_t)c.y * F + f;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int k = g;
  for (; k + 3 * G < c.z; k += 4 * G) {
    s0 += p[(int64_t)k * F];
    s1 += p[(int64_t)(k + G) * F];
    s2 += p[(int64_t)(k + 2 * G) * F];
    s3 += p[(int64_t)(k + 3 * G) * F];
  }
  for (; k < c.z; k += G) s0 += p[(int64_t)k * F];
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  if (g == 0) {
    float t = red[f];
    for (int q = 1; q < G; ++q) t += red[q * F + f];
    Y[(int64_t)c.x * F + f] = t;
  }
}

// dvals[j] = dot(dY[row_j, :], X[col_j, :]); wave per work item, G = F/4
// lanes per entry (float4 each), NZ = 64/G lane groups; the item's dY row is
// loaded once and reused for all its entries. A step covers NZ*U entries:
// group `sub` takes the U consecutive entries from t0 + sub*U, with all U
// X rows in flight before the dot products (U = 8: BF16 F = 64 215 -> 201 us
// against 4). After the group sums, lane q < U of each group holds entry
// t0 + sub*U + q, so one store writes the step's NZ*U results contiguously.
template <class TX, int F>
__global__ __launch_bounds__(kTB) void sddmm_items(const int32_t* __restrict__ col,
                                                   const TX* __restrict__ dY,
                                                   const TX* __restrict__ X,
                                                   const int4* __restrict__ items, int64_t nitems,
                                                   float* __restrict__ dvals) {
  constexpr int G = F / 4, NZ = kWave / G, U = 8;
  static_assert(G >= U && G % 16 == 0, "F must be a multiple of 64");
  const int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (w >= nitems) return;
  const int4 it = items[w];
  const int q = lane % G, sub = lane / G;
  float dy[4];
  load_feats4<TX>(dY + (int64_t)it.x * F + q * 4, dy);
  for (int base = it.y; base < it.z; base += kWave) {
    const int cl = base + lane < it.z ? ld_stream(&col[base + lane]) : 0;
    const int n = min(kWave, it.z - base);
    for (int t0 = 0; t0 < n; t0 += NZ * U) {
      const int j0 = t0 + sub * U;
      float xv[U][4];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int c = __shfl(cl, min(j0 + u, kWave - 1), 64);
        load_feats4<TX>(X + (int64_t)c * F + q * 4, xv[u]);
      }
      float r = 0.f;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const float p = group_sum<G>(dy[0] * xv[u][0] + dy[1] * xv[u][1] + dy[2] * xv[u][2] +
                                     dy[3] * xv[u][3]);
        r = q == u ? p : r;
      }
      if (q < U && j0 + q < n) dvals[base + j0 + q] = r;
    }
  }
}

// ---- work-list construction ----------------------------------------------

__global__ void chunk_counts(const int32_t* __restrict__ ptr, int64_t n, int32_t* __restrict__ nch,
                             int32_t* __restrict__ nslot, int32_t* __restrict__ multi) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r > n) return;
  if (r == n) {
    nch[r] = 0;
    nslot[r] = 0;
    multi[r] = 0;
    return;
  }
  const int len = ptr[r + 1] - ptr[r];
  const int c = len <= kChunk ? 1 : (len + kChunk - 1) / kChunk;
  nch[r] = c;
  nslot[r] = c > 1 ? c : 0;
  multi[r] = c > 1;
}

__global__ void fill_items(const int32_t* __restrict__ ptr, int64_t n,
                           const int32_t* __restrict__ itemptr, const int32_t* __restrict__ slotptr,
                           const int32_t* __restrict__ combpos, const int32_t* __restrict__ multi,
                           int4* __restrict__ items, int4* __restrict__ comb) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = ptr[r], e = ptr[r + 1];
  const int c = itemptr[r + 1] - itemptr[r];
  for (int k = 0; k < c; ++k)
    items[itemptr[r] + k] = make_int4((int)r, s + k * kChunk, min(e, s + (k + 1) * kChunk),
                                      c > 1 ? slotptr[r] + k : -1);
  if (multi[r]) comb[combpos[r]] = make_int4((int)r, slotptr[r], c, 0);
}

// wave per row (lanes stride the row), so a long row is not one thread's loop
__global__ __launch_bounds__(kTB) void csc_keys(const int32_t* __restrict__ rowptr,
                                                const int32_t* __restrict__ col, int64_t n,
                                                int64_t* __restrict__ key,
                                                int32_t* __restrict__ idx, int64_t first) {
  const int64_t r = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= n) return;
  for (int j = rowptr[r] + lane, e = rowptr[r + 1]; j < e; j += kWave) {
    key[j] = (int64_t)col[j] * n + r;
    idx[j] = j;
  }
}

// Row indices and column pointers from the (column, row)-sorted keys:
// colptr[c] is the first k whose column is >= c, written by the thread at the
// boundary where the column changes (k = nnz closes the last columns), so
// every entry is written once and no counters are needed.
__global__ void csc_fill(const int64_t* __restrict__ skey, int64_t nnz, int64_t n, int64_t ncols,
                         int32_t* __restrict__ rowidx, int32_t* __restrict__ colptr) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nnz) return;
  const int64_t c = k < nnz ? skey[k] / n : ncols;
  const int64_t cp = k > 0 ? skey[k - 1] / n : -1;
  if (k < nnz) rowidx[k] = (int32_t)(skey[k] % n);
  for (int64_t q = cp + 1; q <= c; ++q) colptr[q] = (int32_t)k;
}

}  // namespace

static bsp_status build_items(bsp_spmm_plan_s::Side& sd, hipStream_t s) {
  const int64_t n = sd.nrows;
  DevBuf<int32_t> nch, nslot, multi, itemptr, slotptr, combpos;
  BSP_TRY(nch.alloc(n + 1));
  BSP_TRY(nslot.alloc(n + 1));
  BSP_TRY(multi.alloc(n + 1));
  BSP_TRY(itemptr.alloc(n + 1));
  BSP_TRY(slotptr.alloc(n + 1));
  BSP_TRY(combpos.alloc(n + 1));
  chunk_counts<<<cdiv(n + 1, kTB), kTB, 0, s>>>(sd.ptr, n, nch.get(), nslot.get(), multi.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(exclusive_scan_i32(nch.get(), itemptr.get(), n + 1, s));
  BSP_TRY(exclusive_scan_i32(nslot.get(), slotptr.get(), n + 1, s));
  BSP_TRY(exclusive_scan_i32(multi.get(), combpos.get(), n + 1, s));
  int32_t ni = 0, ns = 0, nc = 0;
  BSP_TRY(fetch(itemptr.get() + n, &ni, s));
  BSP_TRY(fetch(slotptr.get() + n, &ns, s));
  BSP_TRY(fetch(combpos.get() + n, &nc, s));
  sd.nitems = ni;
  sd.nslots = ns;
  sd.ncombine = nc;
  BSP_TRY(sd.items.alloc(ni));
  BSP_TRY(sd.combine.alloc(nc));
  if (n > 0)
    fill_items<<<cdiv(n, kTB), kTB, 0, s>>>(sd.ptr, n, itemptr.get(), slotptr.get(), combpos.get(),
                                            multi.get(), sd.items.get(), sd.combine.get());
  BSP_HIP(hipGetLastError());
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

static bsp_status build_spmm_plan(bsp_spmm_plan_s* p, hipStream_t s) {
  const bsp_csr& A = p->A;
  p->fwd.ptr = A.rowptr;
  p->fwd.idx = A.colidx;
  p->fwd.nrows = A.nrows;
  BSP_TRY(build_items(p->fwd, s));
  // CSC of the pattern: stable sort of (col, row) keys
  const int64_t nnz = A.nnz;
  DevBuf<int64_t> key, skey;
  DevBuf<int32_t> idx;
  BSP_TRY(key.alloc(nnz));
  BSP_TRY(skey.alloc(nnz));
  BSP_TRY(idx.alloc(nnz));
  BSP_TRY(p->perm.alloc(nnz));
  BSP_TRY(p->rowidx.alloc(nnz));
  BSP_TRY(p->colptr.alloc(A.ncols + 1));
  if (nnz > 0) {
    in_slices(A.nrows, kWave, [&](int64_t f, int64_t c) {
      csc_keys<<<cdiv(c * kWave, kTB), kTB, 0, s>>>(A.rowptr, A.colidx, A.nrows, key.get(),
                                                    idx.get(), f);
    });
    BSP_HIP(hipGetLastError());
    int bits = 1;
    while (bits < 63 && (int64_t(1) << bits) <= A.ncols * A.nrows) ++bits;
    BSP_TRY(sort_pairs_i64key(key.get(), skey.get(), idx.get(), p->perm.get(), nnz, bits, s));
  }
  csc_fill<<<cdiv(nnz + 1, kTB), kTB, 0, s>>>(skey.get(), nnz, A.nrows, A.ncols, p->rowidx.get(),
                                               p->colptr.get());
  BSP_HIP(hipGetLastError());
  p->bwd.ptr = p->colptr.get();
  p->bwd.idx = p->rowidx.get();
  p->bwd.perm = p->perm.get();
  p->bwd.nrows = A.ncols;
  BSP_TRY(build_items(p->bwd, s));
  return BSP_OK;
}

// One traversal of `sd`: Y (or chunk partials, then their combine) from
// the gathered rows of X; with dvals set (transposed side only) also the
// fused SDDMM against the item rows of D.
template <class TV, class TX, bool GATHER>
static bsp_status run_spmm_t(bsp_spmm_plan_s* p, const bsp_spmm_plan_s::Side& sd, const void* V,
                             const void* X, int64_t F, float* Y, const void* D, float* dvals,
                             hipStream_t s) {
  const auto* val = static_cast<const TV*>(V);
  const auto* x = static_cast<const TX*>(X);
  const auto* d = static_cast<const TX*>(D);
  const int64_t grid = cdiv(sd.nitems * kWave, kTB);
  if (sd.nslots * F > p->partial_cap) {
    BSP_HIP(hipStreamSynchronize(s));
    BSP_TRY(p->partial.alloc(sd.nslots * F));
    p->partial_cap = sd.nslots * F;
  }
  if (sd.nitems == 0) return BSP_OK;
  auto launch = [&](auto fw) -> bsp_status {
    constexpr int FW = decltype(fw)::value;
    if constexpr (GATHER) {
      if (dvals) {
        spmm_items<TV, TX, FW, true, true><<<grid, kTB, 0, s>>>(
            sd.idx, val, sd.perm, x, Y, sd.items.get(), sd.nitems, p->partial.get(), d, dvals);
        return BSP_OK;
      }
    }
    spmm_items<TV, TX, FW, GATHER><<<grid, kTB, 0, s>>>(sd.idx, val, sd.perm, x, Y,
                                                        sd.items.get(), sd.nitems,
                                                        p->partial.get());
    return BSP_OK;
  };
  switch (F) {
    case 64: BSP_TRY(launch(std::integral_constant<int, 64>{})); break;
    case 128: BSP_TRY(launch(std::integral_constant<int, 128>{})); break;
    case 256: BSP_TRY(launch(std::integral_constant<int, 256>{})); break;
    default:
      return fail(BSP_ERR_UNSUPPORTED, "SpMM feature width must be 64, 128 or 256");
  }
  BSP_HIP(hipGetLastError());
  if (sd.ncombine > 0) {
    spmm_combine<<<sd.ncombine, kTB, 0, s>>>(sd.combine.get(), sd.ncombine, p->partial.get(), F,
                                              Y);
    BSP_HIP(hipGetLastError());
  }
  return BSP_OK;
}

template <bool GATHER>
static bsp_status run_spmm(bsp_spmm_plan_s* p, const bsp_spmm_plan_s::Side& sd, const void* V,
                           bsp_dtype vt, const void* X, bsp_dtype xt, int64_t F, float* Y,
                           hipStream_t s, const void* D = nullptr, float* dvals = nullptr) {
  const bool vb = vt == BSP_BF16, xb = xt == BSP_BF16;
  if (vb && xb) return run_spmm_t<uint16_t, uint16_t, GATHER>(p, sd, V, X, F, Y, D, dvals, s);
  if (vb) return run_spmm_t<uint16_t, float, GATHER>(p, sd, V, X, F, Y, D, dvals, s);
  if (xb) return run_spmm_t<float, uint16_t, GATHER>(p, sd, V, X, F, Y, D, dvals, s);
  return run_spmm_t<float, float, GATHER>(p, sd, V, X, F, Y, D, dvals, s);
}

}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_spmm_plan_create(const bsp_csr* A, bsp_stream s, bsp_spmm_plan* out) {
  BSP_REQUIRE(A && out, "null argument");
  BSP_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, "negative CSR size");
  if (!sizes_fit(A->nrows, A->ncols, A->nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(A->nrows == 0 || A->rowptr, "null rowptr");
  BSP_REQUIRE(A->nnz == 0 || A->colidx, "null colidx");
  auto p = std::make_unique<bsp_spmm_plan_s>();
  p->A = *A;
  BSP_TRY(build_spmm_plan(p.get(), (hipStream_t)s));
  *out = p.release();
  return BSP_OK;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15) == 0; }

bsp_status bsp_spmm(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* X,
                    bsp_dtype xdtype, int64_t F, float* Y, bsp_stream s) {
  BSP_REQUIRE(p && (X || p->A.ncols == 0) && (Y || p->A.nrows == 0), "null argument");
  BSP_REQUIRE(aligned16(X) && aligned16(Y), "X and Y must be 16-byte aligned");
  BSP_REQUIRE(val || p->A.nnz == 0, "null values");
  BSP_REQUIRE(xdtype == BSP_F32 || xdtype == BSP_BF16, "X must be FP32 or BF16");
  BSP_REQUIRE(vdtype == BSP_F32 || vdtype == BSP_BF16, "values must be FP32 or BF16");
  return run_spmm<false>(p, p->fwd, val, vdtype, X, xdtype, F, Y, (hipStream_t)s);
}

bsp_status bsp_spmm_t(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* dY,
                      bsp_dtype dtype, int64_t F, float* dX, bsp_stream s) {
  BSP_REQUIRE(p && (dY || p->A.nrows == 0) && (dX || p->A.ncols == 0), "null argument");
  BSP_REQUIRE(aligned16(dY) && aligned16(dX), "dY and dX must be 16-byte aligned");
  BSP_REQUIRE(val || p->A.nnz == 0, "null values");
  BSP_REQUIRE(dtype == BSP_F32 || dtype == BSP_BF16, "dY must be FP32 or BF16");
  BSP_REQUIRE(vdtype == BSP_F32 || vdtype == BSP_BF16, "values must be FP32 or BF16");
  return run_spmm<true>(p, p->bwd, val, vdtype, dY, dtype, F, dX, (hipStream_t)s);
}

bsp_status bsp_sddmm(bsp_spmm_plan p, const void* dY, const void* X, bsp_dtype dtype, int64_t F,
                     float* dvals, bsp_stream s) {
  BSP_REQUIRE(p && dvals, "null argument");
  BSP_REQUIRE(aligned16(dY) && aligned16(X), "dY and X must be 16-byte aligned");
  BSP_REQUIRE(dtype == BSP_F32 || dtype == BSP_BF16, "operands must be FP32 or BF16");
  const auto& sd = p->fwd;
  if (sd.nitems == 0 || p->A.nnz == 0) return BSP_OK;
  const int64_t grid = cdiv(sd.nitems * kWave, kTB);
  hipStream_t st = (hipStream_t)s;
  auto launch = [&](auto tag, auto fw) -> bsp_status {
    using T = decltype(tag);
    constexpr int FW = decltype(fw)::value;
    sddmm_items<T, FW><<<grid, kTB, 0, st>>>(p->A.colidx, static_cast<const T*>(dY),
                                             static_cast<const T*>(X), sd.items.get(), sd.nitems,
                                             dvals);
    return BSP_OK;
  };
  using F64c = std::integral_constant<int, 64>;
  using F128c = std::integral_constant<int, 128>;
  using F256c = std::integral_constant<int, 256>;
  const bool bf = dtype == BSP_BF16;
  switch (F) {
    case 64: bf ? launch(uint16_t{}, F64c{}) : launch(float{}, F64c{}); break;
    case 128: bf ? launch(uint16_t{}, F128c{}) : launch(float{}, F128c{}); break;
    case 256: bf ? launch(uint16_t{}, F256c{}) : launch(float{}, F256c{}); break;
    default: return fail(BSP_ERR_UNSUPPORTED, "SDDMM feature width must be 64, 128 or 256");
  }
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_spmm_t_sddmm(bsp_spmm_plan p, const void* val, bsp_dtype vdtype, const void* dY,
                           const void* X, bsp_dtype dtype, int64_t F, float* dX, float* dvals,
                           bsp_stream s) {
  BSP_REQUIRE(p && (dY || p->A.nrows == 0) && (X || p->A.ncols == 0), "null argument");
  BSP_REQUIRE((dX || p->A.ncols == 0) && (dvals || p->A.nnz == 0), "null output");
  BSP_REQUIRE(aligned16(dY) && aligned16(X) && aligned16(dX),
              "dY, X and dX must be 16-byte aligned");
  BSP_REQUIRE(val || p->A.nnz == 0, "null values");
  BSP_REQUIRE(dtype == BSP_F32 || dtype == BSP_BF16, "dY and X must be FP32 or BF16");
  BSP_REQUIRE(vdtype == BSP_F32 || vdtype == BSP_BF16, "values must be FP32 or BF16");
  hipStream_t st = (hipStream_t)s;
  const int64_t nnz = p->A.nnz;
  if (nnz == 0) return run_spmm<true>(p, p->bwd, val, vdtype, dY, dtype, F, dX, st);
  if (!p->iperm.get()) {
    BSP_HIP(hipStreamSynchronize(st));
    BSP_TRY(p->iperm.alloc(nnz));
    BSP_TRY(p->dvt.alloc(nnz));
    invert_perm<<<cdiv(nnz, kTB), kTB, 0, st>>>(p->perm.get(), nnz, p->iperm.get());
    BSP_HIP(hipGetLastError());
  }
  BSP_TRY(run_spmm<true>(p, p->bwd, val, vdtype, dY, dtype, F, dX, st, X, p->dvt.get()));
  to_csr_order<<<cdiv(nnz, kTB), kTB, 0, st>>>(p->iperm.get(), p->dvt.get(), nnz, dvals);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_spmm_plan_destroy(bsp_spmm_plan p) {
  delete p;
  return BSP_OK;
}

}  // extern "C"

// FP64 CSR SpMV on gfx950: the library's adaptive kernel and the three
// reference CSR kernel families (scalar, balanced, short/long split), each
// with its schedule built on the GPU.
#include <algorithm>
#include <memory>

#include "bsp_hip.hpp"
#include "csr_adaptive.hpp"
#include "csr_plan.hpp"

namespace bsp {

// ---------------------------------------------------------------------------
// Adaptive kernel (block routine in csr_adaptive.hpp).

// (rshift > 0: a row weighs 2^-rshift units instead of one, so a window holds
// more rows. Rows longer than `longrow` take a block of their own, so a
// block has fewer than chunk + longrow entries.)
__global__ void csr_block_flags(const int32_t* __restrict__ rowptr, int64_t nrows, int chunk,
                                int rshift, int longrow, int32_t* __restrict__ flags) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int32_t s = rowptr[r], e = rowptr[r + 1];
  int f = 1;
  if (r > 0) {
    int32_t sp = rowptr[r - 1];
    bool lng = (e - s) > longrow, lngp = (s - sp) > longrow;
    int64_t key = ((int64_t)s + (r >> rshift)) / chunk,
            keyp = ((int64_t)sp + ((r - 1) >> rshift)) / chunk;
    f = (key != keyp) || lng || lngp;
  }
  flags[r] = f;
}

__global__ void scatter_flagged(const int32_t* __restrict__ flags, const int32_t* __restrict__ pos,
                                int64_t n, int32_t* __restrict__ out) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  if (flags[r]) out[pos[r]] = (int32_t)r;
}

__global__ void set_value_i32(int32_t* p, int32_t v) { *p = v; }

template <int CHUNK, bool ACCUM>
__global__ __launch_bounds__(kWG) void csr_adaptive_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
    const int2* __restrict__ blk, int64_t nblk) {
  // y through the caches (a caller's next use of y, or its next SpMV into
  // the same buffer, finds it there; non-temporal stores cost 0.6 us more on
  // the 27-point 64^3 launch)
  adaptive_block<CHUNK, ACCUM, false>(rowptr, col, val, x, y, blk, xcd_remap(blockIdx.x, nblk));
}

// The same with 16-bit column offsets from each block's first row.
template <int CHUNK>
__global__ __launch_bounds__(kWG) void csr_adaptive16_f64(
    const int32_t* __restrict__ rowptr, const int16_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
    const int2* __restrict__ blk, int64_t nblk) {
  adaptive_block<CHUNK, false, false, int16_t>(rowptr, col, val, x, y, blk,
                                               xcd_remap(blockIdx.x, nblk));
}

// Flags (bad = 1) a block with a column index farther than int16 from the
// block's first row; wave per block.
__global__ __launch_bounds__(kWG) void narrow_check(const int32_t* __restrict__ col,
                                                    const int2* __restrict__ blk, int64_t nblk,
                                                    int* __restrict__ bad, int64_t first) {
  const int64_t b = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (b >= nblk) return;
  const int2 s0 = blk[b], s1 = blk[b + 1];
  bool ok = true;
  for (int j = s0.y + lane; j < s1.y; j += kWave) {
    const int d = col[j] - s0.x;
    ok = ok && d >= -32768 && d <= 32767;
  }
  if (!ok) atomicOr(bad, 1);
}

// col16[j] = col[j] - (first row of j's block); wave per block.
__global__ __launch_bounds__(kWG) void narrow_fill(const int32_t* __restrict__ col,
                                                   const int2* __restrict__ blk, int64_t nblk,
                                                   int16_t* __restrict__ col16, int64_t first) {
  const int64_t b = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (b >= nblk) return;
  const int2 s0 = blk[b], s1 = blk[b + 1];
  for (int j = s0.y + lane; j < s1.y; j += kWave) col16[j] = (int16_t)(col[j] - s0.x);
}

// All column segments in one launch, each pinned to a set of XCDs:
// workgroup i runs on XCD i % 8, XCD k serves segment k % nseg, so an XCD
// only ever gathers from its own segment's slice of x, which stays in its
// L2. Each segment writes its own partial y; sum_partials adds them in
// segment order.
struct SegDesc {
  const int32_t* rowptr;
  const int32_t* col;
  const double* val;
  const int2* blk;
  int64_t nblk;
  double* out;
  const uint16_t* slot;  // sorted-block plans: each entry's row-order position in its block
};

template <int CHUNK>
__global__ __launch_bounds__(kWG) void csr_adaptive_xcd_f64(const SegDesc* __restrict__ segs,
                                                            int nseg, const double* __restrict__ x) {
  const int xcd = blockIdx.x % 8;
  const int64_t local = blockIdx.x / 8;
  const int s = xcd % nseg, q = xcd / nseg, reps = 8 / nseg;
  const SegDesc d = segs[s];
  const int64_t b = local * reps + q;
  if (b >= d.nblk) return;
  // partial sums stay cached (sum_partials reads them back at once)
  adaptive_block<CHUNK, false, false>(d.rowptr, d.col, d.val, x, d.out, d.blk, b);
}

// Sorted-block segmented plans (seg_mode 2). A segment's blocks are cut on
// rowptr[r] + r/2 with windows of kSortChunk units; a row longer than
// kSortLong has a block of its own, so a block holds fewer than
// kSortChunk + kSortLong = kSortCap entries. Within a block the entries are
// stored in column order, each with a 2-byte slot: its position in row
// order. A wave's gathers then fall on nearby x entries of the XCD's slice
// and share cache lines, while the products go back to row order through
// LDS (prod[slot]) for the row sums. Every row is summed in row order from
// LDS, by one lane (up to kSortShort entries: the CPU loop's order) or by a
// wave (strided lanes, then a fixed butterfly). 512-lane workgroups, two per
// CU: the launch asks for kSortPad bytes of LDS it does not use, which
// keeps a third from being resident (power-law 2^20, four 2 MB segments:
// 140.9 us with two per CU, 159.4 us with three, 197 us with one; blocks of
// 4,096 / 5,120 / 8,192 entries: 162 / 157 / 152-155 us; the row-order
// plan 155 us. Rows weighing 1/2 unit in the cut: 138.9 us, against 140.6
// with 1/4 and 138.7 with 1; rows of more than 256 / 512 / 1,536 entries
// alone: 155.9 / 145.8 / 139.5 us).
constexpr int kSortCap = 6144;
constexpr int kSortWG = 512;
constexpr int kSortLong = 1024;
constexpr int kSortChunk = kSortCap - kSortLong;
constexpr int kSortRowShift = 1;
constexpr int kSortShort = 32;
constexpr int kSortRows = 2;  // row bounds a lane loads with the entries
constexpr int kSortPad = 24 * 1024;

__global__ __launch_bounds__(kSortWG) void csr_sorted_xcd_f64(const SegDesc* __restrict__ segs,
                                                              int nseg,
                                                              const double* __restrict__ x) {
#pragma clang fp contract(off)
  __shared__ double prod[kSortCap];
  __shared__ int longrow[kSortCap / (kSortShort + 1) + 1];
  __shared__ int nlong;
  __shared__ double red[kSortWG / kWave];
  const int xcd = blockIdx.x % 8;
  const int s = xcd % nseg, q = xcd / nseg, reps = 8 / nseg;
  const SegDesc d = segs[s];
  const int64_t b = (int64_t)(blockIdx.x / 8) * reps + q;
  if (b >= d.nblk) return;
  const int2 s0 = d.blk[b], s1 = d.blk[b + 1];
  const int r0 = s0.x, p0 = s0.y, r1 = s1.x, p1 = s1.y;
  const int tid = threadIdx.x, nnz = p1 - p0;
  if (nnz >= kSortCap) {
    // one long row (only a row longer than kSortLong has a block this
    // large; its entries are in column order, which is row order): the
    // workgroup strides over it
    double a0 = 0.0, a1 = 0.0;
    int i = p0 + tid;
    for (; i + kSortWG < p1; i += 2 * kSortWG) {
      a0 += ld_stream(&d.val[i]) * x[ld_stream(&d.col[i])];
      a1 += ld_stream(&d.val[i + kSortWG]) * x[ld_stream(&d.col[i + kSortWG])];
    }
    if (i < p1) a0 += ld_stream(&d.val[i]) * x[ld_stream(&d.col[i])];
    const double v = group_sum<kWave>(a0 + a1);
    if ((tid & (kWave - 1)) == 0) red[tid / kWave] = v;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < kSortWG / kWave; ++w) t += red[w];
      d.out[r0] = t;
    }
    return;
  }
  constexpr int IT = kSortCap / kSortWG;
  int c[IT], sl[IT];
  double v[IT], xv[IT];
#pragma unroll
  for (int k = 0; k < IT; ++k) {
    const int i = tid + k * kSortWG;
    if (i < nnz) {
      c[k] = ld_stream(&d.col[p0 + i]);
      v[k] = ld_stream(&d.val[p0 + i]);
      sl[k] = ld_stream(&d.slot[p0 + i]);
    }
  }
  // the first rows' bounds load with the entries, not after the barrier
  int rb[kSortRows], re_[kSortRows];
#pragma unroll
  for (int k = 0; k < kSortRows; ++k) {
    const int r = r0 + tid + k * kSortWG;
    if (r < r1) {
      rb[k] = d.rowptr[r];
      re_[k] = d.rowptr[r + 1];
    }
  }
  if (tid == 0) nlong = 0;
#pragma unroll
  for (int k = 0; k < IT; ++k)
    if (tid + k * kSortWG < nnz) xv[k] = x[c[k]];
#pragma unroll
  for (int k = 0; k < IT; ++k)
    if (tid + k * kSortWG < nnz) prod[sl[k]] = v[k] * xv[k];
  __syncthreads();
  auto row = [&](int r, int rs, int re) {
    if (re - rs > kSortShort) {
      longrow[atomicAdd(&nlong, 1)] = r;  // list order varies; each row's sum does not
      return;
    }
    double a = 0.0;
    for (int j = rs; j < re; ++j) a += prod[j];
    d.out[r] = a;
  };
#pragma unroll
  for (int k = 0; k < kSortRows; ++k) {
    const int r = r0 + tid + k * kSortWG;
    if (r < r1) row(r, rb[k] - p0, re_[k] - p0);
  }
  for (int r = r0 + tid + kSortRows * kSortWG; r < r1; r += kSortWG)
    row(r, d.rowptr[r] - p0, d.rowptr[r + 1] - p0);
  __syncthreads();
  const int w = tid / kWave, lane = tid & (kWave - 1);
  for (int t = w; t < nlong; t += kSortWG / kWave) {
    const int r = longrow[t];
    const int rs = d.rowptr[r] - p0, re = d.rowptr[r + 1] - p0;
    double a = 0.0;
    for (int j = rs + lane; j < re; j += kWave) a += prod[j];
    a = group_sum<kWave>(a);
    if (lane == 0) d.out[r] = a;
  }
}

// Sort keys of a segment's entries, a workgroup per block: (global block id,
// column) and the entry's global index; gstart[block] = its first entry.
__global__ __launch_bounds__(kWG) void sorted_keys(const int2* __restrict__ blk,
                                                   const int32_t* __restrict__ col,
                                                   int64_t block0, int64_t entry0,
                                                   int64_t* __restrict__ keys,
                                                   int32_t* __restrict__ idx,
                                                   int64_t* __restrict__ gstart, int64_t first) {
  const int64_t b = first + blockIdx.x;
  const int p0 = blk[b].y, p1 = blk[b + 1].y;
  if (threadIdx.x == 0) gstart[block0 + b] = entry0 + p0;
  for (int i = p0 + (int)threadIdx.x; i < p1; i += kWG) {
    keys[entry0 + i] = ((block0 + b) << 32) | (uint32_t)col[i];
    idx[entry0 + i] = (int32_t)(entry0 + i);
  }
}

// Entries in their sorted order, with their slots.
__global__ __launch_bounds__(kWG) void sorted_apply(const int64_t* __restrict__ keys,
                                                    const int32_t* __restrict__ idx,
                                                    const int64_t* __restrict__ gstart,
                                                    const int32_t* __restrict__ col,
                                                    const double* __restrict__ val, int64_t n,
                                                    int32_t* __restrict__ ocol,
                                                    double* __restrict__ oval,
                                                    uint16_t* __restrict__ slot) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t j = idx[i];
  ocol[i] = col[j];
  oval[i] = val[j];
  slot[i] = (uint16_t)(j - gstart[keys[i] >> 32]);
}

// Data movement of the column-segmented launch without its arithmetic: on
// the same XCD pinning (workgroup i on XCD i % 8 serves segment (i % 8) %
// nseg), each workgroup streams 2048 consecutive entries of its segment's
// column indices (and values, VALS) and gathers x at every one; a sum per
// workgroup is stored so nothing is optimised away.
template <bool VALS>
__global__ __launch_bounds__(kWG) void probe_gather(const SegDesc* __restrict__ segs, int nseg,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ sink) {
  __shared__ double red[kWG / kWave];
  const int xcd = blockIdx.x % 8;
  const int64_t local = blockIdx.x / 8;
  const int s = xcd % nseg, q = xcd / nseg, reps = 8 / nseg;
  const SegDesc d = segs[s];
  const int64_t nnz = d.blk[d.nblk].y;  // the segment's last window bound
  const int64_t b0 = (local * reps + q) * 2048;
  double acc = 0.0;
  if (b0 < nnz) {
    constexpr int IT = 2048 / kWG;
    int c[IT];
    double v[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
      const int64_t j = min(b0 + threadIdx.x + k * kWG, nnz - 1);
      c[k] = ld_stream(&d.col[j]);
      v[k] = VALS ? ld_stream(&d.val[j]) : 1.0;
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) acc += v[k] * x[c[k]];
  }
  acc = group_sum<kWave>(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sink[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Random 8-byte gathers alone (no streams, no arithmetic beyond a sum): the
// rate at which the L2 serves them. Workgroup i, on XCD i % 8, gathers from
// slice (i % 8) % nslices of x (slice length a power of two), each thread
// kProbePer times at xorshift-generated positions; one sum per workgroup is
// stored so nothing is optimised away.
constexpr int kProbePer = 32;
__global__ __launch_bounds__(kWG) void probe_random_gather(const double* __restrict__ x,
                                                           int nslices, int slice_log2,
                                                           double* __restrict__ sink) {
  __shared__ double red[kWG / kWave];
  const int s = (blockIdx.x % 8) % nslices;
  const double* xs = x + ((int64_t)s << slice_log2);
  const uint32_t mask = (1u << slice_log2) - 1;
  uint32_t h = (uint32_t)(blockIdx.x * kWG + threadIdx.x) * 2654435761u + 12345u;
  double acc = 0.0;
#pragma unroll 8
  for (int k = 0; k < kProbePer; ++k) {
    h ^= h << 13;
    h ^= h >> 17;
    h ^= h << 5;
    acc += xs[h & mask];
  }
  acc = group_sum<kWave>(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sink[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Whole-line reads from the same slices: each wave reads kProbePer 1 KB
// runs (16 bytes per lane, 64 consecutive lanes) at xorshift-chosen
// 1 KB-aligned positions of its slice. Against probe_random_gather, which
// moves one 64-byte sector per 8-byte gather, this is the rate at which
// the L2 feeds the CUs full lines.
__global__ __launch_bounds__(kWG) void probe_line_read(const double* __restrict__ x, int nslices,
                                                       int slice_log2,
                                                       double* __restrict__ sink) {
  __shared__ double red[kWG / kWave];
  const int s = (blockIdx.x % 8) % nslices;
  const double* xs = x + ((int64_t)s << slice_log2);
  const uint32_t mask = (1u << (slice_log2 - 7)) - 1;  // 1 KB runs of 128 doubles
  const int lane = threadIdx.x & (kWave - 1);
  uint32_t h = (uint32_t)(blockIdx.x * (kWG / kWave) + threadIdx.x / kWave) * 2654435761u + 12345u;
  double acc = 0.0;
#pragma unroll 8
  for (int k = 0; k < kProbePer; ++k) {
    h ^= h << 13;
    h ^= h >> 17;
    h ^= h << 5;
    const dbl2v v = *reinterpret_cast<const dbl2v*>(xs + ((int64_t)(h & mask) << 7) + 2 * lane);
    acc += v.x + v.y;
  }
  acc = group_sum<kWave>(acc);
  if ((threadIdx.x & (kWave - 1)) == 0) red[threadIdx.x / kWave] = acc;
  __syncthreads();
  if (threadIdx.x == 0) sink[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void sum_partials(const double* __restrict__ part, int nseg, int64_t n,
                             double* __restrict__ y) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int k = 0; k < nseg; ++k) s += ld_stream(&part[k * n + r]);
  y[r] = s;
}

// The same sums two rows per lane with 16-byte loads and store (even n, so
// every partial vector starts 16-byte aligned; y 16-byte aligned).
__global__ void sum_partials2(const double* __restrict__ part, int nseg, int64_t n,
                              double* __restrict__ y) {
  typedef double d2 __attribute__((ext_vector_type(2)));
  const int64_t r = 2 * ((int64_t)blockIdx.x * blockDim.x + threadIdx.x);
  if (r >= n) return;
  d2 s = {0.0, 0.0};
  for (int k = 0; k < nseg; ++k) s += ld_stream(reinterpret_cast<const d2*>(&part[k * n + r]));
  *reinterpret_cast<d2*>(&y[r]) = s;
}

template <int CHUNK>
static void launch_adaptive_t(const CsrSegment& g, bool accum, const double* x, double* y,
                              hipStream_t s) {
  const int2* blk = reinterpret_cast<const int2*>(g.blk.get());
  if (accum)
    csr_adaptive_f64<CHUNK, true><<<g.nblk, kWG, 0, s>>>(g.rowptr, g.col, g.val, x, y, blk, g.nblk);
  else
    csr_adaptive_f64<CHUNK, false><<<g.nblk, kWG, 0, s>>>(g.rowptr, g.col, g.val, x, y, blk, g.nblk);
}

// The window (chunk) is a plan-time value up to 2048; the kernel's LDS
// capacity is the next of 1024 / 2048 (a window of c units gives blocks of at
// most c rows and 1.5c entries, or a single row).
static void launch_adaptive(const CsrSegment& g, int chunk, bool accum, const double* x, double* y,
                            hipStream_t s) {
  if (chunk > 1024)
    launch_adaptive_t<2048>(g, accum, x, y, s);
  else
    launch_adaptive_t<1024>(g, accum, x, y, s);
}

// {first row, first entry} of every block; blk[nblk] = {nrows, nnz}.
__global__ void blk_pairs(const int32_t* __restrict__ starts, const int32_t* __restrict__ rowptr,
                          int64_t nblk, int64_t nrows, int2* __restrict__ blk) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > nblk) return;
  const int r = b == nblk ? (int)nrows : starts[b];
  blk[b] = make_int2(r, rowptr[r]);
}

// ---------------------------------------------------------------------------
// Scalar: thread per row (src/csrspmv.cu:39-57). Products are rounded before
// the add (no contraction), left to right: bit-identical to the CPU loop of
// src/ref.cpp:26-37.
__global__ __launch_bounds__(kWG) void csr_scalar_f64(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ col,
                                                      const double* __restrict__ val,
                                                      const double* __restrict__ x,
                                                      double* __restrict__ y, int64_t nrows) {
#pragma clang fp contract(off)
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s = 0.0;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) s += val[j] * x[col[j]];
  y[r] = s;
}

// ---------------------------------------------------------------------------
// Balanced (src/csrspmv.cu:59-102): one thread per 16-nonzero chunk of a row.
// Chunk partials land in a workspace and a second pass adds each row's
// partials in chunk order, so no atomics and a fixed summation order.
constexpr int kBalChunk = 16;

__global__ void bal_chunk_count(const int32_t* __restrict__ rowptr, int64_t nrows,
                                int32_t* __restrict__ cnt) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < nrows) cnt[r] = (rowptr[r + 1] - rowptr[r] + kBalChunk - 1) / kBalChunk;
  if (r == nrows) cnt[r] = 0;
}

__global__ void bal_chunk_rows(const int32_t* __restrict__ cptr, int64_t nrows,
                               int32_t* __restrict__ crow) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  for (int c = cptr[r]; c < cptr[r + 1]; ++c) crow[c] = (int32_t)r;
}

__global__ __launch_bounds__(kWG) void bal_partials_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, const int32_t* __restrict__ cptr,
    const int32_t* __restrict__ crow, int64_t nchunk, double* __restrict__ part) {
  int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nchunk) return;
  const int r = crow[c];
  const int s = rowptr[r] + (int)(c - cptr[r]) * kBalChunk;
  const int e = min(s + kBalChunk, rowptr[r + 1]);
  double acc = 0.0;
  for (int j = s; j < e; ++j) acc += val[j] * x[col[j]];
  part[c] = acc;
}

__global__ __launch_bounds__(kWG) void bal_combine_f64(const int32_t* __restrict__ cptr,
                                                       const double* __restrict__ part,
                                                       int64_t nrows, double* __restrict__ y) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  double s = 0.0;
  for (int c = cptr[r]; c < cptr[r + 1]; ++c) s += part[c];
  y[r] = s;
}

// ---------------------------------------------------------------------------
// Short/long split (src/spmv.cu:131-196): rows with at most 128 nonzeros are
// thread-per-row, longer ones wave-per-row. Both kernels read the original
// CSR through the row maps; the reference's separate short/long CSR copies
// are not materialised.
constexpr int kSplitThreshold = 128;

__global__ void split_flags(const int32_t* __restrict__ rowptr, int64_t nrows,
                            int32_t* __restrict__ is_long, int32_t* __restrict__ is_short) {
  int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nrows) return;
  int lng = (rowptr[r + 1] - rowptr[r]) > kSplitThreshold;
  is_long[r] = lng;
  is_short[r] = !lng;
}

__global__ __launch_bounds__(kWG) void split_short_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
    const int32_t* __restrict__ map, int64_t n) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int r = map[t];
  double s = 0.0;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) s += val[j] * x[col[j]];
  y[r] = s;
}

__global__ __launch_bounds__(kWG) void split_long_f64(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
    const int32_t* __restrict__ map, int64_t n) {
  int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (w >= n) return;
  const int r = map[w];
  double a0 = 0.0, a1 = 0.0;
  int j = rowptr[r] + lane;
  const int e = rowptr[r + 1];
  for (; j + kWave < e; j += 2 * kWave) {
    a0 += ld_stream(&val[j]) * x[ld_stream(&col[j])];
    a1 += ld_stream(&val[j + kWave]) * x[ld_stream(&col[j + kWave])];
  }
  if (j < e) a0 += ld_stream(&val[j]) * x[ld_stream(&col[j])];
  double s = group_sum<kWave>(a0 + a1);
  if (lane == 0) y[r] = s;
}

// FP16 split SpMV (src/spmv.cu:13-27, 44-75): half values and x, products
// and sums in FP32 (the reference accumulates in half), y = half(sum)
// returned as float, the rounding of ref_spmv_fp16 (src/ref.cpp:3-13).
__device__ __forceinline__ float h2f(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}
__device__ __forceinline__ float round_half(float v) { return (float)(_Float16)v; }

__global__ __launch_bounds__(kWG) void split_short_f16(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const uint16_t* __restrict__ val, const uint16_t* __restrict__ x, float* __restrict__ y,
    const int32_t* __restrict__ map, int64_t n) {
#pragma clang fp contract(off)
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int r = map[t];
  float s = 0.f;
  for (int j = rowptr[r], e = rowptr[r + 1]; j < e; ++j) s += h2f(val[j]) * h2f(x[col[j]]);
  y[r] = round_half(s);
}

__global__ __launch_bounds__(kWG) void split_long_f16(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ col,
    const uint16_t* __restrict__ val, const uint16_t* __restrict__ x, float* __restrict__ y,
    const int32_t* __restrict__ map, int64_t n) {
  int64_t w = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (w >= n) return;
  const int r = map[w];
  float s = 0.f;
  for (int j = rowptr[r] + lane, e = rowptr[r + 1]; j < e; j += kWave)
    s += h2f(val[j]) * h2f(x[col[j]]);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) y[r] = round_half(s);
}

bsp_status csr_plan_run_f16(const CsrPlan* p, const uint16_t* x, float* y, hipStream_t s) {
  const bsp_csr& A = p->A;
  if (p->method != BSP_CSR_SPLIT) return fail(BSP_ERR_UNSUPPORTED, "FP16 SpMV uses the split plan");
  if (A.nrows == 0) return BSP_OK;
  const auto* v = static_cast<const uint16_t*>(A.val);
  if (p->nblk > 0)
    split_short_f16<<<cdiv(p->nblk, kWG), kWG, 0, s>>>(A.rowptr, A.colidx, v, x, y, p->a.get(),
                                                       p->nblk);
  if (p->nlong > 0)
    split_long_f16<<<cdiv(p->nlong * kWave, kWG), kWG, 0, s>>>(A.rowptr, A.colidx, v, x, y,
                                                                p->b.get(), p->nlong);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

// ---------------------------------------------------------------------------
// Plan construction.

static bsp_status stable_partition(const int32_t* flags, int64_t n, DevBuf<int32_t>& out,
                                   int64_t* count, hipStream_t s) {
  DevBuf<int32_t> pos;
  BSP_TRY(pos.alloc(n));
  BSP_TRY(exclusive_scan_i32(flags, pos.get(), n, s));
  int32_t last_pos = 0, last_flag = 0;
  BSP_TRY(fetch(pos.get() + n - 1, &last_pos, s));
  BSP_TRY(fetch(flags + n - 1, &last_flag, s));
  *count = (int64_t)last_pos + last_flag;
  BSP_TRY(out.alloc(*count + 1));
  scatter_flagged<<<cdiv(n, kWG), kWG, 0, s>>>(flags, pos.get(), n, out.get());
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}


// Column-segment support for the adaptive layout ----------------------------

constexpr int64_t kDefaultSegCols = int64_t(1) << 19;  // 4 MB of FP64 x (one XCD's L2)
constexpr int64_t kXcdSegCols = int64_t(1) << 18;      // 2 MB slices when XCDs are pinned

// The three plan-builder passes below run a wave per row (lanes stride the
// row's entries), so a 5,000-entry row is 79 steps of one wave rather than
// one thread's loop.

// Validation of a device CSR, wave per row: bit 1 of *bad for a row pointer
// out of order or out of [0, nnz], bit 2 for a column index outside
// [0, ncols), bit 4 for columns not strictly ascending in a row.
__global__ __launch_bounds__(kWG) void csr_check_rows(const int32_t* __restrict__ rowptr,
                                                      const int32_t* __restrict__ col, int64_t n,
                                                      int64_t ncols, int64_t nnz,
                                                      int* __restrict__ bad, int64_t first) {
  const int64_t r = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= n) return;
  const int64_t s0 = rowptr[r], e0 = rowptr[r + 1];
  if (s0 > e0 || s0 < 0 || e0 > nnz || (r == 0 && s0 != 0) || (r == n - 1 && e0 != nnz)) {
    if (lane == 0) atomicOr(bad, 1);
    return;
  }
  int f = 0;
  for (int64_t j = s0 + lane; j < e0; j += kWave) {
    const int c = col[j];
    if (c < 0 || c >= ncols) f |= 2;
    if (j > s0 && col[j - 1] >= c) f |= 4;
  }
  if (f) atomicOr(bad, f);
}

// Counts entries farther than `half` columns from the (scaled) diagonal and
// rows whose columns are not strictly ascending.
__global__ __launch_bounds__(kWG) void csr_scatter_stats(const int32_t* __restrict__ rowptr,
                                                         const int32_t* __restrict__ col,
                                                         int64_t nrows, int64_t ncols, int64_t half,
                                                         unsigned long long* far_count,
                                                         int* unsorted) {
  // a fixed grid of waves strides over the rows, so the totals take one
  // atomic per wave, not one per row
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t nw = (int64_t)gridDim.x * blockDim.x / kWave;
  unsigned long long far = 0;
  int bad = 0;
  for (int64_t r = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; r < nrows; r += nw) {
    const int64_t d = r * ncols / nrows;
    const int j0 = rowptr[r], e = rowptr[r + 1];
    for (int j = j0 + lane; j < e; j += kWave) {
      const int c = col[j];
      far += (c > d + half || c < d - half);
      bad |= j > j0 && c <= col[j - 1];
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) far += __shfl_xor(far, o, 64);
  bad = __any(bad);
  if (lane == 0 && far) atomicAdd(far_count, far);
  if (lane == 0 && bad) atomicOr(unsorted, 1);
}

static int scatter_stats_grid(int64_t nrows) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(cdiv(nrows * kWave, kWG), 2048));
}

// cnt[s*(nrows+1) + r] = entries of row r in column segment s (columns
// ascend, so they form one run): the lane at the run's first entry j adds -j,
// the lane at its last entry adds j + 1 (cnt starts at zero).
__global__ __launch_bounds__(kWG) void seg_count(const int32_t* __restrict__ rowptr,
                                                 const int32_t* __restrict__ col, int64_t nrows,
                                                 int64_t width, int32_t* __restrict__ cnt,
                                                 int64_t first) {
  const int64_t r = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= nrows) return;
  const int j0 = rowptr[r], e = rowptr[r + 1];
  for (int j = j0 + lane; j < e; j += kWave) {
    const int64_t sg = col[j] / width;
    int32_t* c = cnt + sg * (nrows + 1) + r;
    if (j == j0 || col[j - 1] / width != sg) atomicAdd(c, -j);
    if (j == e - 1 || col[j + 1] / width != sg) atomicAdd(c, j + 1);
  }
}

// Copies each row's run of segment-s entries (contiguous: columns ascend) to
// segment s's sub-CSR: entry j of the row goes to base[s] + segrp[s][r] plus
// its offset in the run, the run starting after the row's entries of the
// segments before s.
__global__ __launch_bounds__(kWG) void seg_scatter(const int32_t* __restrict__ rowptr,
                                                   const int32_t* __restrict__ col,
                                                   const double* __restrict__ val, int64_t nrows,
                                                   int64_t width, const int32_t* __restrict__ segrp,
                                                   const int64_t* __restrict__ base,
                                                   int32_t* __restrict__ ocol,
                                                   double* __restrict__ oval, int64_t first) {
  const int64_t r = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (r >= nrows) return;
  const int j0 = rowptr[r], e = rowptr[r + 1];
  for (int j = j0 + lane; j < e; j += kWave) {
    const int c = col[j];
    const int64_t sg = c / width;
    int64_t before = 0;
    for (int64_t q = 0; q < sg; ++q)
      before += segrp[q * (nrows + 1) + r + 1] - segrp[q * (nrows + 1) + r];
    const int64_t dst = base[sg] + segrp[sg * (nrows + 1) + r] + (j - j0 - before);
    ocol[dst] = c;
    oval[dst] = val[j];
  }
}

static bsp_status adaptive_partition(CsrSegment& g, int64_t n, int chunk, hipStream_t s,
                                     int rshift = 0, int longrow = 0) {
  DevBuf<int32_t> flags;
  BSP_TRY(flags.alloc(n));
  csr_block_flags<<<cdiv(n, kWG), kWG, 0, s>>>(g.rowptr, n, chunk, rshift,
                                               longrow > 0 ? longrow : chunk / 2, flags.get());
  BSP_HIP(hipGetLastError());
  DevBuf<int32_t> starts;
  BSP_TRY(stable_partition(flags.get(), n, starts, &g.nblk, s));
  BSP_TRY(g.blk.alloc(2 * (g.nblk + 1)));
  blk_pairs<<<cdiv(g.nblk + 1, kWG), kWG, 0, s>>>(starts.get(), g.rowptr, g.nblk, n,
                                                  reinterpret_cast<int2*>(g.blk.get()));
  BSP_HIP(hipGetLastError());
  BSP_HIP(hipStreamSynchronize(s));
  return BSP_OK;
}

// Sorted blocks: every block's entries reordered by column (stable), with
// each entry's row-order position in its block kept as a 2-byte slot. One
// radix sort over all segments on (global block, column) keys.
static bsp_status sort_blocks(CsrPlan* p, hipStream_t s) {
  const int64_t total = p->A.nnz;
  if (total == 0) return BSP_OK;
  int64_t nblk = 0;
  for (const auto& g : p->seg) nblk += g.nblk;
  DevBuf<int64_t> keys, keys2, gstart;
  DevBuf<int32_t> idx, idx2;
  BSP_TRY(keys.alloc(total));
  BSP_TRY(keys2.alloc(total));
  BSP_TRY(idx.alloc(total));
  BSP_TRY(idx2.alloc(total));
  BSP_TRY(gstart.alloc(nblk));
  int64_t block0 = 0;
  for (const auto& g : p->seg) {
    const int64_t entry0 = g.col - p->seg_col.get();
    if (g.nblk > 0)
      in_slices(g.nblk, kWG, [&](int64_t f, int64_t c) {
        sorted_keys<<<c, kWG, 0, s>>>(reinterpret_cast<const int2*>(g.blk.get()), g.col, block0,
                                      entry0, keys.get(), idx.get(), gstart.get(), f);
      });
    block0 += g.nblk;
  }
  BSP_HIP(hipGetLastError());
  int bits = 1;
  while ((int64_t(1) << bits) < nblk) ++bits;
  BSP_TRY(sort_pairs_i64key(keys.get(), keys2.get(), idx.get(), idx2.get(), total, 32 + bits, s));
  DevBuf<int32_t> col;
  DevBuf<double> val;
  BSP_TRY(col.alloc(total));
  BSP_TRY(val.alloc(total));
  BSP_TRY(p->slot.alloc(total));
  sorted_apply<<<cdiv(total, kWG), kWG, 0, s>>>(keys2.get(), idx2.get(), gstart.get(),
                                                p->seg_col.get(), p->seg_val.get(), total,
                                                col.get(), val.get(), p->slot.get());
  BSP_HIP(hipGetLastError());
  BSP_HIP(hipStreamSynchronize(s));
  // the blocks keep their entry ranges: segment pointers move with the arrays
  for (auto& g : p->seg) {
    const int64_t entry0 = g.col - p->seg_col.get();
    g.col = col.get() + entry0;
    g.val = val.get() + entry0;
  }
  p->seg_col = std::move(col);
  p->seg_val = std::move(val);
  return BSP_OK;
}

static bsp_status adaptive_build(CsrPlan* p, hipStream_t s) {
  const bsp_csr& A = p->A;
  const int64_t n = A.nrows;
  int64_t W = p->opt.seg_cols;
  if (W == 0 && A.ncols > kDefaultSegCols && A.nnz > 0) {
    DevBuf<unsigned long long> far;
    DevBuf<int> unsorted;
    BSP_TRY(far.alloc(1));
    BSP_TRY(unsorted.alloc(1));
    BSP_HIP(hipMemsetAsync(far.get(), 0, sizeof(unsigned long long), s));
    BSP_HIP(hipMemsetAsync(unsorted.get(), 0, sizeof(int), s));
    csr_scatter_stats<<<scatter_stats_grid(n), kWG, 0, s>>>(A.rowptr, A.colidx, n, A.ncols,
                                                            kDefaultSegCols / 2, far.get(),
                                                            unsorted.get());
    BSP_HIP(hipGetLastError());
    unsigned long long nfar = 0;
    int uns = 0;
    BSP_TRY(fetch(far.get(), &nfar, s));
    BSP_TRY(fetch(unsorted.get(), &uns, s));
    W = (!uns && (int64_t)nfar * 4 > A.nnz) ? kDefaultSegCols : 0;
    if (W > 0 && (p->opt.seg_mode == 0 || p->opt.seg_mode == 2)) {
      // one launch, one segment per XCD group: the smallest power-of-two
      // segment count (<= 8) whose x slices are at most 2 MB
      int64_t S = 1;
      while (S < 8 && cdiv(A.ncols, S) > kXcdSegCols) S *= 2;
      if (cdiv(A.ncols, S) <= kXcdSegCols) W = cdiv(A.ncols, S);
    }
  } else if (W > 0 && A.nnz > 0) {
    DevBuf<unsigned long long> far;
    DevBuf<int> unsorted;
    BSP_TRY(far.alloc(1));
    BSP_TRY(unsorted.alloc(1));
    BSP_HIP(hipMemsetAsync(far.get(), 0, sizeof(unsigned long long), s));
    BSP_HIP(hipMemsetAsync(unsorted.get(), 0, sizeof(int), s));
    csr_scatter_stats<<<scatter_stats_grid(n), kWG, 0, s>>>(A.rowptr, A.colidx, n, A.ncols, W / 2,
                                                            far.get(), unsorted.get());
    BSP_HIP(hipGetLastError());
    int uns = 0;
    BSP_TRY(fetch(unsorted.get(), &uns, s));
    if (uns) return fail(BSP_ERR_UNSORTED, "column segmentation needs ascending columns per row");
  }
  if (W < 0) W = 0;
  const int64_t nseg = W > 0 ? cdiv(A.ncols, W) : 1;
  p->seg.clear();
  p->seg.resize(nseg);
  if (nseg == 1) {
    p->seg_width = 0;
    CsrSegment& g = p->seg[0];
    g.rowptr = A.rowptr;
    g.col = A.colidx;
    g.val = static_cast<const double*>(A.val);
    g.nnz = A.nnz;
  } else {
    p->seg_width = W;
    DevBuf<int32_t> cnt;
    BSP_TRY(cnt.alloc(nseg * (n + 1)));
    BSP_HIP(hipMemsetAsync(cnt.get(), 0, sizeof(int32_t) * nseg * (n + 1), s));
    in_slices(n, kWave, [&](int64_t f, int64_t c) {
      seg_count<<<cdiv(c * kWave, kWG), kWG, 0, s>>>(A.rowptr, A.colidx, n, W, cnt.get(), f);
    });
    BSP_HIP(hipGetLastError());
    DevBuf<int32_t> segrp;
    BSP_TRY(segrp.alloc(nseg * (n + 1)));
    std::vector<int64_t> base(nseg + 1, 0);
    for (int64_t k = 0; k < nseg; ++k) {
      BSP_TRY(exclusive_scan_i32(cnt.get() + k * (n + 1), segrp.get() + k * (n + 1), n + 1, s));
      int32_t tot = 0;
      BSP_TRY(fetch(segrp.get() + k * (n + 1) + n, &tot, s));
      base[k + 1] = base[k] + tot;
    }
    DevBuf<int64_t> dbase;
    BSP_TRY(dbase.alloc(nseg));
    BSP_HIP(hipMemcpyAsync(dbase.get(), base.data(), sizeof(int64_t) * nseg,
                           hipMemcpyHostToDevice, s));
    BSP_TRY(p->seg_col.alloc(A.nnz));
    BSP_TRY(p->seg_val.alloc(A.nnz));
    in_slices(n, kWave, [&](int64_t f, int64_t c) {
      seg_scatter<<<cdiv(c * kWave, kWG), kWG, 0, s>>>(
          A.rowptr, A.colidx, static_cast<const double*>(A.val), n, W, segrp.get(), dbase.get(),
          p->seg_col.get(), p->seg_val.get(), f);
    });
    BSP_HIP(hipGetLastError());
    for (int64_t k = 0; k < nseg; ++k) {
      CsrSegment& g = p->seg[k];
      BSP_TRY(g.own_rowptr.alloc(n + 1));
      BSP_HIP(hipMemcpyAsync(g.own_rowptr.get(), segrp.get() + k * (n + 1),
                             sizeof(int32_t) * (n + 1), hipMemcpyDeviceToDevice, s));
      g.rowptr = g.own_rowptr.get();
      g.col = p->seg_col.get() + base[k];
      g.val = p->seg_val.get() + base[k];
      g.nnz = base[k + 1] - base[k];
    }
    BSP_HIP(hipStreamSynchronize(s));
  }
  // Default window: the largest of 2048, 1024, 512, 256 units that still
  // gives at least 8 blocks per CU (larger blocks amortise the per-block
  // start-up; measured on the 27-point 64^3 stencil: 2048 -> 21.2 us,
  // 1024 -> 22.9 us, 512 -> 29.8 us); segmented plans always use 2048.
  if (p->opt.chunk >= 128 && p->opt.chunk <= 2048) {
    p->chunk = p->opt.chunk;
  } else if (nseg > 1) {
    p->chunk = 2048;
  } else {
    int dev = 0, cus = 0;
    BSP_HIP(hipGetDevice(&dev));
    BSP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t units = n + A.nnz, want = 8 * (int64_t)std::max(cus, 1);
    p->chunk = 256;
    for (int c : {2048, 1024, 512})
      if (units / c >= want) {
        p->chunk = c;
        break;
      }
  }
  // sorted blocks (seg_mode 2): only for the single XCD-pinned launch
  p->sorted = p->opt.seg_mode == 2 && (nseg == 2 || nseg == 4 || nseg == 8);
  p->nblk = 0;
  for (auto& g : p->seg) {
    if (p->sorted)
      BSP_TRY(adaptive_partition(g, n, kSortChunk, s, kSortRowShift, kSortLong));
    else
      BSP_TRY(adaptive_partition(g, n, p->chunk, s));
    p->nblk += g.nblk;
  }
  if (p->sorted) BSP_TRY(sort_blocks(p, s));
  // 16-bit column offsets, when asked for and every block allows them (its
  // kernel loads values two at a time: 16-byte aligned values needed). Not
  // the default: on the 27-point 64^3 stencil the plan reads 16% fewer bytes
  // but runs 24.2 us against 22.4 us event-timed (the launch is bound by
  // each workgroup's chain of dependent loads, not by the bytes)
  p->narrow = false;
  if (nseg == 1 && p->opt.narrow_cols > 0 && A.nnz > 0 &&
      reinterpret_cast<uintptr_t>(A.val) % 16 == 0) {
    const CsrSegment& g = p->seg[0];
    const auto* blk = reinterpret_cast<const int2*>(g.blk.get());
    DevBuf<int> bad;
    BSP_TRY(bad.alloc(1));
    BSP_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
    in_slices(g.nblk, kWave, [&](int64_t f, int64_t c) {
      narrow_check<<<cdiv(c * kWave, kWG), kWG, 0, s>>>(g.col, blk, g.nblk, bad.get(), f);
    });
    BSP_HIP(hipGetLastError());
    int b = 1;
    BSP_TRY(fetch(bad.get(), &b, s));
    if (!b) {
      // one element of padding: the kernel loads offsets two at a time
      BSP_TRY(p->col16.alloc(A.nnz + 2));
      BSP_HIP(hipMemsetAsync(p->col16.get() + A.nnz, 0, 2 * sizeof(int16_t), s));
      in_slices(g.nblk, kWave, [&](int64_t f, int64_t c) {
        narrow_fill<<<cdiv(c * kWave, kWG), kWG, 0, s>>>(g.col, blk, g.nblk, p->col16.get(), f);
      });
      BSP_HIP(hipGetLastError());
      BSP_HIP(hipStreamSynchronize(s));
      p->narrow = true;
    }
  }
  // XCD-pinned single launch for 2, 4 or 8 segments (unless sequential passes
  // were asked for)
  p->xcd_mode = false;
  if (nseg == 2 || nseg == 4 || nseg == 8) {
    if (p->opt.seg_mode == 0 || p->opt.seg_mode == 2) {
      p->xcd_mode = true;
      BSP_TRY(p->part.alloc(nseg * n));
      std::vector<SegDesc> tab(nseg);
      int64_t most = 0;
      for (int64_t k = 0; k < nseg; ++k) {
        const CsrSegment& g = p->seg[k];
        tab[k] = SegDesc{g.rowptr, g.col, g.val, reinterpret_cast<const int2*>(g.blk.get()),
                         g.nblk, p->part.get() + k * n,
                         p->sorted ? p->slot.get() + (g.col - p->seg_col.get()) : nullptr};
        most = std::max(most, cdiv(g.nblk, 8 / nseg));
      }
      p->xcd_grid = 8 * most;
      BSP_TRY(p->segtab.alloc(nseg * sizeof(SegDesc)));
      BSP_HIP(hipMemcpyAsync(p->segtab.get(), tab.data(), nseg * sizeof(SegDesc),
                             hipMemcpyHostToDevice, s));
      BSP_HIP(hipStreamSynchronize(s));
    }
  }
  return BSP_OK;
}

bsp_status csr_plan_build(CsrPlan* p, hipStream_t s) {
  const bsp_csr& A = p->A;
  const int64_t n = A.nrows;
  if (n == 0) return BSP_OK;
  switch (p->method) {
    case BSP_CSR_ADAPTIVE:
      BSP_TRY(adaptive_build(p, s));
      break;
    case BSP_CSR_SCALAR:
      p->nblk = cdiv(n, kWG);
      break;
    case BSP_CSR_BALANCED: {
      BSP_TRY(p->a.alloc(n + 1));
      DevBuf<int32_t> cnt;
      BSP_TRY(cnt.alloc(n + 1));
      bal_chunk_count<<<cdiv(n + 1, kWG), kWG, 0, s>>>(A.rowptr, n, cnt.get());
      BSP_HIP(hipGetLastError());
      BSP_TRY(exclusive_scan_i32(cnt.get(), p->a.get(), n + 1, s));
      int32_t nchunk = 0;
      BSP_TRY(fetch(p->a.get() + n, &nchunk, s));
      p->nblk = nchunk;
      BSP_TRY(p->b.alloc(nchunk));
      BSP_TRY(p->work.alloc(nchunk));
      bal_chunk_rows<<<cdiv(n, kWG), kWG, 0, s>>>(p->a.get(), n, p->b.get());
      BSP_HIP(hipGetLastError());
      BSP_HIP(hipStreamSynchronize(s));
      break;
    }
    case BSP_CSR_SPLIT: {
      DevBuf<int32_t> is_long, is_short;
      BSP_TRY(is_long.alloc(n));
      BSP_TRY(is_short.alloc(n));
      split_flags<<<cdiv(n, kWG), kWG, 0, s>>>(A.rowptr, n, is_long.get(), is_short.get());
      BSP_HIP(hipGetLastError());
      BSP_TRY(stable_partition(is_short.get(), n, p->a, &p->nblk, s));
      BSP_TRY(stable_partition(is_long.get(), n, p->b, &p->nlong, s));
      BSP_HIP(hipStreamSynchronize(s));
      break;
    }
    default:
      return fail(BSP_ERR_INVALID, "unknown CSR SpMV method " + std::to_string(p->method));
  }
  return BSP_OK;
}

bsp_status csr_plan_probe(const CsrPlan* p, int with_values, const double* x, double* sink,
                          int64_t sink_len, int64_t* grid_out, hipStream_t s) {
  if (p->method != BSP_CSR_ADAPTIVE || !p->xcd_mode)
    return fail(BSP_ERR_UNSUPPORTED, "the gather probe runs over an XCD-pinned segmented plan");
  const int nseg = (int)p->seg.size();
  int64_t most = 0;
  for (int k = 0; k < nseg; ++k)
    most = std::max(most, cdiv(cdiv(p->seg[k].nnz, 2048), 8 / nseg));
  const int64_t grid = 8 * most;
  if (grid_out) *grid_out = grid;
  if (!sink || grid == 0) return BSP_OK;  // grid query
  BSP_REQUIRE(sink_len >= grid, "sink must hold one double per workgroup");
  const auto* tab = reinterpret_cast<const SegDesc*>(p->segtab.get());
  if (with_values)
    probe_gather<true><<<grid, kWG, 0, s>>>(tab, nseg, x, sink);
  else
    probe_gather<false><<<grid, kWG, 0, s>>>(tab, nseg, x, sink);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status gather_probe(const double* x, int64_t n, int nslices, int64_t gathers, double* sink,
                        int64_t sink_len, int64_t* grid_out, hipStream_t s) {
  BSP_REQUIRE(nslices == 1 || nslices == 2 || nslices == 4 || nslices == 8,
              "slices must be 1, 2, 4 or 8");
  BSP_REQUIRE(n > 0 && n % nslices == 0 && gathers >= 0, "bad probe sizes");
  const int64_t len = n / nslices;
  int lg = 0;
  while ((int64_t(1) << lg) < len) ++lg;
  BSP_REQUIRE((int64_t(1) << lg) == len && lg <= 30, "slice length must be a power of two");
  const int64_t grid = cdiv(cdiv(gathers, kProbePer), kWG);
  BSP_REQUIRE(grid * kWG < kDispatchLimit, "too many gathers for one launch");
  if (grid_out) *grid_out = grid;
  if (!sink || grid == 0) return BSP_OK;  // grid query
  BSP_REQUIRE(sink_len >= grid, "sink must hold one double per workgroup");
  probe_random_gather<<<grid, kWG, 0, s>>>(x, nslices, lg, sink);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status line_probe(const double* x, int64_t n, int nslices, int64_t nbytes, double* sink,
                      int64_t sink_len, int64_t* grid_out, hipStream_t s) {
  BSP_REQUIRE(nslices == 1 || nslices == 2 || nslices == 4 || nslices == 8,
              "slices must be 1, 2, 4 or 8");
  BSP_REQUIRE(n > 0 && n % nslices == 0 && nbytes >= 0, "bad probe sizes");
  const int64_t len = n / nslices;
  int lg = 0;
  while ((int64_t(1) << lg) < len) ++lg;
  BSP_REQUIRE((int64_t(1) << lg) == len && lg >= 7 && lg <= 30,
              "slice length must be a power of two of at least 128 doubles");
  const int64_t grid = cdiv(cdiv(nbytes, int64_t(16) * kProbePer), kWG);
  BSP_REQUIRE(grid * kWG < kDispatchLimit, "too many bytes for one launch");
  if (grid_out) *grid_out = grid;
  if (!sink || grid == 0) return BSP_OK;  // grid query
  BSP_REQUIRE(sink_len >= grid, "sink must hold one double per workgroup");
  probe_line_read<<<grid, kWG, 0, s>>>(x, nslices, lg, sink);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status csr_check(const bsp_csr& A, hipStream_t s, int* flags) {
  DevBuf<int> bad;
  BSP_TRY(bad.alloc(1));
  BSP_HIP(hipMemsetAsync(bad.get(), 0, sizeof(int), s));
  in_slices(A.nrows, kWave, [&](int64_t f, int64_t c) {
    csr_check_rows<<<cdiv(c * kWave, kWG), kWG, 0, s>>>(A.rowptr, A.colidx, A.nrows, A.ncols,
                                                        A.nnz, bad.get(), f);
  });
  BSP_HIP(hipGetLastError());
  BSP_TRY(fetch(bad.get(), flags, s));
  return BSP_OK;
}

bsp_status csr_plan_run(const CsrPlan* p, const double* x, double* y, hipStream_t s) {
  const bsp_csr& A = p->A;
  const int64_t n = A.nrows;
  if (n == 0) return BSP_OK;
  const auto* rp = A.rowptr;
  const auto* ci = A.colidx;
  const auto* v = static_cast<const double*>(A.val);
  switch (p->method) {
    case BSP_CSR_ADAPTIVE:
      if (p->xcd_mode) {
        const int nseg = (int)p->seg.size();
        const auto* tab = reinterpret_cast<const SegDesc*>(p->segtab.get());
        if (p->sorted)
          csr_sorted_xcd_f64<<<p->xcd_grid, kSortWG, kSortPad, s>>>(tab, nseg, x);
        else if (p->chunk > 1024)
          csr_adaptive_xcd_f64<2048><<<p->xcd_grid, kWG, 0, s>>>(tab, nseg, x);
        else
          csr_adaptive_xcd_f64<1024><<<p->xcd_grid, kWG, 0, s>>>(tab, nseg, x);
        BSP_HIP(hipGetLastError());
        if (n % 2 == 0 && ((uintptr_t)y & 15) == 0)
          sum_partials2<<<cdiv(n / 2, kWG), kWG, 0, s>>>(p->part.get(), nseg, n, y);
        else
          sum_partials<<<cdiv(n, kWG), kWG, 0, s>>>(p->part.get(), nseg, n, y);
        break;
      }
      if (p->narrow) {
        const CsrSegment& g = p->seg[0];
        const int2* blk = reinterpret_cast<const int2*>(g.blk.get());
        if (g.nblk == 0) break;
        if (p->chunk > 1024)
          csr_adaptive16_f64<2048><<<g.nblk, kWG, 0, s>>>(g.rowptr, p->col16.get(), g.val, x, y,
                                                          blk, g.nblk);
        else
          csr_adaptive16_f64<1024><<<g.nblk, kWG, 0, s>>>(g.rowptr, p->col16.get(), g.val, x, y,
                                                          blk, g.nblk);
        break;
      }
      for (size_t k = 0; k < p->seg.size(); ++k) {
        const CsrSegment& g = p->seg[k];
        if (g.nblk == 0) continue;
        launch_adaptive(g, p->chunk, k > 0, x, y, s);
      }
      break;
    case BSP_CSR_SCALAR:
      csr_scalar_f64<<<cdiv(n, kWG), kWG, 0, s>>>(rp, ci, v, x, y, n);
      break;
    case BSP_CSR_BALANCED:
      if (p->nblk > 0)
        bal_partials_f64<<<cdiv(p->nblk, kWG), kWG, 0, s>>>(rp, ci, v, x, p->a.get(), p->b.get(),
                                                             p->nblk, p->work.get());
      bal_combine_f64<<<cdiv(n, kWG), kWG, 0, s>>>(p->a.get(), p->work.get(), n, y);
      break;
    case BSP_CSR_SPLIT:
      if (p->nblk > 0)
        split_short_f64<<<cdiv(p->nblk, kWG), kWG, 0, s>>>(rp, ci, v, x, y, p->a.get(), p->nblk);
      if (p->nlong > 0)
        split_long_f64<<<cdiv(p->nlong * kWave, kWG), kWG, 0, s>>>(rp, ci, v, x, y, p->b.get(),
                                                                    p->nlong);
      break;
    default:
      return fail(BSP_ERR_INVALID, "unknown CSR SpMV method");
  }
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

}  // namespace bsp

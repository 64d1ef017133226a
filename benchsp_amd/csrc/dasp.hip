// DASP layout: GPU builder (bit-exact with the reference's host builder,
// src/dasp.cu:814-1503) and a fused FP64 SpMV kernel over it.
#include <memory>
#include <vector>

#include "bsp_hip.hpp"
#include "dasp_kernel.hpp"

struct bsp_dasp_s {
  bsp_dasp_view v{};
  bsp::DevBuf<int32_t> order_rid, short_cid, long_rpt, long_cid, rid_by_warp, block_ptr, irreg_rpt,
      reg_cid, irreg_cid, ridA;
  bsp::DevBuf<double> short_val, long_val, reg_val, irreg_val;
};

namespace bsp {
namespace {

constexpr int kTB = 256;
constexpr int BS = 8, MM = 8, MK = 4;  // BlockSize, MMA_M, MMA_K (dasp.cu:27-33)
enum { C1 = 0, C2, C3, C4, CLONG, CBLOCK, CZERO, NCLASS };

__global__ void dasp_classify(const int32_t* __restrict__ rowptr, int64_t n, int32_t block_longest,
                              int32_t* __restrict__ flags /* NCLASS*n */) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int L = rowptr[i + 1] - rowptr[i];
  int c;  // dasp.cu:827-858 (the test order there: 1, 3, 2, 0, 4, long, block)
  if (L == 1) c = C1;
  else if (L == 3) c = C3;
  else if (L == 2) c = C2;
  else if (L == 0) c = CZERO;
  else if (L == 4) c = C4;
  else if (L >= block_longest) c = CLONG;
  else c = CBLOCK;
#pragma unroll
  for (int k = 0; k < NCLASS; ++k) flags[(int64_t)k * n + i] = (k == c);
}

__global__ void scatter_class(const int32_t* __restrict__ flags, const int32_t* __restrict__ pos,
                              int64_t n, int32_t* __restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) out[pos[i]] = (int32_t)i;
}

__global__ void gather_len(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ rows,
                           int64_t m, int32_t* __restrict__ len) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int r = rows[k];
  len[k] = rowptr[r + 1] - rowptr[r];
}

struct Lists {
  const int32_t *r1, *r2, *r3, *r4, *rl, *rb, *rz;
  int32_t s1, s2, s3, s4, nl, nb, nz, c13;  // s1/s3 after removing common_13
};

// dasp.cu:1290-1306
__global__ void dasp_order(Lists Ls, int64_t n, int32_t* __restrict__ order) {
  int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int64_t q = k;
  int32_t r;
  if (q < Ls.nl) { r = Ls.rl[q]; goto done; }
  q -= Ls.nl;
  if (q < Ls.nb) { r = Ls.rb[q]; goto done; }
  q -= Ls.nb;
  if (q < Ls.s1) { r = Ls.r1[q]; goto done; }
  q -= Ls.s1;
  if (q < 2 * (int64_t)Ls.c13) {
    const int64_t t = q / 16, w = q % 16;
    r = w < 8 ? Ls.r1[Ls.s1 + t * 8 + w] : Ls.r3[t * 8 + w - 8];
    goto done;
  }
  q -= 2 * (int64_t)Ls.c13;
  if (q < Ls.s3) { r = Ls.r3[Ls.c13 + q]; goto done; }
  q -= Ls.s3;
  if (q < Ls.s4) { r = Ls.r4[q]; goto done; }
  q -= Ls.s4;
  if (q < Ls.s2) { r = Ls.r2[q]; goto done; }
  q -= Ls.s2;
  r = Ls.rz[q];
done:
  order[k] = r;
}

// Short layout (dasp.cu:969-1043). One thread per short row.
__global__ void dasp_short_fill(Lists Ls, const int32_t* __restrict__ rowptr,
                                const int32_t* __restrict__ cid, const double* __restrict__ val,
                                int32_t f13, int32_t f34, double* __restrict__ sv,
                                int32_t* __restrict__ sc) {
  int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t total = (int64_t)Ls.s1 + Ls.c13 + Ls.s3 + Ls.s4 + Ls.s2;
  if (q >= total) return;
  auto put = [&](int64_t dst, int r, int t) {
    sv[dst] = val[rowptr[r] + t];
    sc[dst] = cid[rowptr[r] + t];
  };
  if (q < Ls.s1) {
    put(q, Ls.r1[q], 0);
    return;
  }
  q -= Ls.s1;
  if (q < Ls.c13) {
    const int64_t t = q / BS, j = q % BS;
    const int64_t base = Ls.s1 + t * MM * MK + j * MK;
    put(base, Ls.r1[Ls.s1 + q], 0);
    const int r3 = Ls.r3[q];
    for (int e = 0; e < 3; ++e) put(base + 1 + e, r3, e);
    return;
  }
  q -= Ls.c13;
  if (q < Ls.s3) {
    const int r = Ls.r3[Ls.c13 + q];
    for (int e = 0; e < 3; ++e) put(Ls.s1 + f13 + q * MK + e, r, e);
    return;
  }
  q -= Ls.s3;
  if (q < Ls.s4) {
    const int r = Ls.r4[q];
    for (int e = 0; e < 4; ++e) put(Ls.s1 + f13 + (Ls.s3 + q) * MK + e, r, e);
    return;
  }
  q -= Ls.s4;
  {
    const int64_t t = q / (BS * 2), j = q % (BS * 2);
    const int64_t base = Ls.s1 + f13 + f34 + t * MM * MK + (j % BS) * MK + (j / BS) * 2;
    const int r = Ls.r2[q];
    put(base, r, 0);
    put(base + 1, r, 1);
  }
}

__global__ void sum_lengths(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ rows,
                            int32_t m, unsigned long long* __restrict__ tot) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  const int r = rows[i];
  atomicAdd(tot, (unsigned long long)(rowptr[r + 1] - rowptr[r]));
}

__global__ void long_warps(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ rl,
                           int32_t nl, int32_t* __restrict__ cnt) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nl) {
    const int r = rl[i];
    cnt[i] = (rowptr[r + 1] - rowptr[r] + 63) / 64;  // MMA_M*MMA_K*loopNum_long
  }
  if (i == nl) cnt[i] = 0;
}

// wave per long row: contiguous copy from long_rpt[i]*64 (dasp.cu:1354-1369)
__global__ void long_fill(const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cid,
                          const double* __restrict__ val, const int32_t* __restrict__ rl, int32_t nl,
                          const int32_t* __restrict__ lrpt, double* __restrict__ lv,
                          int32_t* __restrict__ lc, int32_t* __restrict__ rbw) {
  const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (i >= nl) return;
  const int r = rl[i];
  const int64_t base = (int64_t)lrpt[i] * 64;
  for (int j = rowptr[r] + lane, z = lane, e = rowptr[r + 1]; j < e; j += kWave, z += kWave) {
    lv[base + z] = val[j];
    lc[base + z] = cid[j];
  }
  for (int w = lrpt[i] + lane; w < lrpt[i + 1]; w += kWave) rbw[w] = (int32_t)i;
}

// Regular slabs per 8-row block and irregular tails (dasp.cu:1382-1413).
__global__ void block_slabs(const int32_t* __restrict__ blen, int32_t nb, int32_t blocknum,
                            double threshold, int32_t* __restrict__ bcnt,
                            int32_t* __restrict__ icnt) {
  int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b > blocknum) return;
  if (b == blocknum) {
    bcnt[b] = 0;
    if (b == blocknum) icnt[nb] = 0;
    return;
  }
  const int rs = (int)b * BS;
  const int re = (b + 1) * BS >= nb ? nb : (int)(b + 1) * BS;
  int k = 1;
  while (true) {
    int block_nnz = 0;
    for (int r = rs; r < re; ++r) {
      const int L = blen[r];
      if (L / MK >= k) block_nnz += MK;
      else if (L / MK == k - 1) block_nnz += L % MK;
    }
    if (block_nnz >= threshold * MK * MM) {
      ++k;
      continue;
    }
    for (int r = rs; r < re; ++r) {
      const int L = blen[r];
      icnt[r] = L - (k - 1) * MK > 0 ? L - (k - 1) * MK : 0;
    }
    break;
  }
  bcnt[b] = (k - 1) * MK * MM;
}

// regular part, slab-major (dasp.cu:1442-1487): row g's entry t of block b
// goes to block_ptr[b] + (t/4)*32 + g*4 + t%4; irregular tail after it.
__global__ __launch_bounds__(kTB) void block_fill(
    const int32_t* __restrict__ rowptr, const int32_t* __restrict__ cid,
    const double* __restrict__ val, const int32_t* __restrict__ ridA, int32_t nb,
    const int32_t* __restrict__ bptr, const int32_t* __restrict__ irpt, double* __restrict__ rv,
    int32_t* __restrict__ rc, double* __restrict__ iv, int32_t* __restrict__ ic, int64_t first) {
  // wave per slot k, lanes striding its entries
  const int64_t k = first + ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (k >= nb) return;
  const int b = (int)(k / BS), g = (int)(k % BS);
  const int blocklen = (bptr[b + 1] - bptr[b]) / BS;
  const int r = ridA[k];
  const int a0 = rowptr[r], L = rowptr[r + 1] - a0;
  for (int t = lane; t < blocklen && t < L; t += kWave) {
    const int64_t d = bptr[b] + (t / MK) * (BS * MK) + g * MK + t % MK;
    rv[d] = val[a0 + t];
    rc[d] = cid[a0 + t];
  }
  const int i0 = irpt[k], il = irpt[k + 1] - i0;
  for (int t = lane; t < il; t += kWave) {
    iv[i0 + t] = val[a0 + L - il + t];
    ic[i0 + t] = cid[a0 + L - il + t];
  }
}

// ---- SpMV -------------------------------------------------------------------
__global__ __launch_bounds__(kTB) void dasp_spmv_f64(dasp::DaspArgs a) {
  dasp::dasp_block<false>(a, a.x, a.y, blockIdx.x);
}

__global__ __launch_bounds__(kTB) void dasp_spmv_f64_mfma(dasp::DaspArgs a) {
  dasp::dasp_block<true>(a, a.x, a.y, blockIdx.x);
}

}  // namespace

static bsp_status build_dasp(const bsp_csr& A, double threshold, int32_t block_longest,
                             hipStream_t s, bsp_dasp_s* h) {
  const int64_t n = A.nrows;
  bsp_dasp_view& v = h->v;
  v.nrows = n;
  v.ncols = A.ncols;
  v.nnz = A.nnz;
  const auto* aval = static_cast<const double*>(A.val);
  // class lists, each in ascending row order
  DevBuf<int32_t> flags, pos;
  BSP_TRY(flags.alloc(NCLASS * n));
  BSP_TRY(pos.alloc(n));
  if (n > 0) {
    dasp_classify<<<cdiv(n, kTB), kTB, 0, s>>>(A.rowptr, n, block_longest, flags.get());
    BSP_HIP(hipGetLastError());
  }
  DevBuf<int32_t> lists[NCLASS];
  int32_t cnt[NCLASS] = {0};
  for (int c = 0; c < NCLASS; ++c) {
    const int32_t* f = flags.get() + (int64_t)c * n;
    if (n > 0) {
      BSP_TRY(exclusive_scan_i32(f, pos.get(), n, s));
      int32_t lp = 0, lf = 0;
      BSP_TRY(fetch(pos.get() + n - 1, &lp, s));
      BSP_TRY(fetch(f + n - 1, &lf, s));
      cnt[c] = lp + lf;
    }
    BSP_TRY(lists[c].alloc(cnt[c] + 1));
    if (n > 0) {
      scatter_class<<<cdiv(n, kTB), kTB, 0, s>>>(f, pos.get(), n, lists[c].get());
      BSP_HIP(hipGetLastError());
    }
  }
  int32_t s1 = cnt[C1], s2 = cnt[C2], s3 = cnt[C3], s4 = cnt[C4];
  const int32_t nl = cnt[CLONG], nb = cnt[CBLOCK], nz = cnt[CZERO];
  // common_13 (dasp.cu:927-937)
  int32_t c13 = s1 < s3 ? s1 : s3;
  if (c13 / BS >= 16) {
    c13 = BS * (c13 / BS);
    s1 -= c13;
    s3 -= c13;
  } else {
    c13 = 0;
  }
  const int32_t sb13 = (c13 + BS - 1) / BS, half2 = (s2 + 1) / 2, sb22 = (half2 + BS - 1) / BS;
  const int32_t sb34 = (s3 + s4 + BS - 1) / BS;
  const int32_t tb13 = (sb13 + 7) / 8, tb22 = (sb22 + 7) / 8, tb34 = (sb34 + 15) / 16;
  const int32_t f13 = tb13 * 8 * MM * MK, f34 = tb34 * 16 * MM * MK, f22 = tb22 * 8 * MM * MK;
  const int32_t fshort = s1 + f13 + f34 + f22;
  v.row_long = nl;
  v.row_block = nb;
  v.row_zero = nz;
  v.short_row_1 = s1;
  v.short_row_2 = s2;
  v.short_row_3 = s3;
  v.short_row_4 = s4;
  v.common_13 = c13;
  v.fill0_nnz_short13 = f13;
  v.fill0_nnz_short34 = f34;
  v.fill0_nnz_short22 = f22;
  v.fill0_nnz_short = fshort;
  v.threadblock13 = tb13;
  v.threadblock34 = tb34;
  v.threadblock22 = tb22;
  // row-block rows: stable sort, longest first (dasp.cu:214-246, 1244)
  DevBuf<int32_t> blen_in, blen;
  BSP_TRY(blen_in.alloc(nb));
  BSP_TRY(blen.alloc(nb));
  BSP_TRY(h->ridA.alloc(nb + 1));
  if (nb > 0) {
    gather_len<<<cdiv(nb, kTB), kTB, 0, s>>>(A.rowptr, lists[CBLOCK].get(), nb, blen_in.get());
    BSP_HIP(hipGetLastError());
    BSP_TRY(sort_pairs_i32(blen_in.get(), blen.get(), lists[CBLOCK].get(), h->ridA.get(), nb, true,
                           32, s));
  }
  Lists Ls{lists[C1].get(), lists[C2].get(), lists[C3].get(), lists[C4].get(),
           lists[CLONG].get(), h->ridA.get(), lists[CZERO].get(),
           s1, s2, s3, s4, nl, nb, nz, c13};
  BSP_TRY(h->order_rid.alloc(n));
  if (n > 0) dasp_order<<<cdiv(n, kTB), kTB, 0, s>>>(Ls, n, h->order_rid.get());
  // short layout
  BSP_TRY(h->short_val.alloc(fshort));
  BSP_TRY(h->short_cid.alloc(fshort));
  BSP_HIP(hipMemsetAsync(h->short_val.get(), 0, sizeof(double) * (fshort ? fshort : 1), s));
  BSP_HIP(hipMemsetAsync(h->short_cid.get(), 0, sizeof(int32_t) * (fshort ? fshort : 1), s));
  const int64_t nshort_items = (int64_t)s1 + c13 + s3 + s4 + s2;
  if (nshort_items > 0)
    dasp_short_fill<<<cdiv(nshort_items, kTB), kTB, 0, s>>>(Ls, A.rowptr, A.colidx, aval, f13, f34,
                                                            h->short_val.get(), h->short_cid.get());
  BSP_HIP(hipGetLastError());
  // long rows (dasp.cu:1330-1369)
  DevBuf<int32_t> lcnt;
  BSP_TRY(lcnt.alloc(nl + 1));
  BSP_TRY(h->long_rpt.alloc(nl + 1));
  long_warps<<<cdiv(nl + 1, kTB), kTB, 0, s>>>(A.rowptr, lists[CLONG].get(), nl, lcnt.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(exclusive_scan_i32(lcnt.get(), h->long_rpt.get(), nl + 1, s));
  int32_t wn = 0;
  BSP_TRY(fetch(h->long_rpt.get() + nl, &wn, s));
  const int32_t bl_long = (wn + 3) / 4;
  v.fill0_nnz_long = bl_long * 4 * 64;
  v.warp_number = bl_long * 4;
  BSP_TRY(h->long_val.alloc(v.fill0_nnz_long));
  BSP_TRY(h->long_cid.alloc(v.fill0_nnz_long));
  BSP_TRY(h->rid_by_warp.alloc(v.warp_number));
  BSP_HIP(hipMemsetAsync(h->long_val.get(), 0, sizeof(double) * (v.fill0_nnz_long ? v.fill0_nnz_long : 1), s));
  BSP_HIP(hipMemsetAsync(h->long_cid.get(), 0, sizeof(int32_t) * (v.fill0_nnz_long ? v.fill0_nnz_long : 1), s));
  BSP_HIP(hipMemsetAsync(h->rid_by_warp.get(), 0, sizeof(int32_t) * (v.warp_number ? v.warp_number : 1), s));
  // long_fill and the fused SpMV are single dispatches: each must hold
  // fewer than 2^32 work-items
  if ((int64_t)nl * kWave >= kDispatchLimit)
    return fail(BSP_ERR_TOO_LARGE, "DASP: too many long rows for one dispatch");
  if (nl > 0)
    long_fill<<<cdiv((int64_t)nl * kWave, kTB), kTB, 0, s>>>(
        A.rowptr, A.colidx, aval, lists[CLONG].get(), nl, h->long_rpt.get(), h->long_val.get(),
        h->long_cid.get(), h->rid_by_warp.get());
  BSP_HIP(hipGetLastError());
  {
    DevBuf<unsigned long long> tot;
    BSP_TRY(tot.alloc(1));
    BSP_HIP(hipMemsetAsync(tot.get(), 0, sizeof(unsigned long long), s));
    if (nl > 0)
      sum_lengths<<<cdiv(nl, kTB), kTB, 0, s>>>(A.rowptr, lists[CLONG].get(), nl, tot.get());
    BSP_HIP(hipGetLastError());
    unsigned long long t = 0;
    BSP_TRY(fetch(tot.get(), &t, s));
    v.nnz_long = (int32_t)t;
  }
  // row blocks (dasp.cu:1374-1487)
  int32_t blocknum = (nb + BS - 1) / BS;
  blocknum = ((blocknum + 3) / 4) * 4;
  v.blocknum = blocknum;
  DevBuf<int32_t> bcnt, icnt;
  BSP_TRY(bcnt.alloc(blocknum + 1));
  BSP_TRY(icnt.alloc(nb + 1));
  BSP_TRY(h->block_ptr.alloc(blocknum + 1));
  BSP_TRY(h->irreg_rpt.alloc(nb + 1));
  BSP_HIP(hipMemsetAsync(icnt.get(), 0, sizeof(int32_t) * (nb + 1), s));
  block_slabs<<<cdiv(blocknum + 1, kTB), kTB, 0, s>>>(blen.get(), nb, blocknum, threshold,
                                                      bcnt.get(), icnt.get());
  BSP_HIP(hipGetLastError());
  BSP_TRY(exclusive_scan_i32(bcnt.get(), h->block_ptr.get(), blocknum + 1, s));
  BSP_TRY(exclusive_scan_i32(icnt.get(), h->irreg_rpt.get(), nb + 1, s));
  int32_t freg = 0, nirr = 0;
  BSP_TRY(fetch(h->block_ptr.get() + blocknum, &freg, s));
  BSP_TRY(fetch(h->irreg_rpt.get() + nb, &nirr, s));
  v.fill0_nnz_reg = freg;
  v.nnz_irreg = nirr;
  BSP_TRY(h->reg_val.alloc(freg));
  BSP_TRY(h->reg_cid.alloc(freg));
  BSP_TRY(h->irreg_val.alloc(nirr));
  BSP_TRY(h->irreg_cid.alloc(nirr));
  BSP_HIP(hipMemsetAsync(h->reg_val.get(), 0, sizeof(double) * (freg ? freg : 1), s));
  BSP_HIP(hipMemsetAsync(h->reg_cid.get(), 0, sizeof(int32_t) * (freg ? freg : 1), s));
  in_slices(nb, kWave, [&](int64_t f, int64_t c) {
    block_fill<<<cdiv(c * kWave, kTB), kTB, 0, s>>>(
        A.rowptr, A.colidx, aval, h->ridA.get(), nb, h->block_ptr.get(), h->irreg_rpt.get(),
        h->reg_val.get(), h->reg_cid.get(), h->irreg_val.get(), h->irreg_cid.get(), f);
  });
  BSP_HIP(hipGetLastError());
  BSP_HIP(hipStreamSynchronize(s));
  v.order_rid = h->order_rid.get();
  v.short_val = h->short_val.get();
  v.short_cid = h->short_cid.get();
  v.long_rpt = h->long_rpt.get();
  v.long_val = h->long_val.get();
  v.long_cid = h->long_cid.get();
  v.rid_by_warp = h->rid_by_warp.get();
  v.block_ptr = h->block_ptr.get();
  v.irreg_rpt = h->irreg_rpt.get();
  v.reg_val = h->reg_val.get();
  v.reg_cid = h->reg_cid.get();
  v.irreg_val = h->irreg_val.get();
  v.irreg_cid = h->irreg_cid.get();
  {
    const dasp::DaspArgs a = dasp::dasp_args(v, nullptr, nullptr, 0);
    if ((a.wg_long + a.wg_reg + a.wg_short) * kTB >= kDispatchLimit)
      return fail(BSP_ERR_TOO_LARGE, "DASP: the fused SpMV launch would exceed one dispatch");
  }
  return BSP_OK;
}

}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_dasp_build(const bsp_csr* A, double threshold, int32_t block_longest, bsp_stream s,
                          bsp_dasp* out) {
  BSP_REQUIRE(A && out, "null argument");
  BSP_REQUIRE(A->nrows >= 0 && A->ncols >= 0 && A->nnz >= 0, "negative CSR size");
  if (!sizes_fit(A->nrows, A->ncols, A->nnz))
    return fail(BSP_ERR_TOO_LARGE, "CSR sizes must fit int32 indices (at most 2^31 - 2^20)");
  BSP_REQUIRE(A->nrows == 0 || A->rowptr, "null rowptr");
  BSP_REQUIRE(A->nnz == 0 || (A->colidx && A->val), "null colidx/val");
  BSP_REQUIRE(A->dtype == BSP_F64, "FP64 values required");
  // threshold <= 0 would never end the slab loop (dasp.cu:1388)
  BSP_REQUIRE(threshold > 0.0 && threshold <= 1.0, "threshold must be in (0, 1]");
  BSP_REQUIRE(block_longest >= 5, "block_longest must be >= 5");
  auto h = std::make_unique<bsp_dasp_s>();
  BSP_TRY(build_dasp(*A, threshold, block_longest, (hipStream_t)s, h.get()));
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_dasp_view_get(bsp_dasp h, bsp_dasp_view* v) {
  BSP_REQUIRE(h && v, "null handle");
  *v = h->v;
  return BSP_OK;
}

bsp_status bsp_dasp_spmv(bsp_dasp h, const double* x, double* y, int permuted, bsp_stream s) {
  return bsp_dasp_spmv_ex(h, BSP_DASP_VALU, x, y, permuted, s);
}

bsp_status bsp_dasp_spmv_ex(bsp_dasp h, int kernel, const double* x, double* y, int permuted,
                            bsp_stream s) {
  BSP_REQUIRE(h, "null handle");
  BSP_REQUIRE(kernel == BSP_DASP_VALU || kernel == BSP_DASP_MFMA, "unknown DASP kernel");
  const bsp_dasp_view& v = h->v;
  BSP_REQUIRE((v.ncols == 0 || x) && (v.nrows == 0 || y), "null x/y");
  if (v.nrows == 0) return BSP_OK;
  const dasp::DaspArgs a = dasp::dasp_args(v, x, y, permuted);
  const int64_t grid = a.wg_long + a.wg_reg + a.wg_short;
  if (grid == 0) return BSP_OK;
  if (kernel == BSP_DASP_MFMA)
    dasp_spmv_f64_mfma<<<grid, kTB, 0, (hipStream_t)s>>>(a);
  else
    dasp_spmv_f64<<<grid, kTB, 0, (hipStream_t)s>>>(a);
  BSP_HIP(hipGetLastError());
  return BSP_OK;
}

bsp_status bsp_dasp_destroy(bsp_dasp h) {
  delete h;
  return BSP_OK;
}

}  // extern "C"

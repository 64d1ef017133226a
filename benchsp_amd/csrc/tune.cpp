// Per-matrix SpMV autotuner: build each candidate layout of one FP64 matrix,
// time its SpMV, keep the fastest, and describe the choice as a record that
// rebuilds the same layout without timing.
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "bsp_hip.hpp"

using namespace bsp;

struct bsp_tuned_s {
  bsp_csr A{};
  bsp_tune_record rec{};
  std::vector<bsp_tune_record> trials;
  bsp_spmv_plan csr = nullptr;
  bsp_bsr bsr = nullptr;
  bsp_dbsr dbsr = nullptr;
  bsp_dasp dasp = nullptr;
  bsp_dia dia = nullptr;
  ~bsp_tuned_s() {
    if (csr) bsp_plan_destroy(csr);
    if (bsr) bsp_bsr_destroy(bsr);
    if (dbsr) bsp_dbsr_destroy(dbsr);
    if (dasp) bsp_dasp_destroy(dasp);
    if (dia) bsp_dia_destroy(dia);
  }
};

namespace {

constexpr double kMinBlockFill = 0.5;  // BSR candidates: average fraction of a block filled
constexpr int kWarmups = 2;

bsp_status check_record(const bsp_tune_record& r) {
  switch (r.format) {
    case BSP_TUNE_CSR:
      BSP_REQUIRE(r.chunk >= 128 && r.chunk <= 2048, "CSR record: chunk must be in 128..2048");
      BSP_REQUIRE(r.kernel == 0 || r.kernel == 1,
                  "CSR record: kernel must be 0 (row order) or 1 (sorted blocks)");
      return BSP_OK;
    case BSP_TUNE_BSR4:
      BSP_REQUIRE(r.kernel >= BSP_BSR_VALU && r.kernel <= BSP_BSR_VALU_BALANCED,
                  "BSR4 record: kernel must be a bsp_bsr_kernel other than AUTO");
      return BSP_OK;
    case BSP_TUNE_BSR8:
      BSP_REQUIRE(r.kernel == BSP_BSR_VALU || r.kernel == BSP_BSR_MFMA,
                  "BSR8 record: kernel must be BSP_BSR_VALU or BSP_BSR_MFMA");
      return BSP_OK;
    case BSP_TUNE_DBSR:
      BSP_REQUIRE(r.kernel == BSP_DBSR_VALU || r.kernel == BSP_DBSR_MFMA,
                  "DBSR record: kernel must be BSP_DBSR_VALU or BSP_DBSR_MFMA");
      return BSP_OK;
    case BSP_TUNE_DASP:
      BSP_REQUIRE(r.kernel == BSP_DASP_VALU || r.kernel == BSP_DASP_MFMA,
                  "DASP record: kernel must be BSP_DASP_VALU or BSP_DASP_MFMA");
      BSP_REQUIRE(r.threshold > 0.0 && r.threshold <= 1.0, "DASP record: threshold in (0, 1]");
      BSP_REQUIRE(r.block_longest >= 5, "DASP record: block_longest >= 5");
      return BSP_OK;
    case BSP_TUNE_DIA:
      return BSP_OK;
    default:
      return fail(BSP_ERR_INVALID, "unknown tune format");
  }
}

// Builds the layout `r` names into `h` (which holds nothing yet).
bsp_status build(const bsp_csr& A, const bsp_tune_record& r, hipStream_t s, bsp_tuned_s* h) {
  BSP_TRY(check_record(r));
  switch (r.format) {
    case BSP_TUNE_CSR: {
      bsp_csr_options o{};
      o.seg_cols = r.seg_cols;
      o.chunk = r.chunk;
      o.seg_mode = r.kernel == 1 ? 2 : 0;
      return bsp_csr_plan_create_ex(&A, BSP_CSR_ADAPTIVE, &o, s, &h->csr);
    }
    case BSP_TUNE_BSR4:
      return bsp_bsr_build(&A, 4, 4, s, &h->bsr);
    case BSP_TUNE_BSR8:
      return bsp_bsr_build(&A, 8, 8, s, &h->bsr);
    case BSP_TUNE_DBSR:
      return bsp_dbsr_build(&A, s, &h->dbsr);
    case BSP_TUNE_DIA:
      return bsp_dia_build(&A, s, &h->dia);
    default:
      return bsp_dasp_build(&A, r.threshold, r.block_longest, s, &h->dasp);
  }
}

bsp_status run(const bsp_tuned_s* h, const double* x, double* y, hipStream_t s) {
  switch (h->rec.format) {
    case BSP_TUNE_CSR: return bsp_csr_spmv(h->csr, x, y, s);
    case BSP_TUNE_BSR4:
    case BSP_TUNE_BSR8: return bsp_bsr_spmv(h->bsr, h->rec.kernel, x, y, s);
    case BSP_TUNE_DBSR: return bsp_dbsr_spmv_ex(h->dbsr, h->rec.kernel, x, y, 0, s);
    case BSP_TUNE_DIA: return bsp_dia_spmv(h->dia, x, y, s);
    default: return bsp_dasp_spmv_ex(h->dasp, h->rec.kernel, x, y, 0, s);
  }
}

// Mean time of `reps` SpMVs after the warmups.
bsp_status time_spmv(const bsp_tuned_s* h, const double* x, double* y, int reps, hipStream_t s,
                     const EventPair& ev, double* us) {
  for (int i = 0; i < kWarmups; ++i) BSP_TRY(run(h, x, y, s));
  BSP_HIP(hipEventRecord(ev.e0, s));
  for (int i = 0; i < reps; ++i) BSP_TRY(run(h, x, y, s));
  BSP_HIP(hipEventRecord(ev.e1, s));
  BSP_HIP(hipEventSynchronize(ev.e1));
  float ms = 0.f;
  BSP_HIP(hipEventElapsedTime(&ms, ev.e0, ev.e1));
  *us = 1e3 * ms / reps;
  return BSP_OK;
}

bsp_tune_record csr_rec(int chunk, int64_t seg) {
  bsp_tune_record r{};
  r.format = BSP_TUNE_CSR;
  r.chunk = chunk;
  r.seg_cols = seg;
  return r;
}

}  // namespace

extern "C" {

bsp_status bsp_autotune(const bsp_csr* A, const bsp_tune_options* opt, bsp_stream stream,
                        bsp_tuned* out) {
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(out, "null output handle");
  BSP_REQUIRE(A->dtype == BSP_F64 && (A->nnz == 0 || A->val), "FP64 values required");
  constexpr int kAllFormats = (1 << (BSP_TUNE_DIA + 1)) - 1;
  const int formats = opt && opt->formats ? opt->formats : kAllFormats;
  const int reps = opt && opt->reps > 0 ? opt->reps : 10;
  BSP_REQUIRE((formats & ~kAllFormats) == 0, "unknown format bits");
  BSP_REQUIRE(!opt || opt->reps >= 0, "reps must be >= 0");
  hipStream_t s = (hipStream_t)stream;

  // candidate list
  std::vector<bsp_tune_record> cands;
  if (formats & (1 << BSP_TUNE_CSR)) {
    for (int c : {512, 1024, 2048}) cands.push_back(csr_rec(c, 0));
    // the unsegmented plan too, when the automatic one segments columns
    bsp_spmv_plan p = nullptr;
    BSP_TRY(bsp_csr_plan_create(A, BSP_CSR_ADAPTIVE, stream, &p));
    int64_t w = 0, k = 0;
    const bsp_status st = bsp_csr_plan_segments(p, &w, &k);
    bsp_plan_destroy(p);
    BSP_TRY(st);
    if (k > 1) {
      for (int c : {1024, 2048}) cands.push_back(csr_rec(c, -1));
      // and the sorted-block segments, where they apply (2, 4 or 8 segments)
      if (k == 2 || k == 4 || k == 8) {
        bsp_tune_record r = csr_rec(2048, 0);
        r.kernel = 1;
        cands.push_back(r);
      }
    }
  }
  // block sizes: 4x4 and 8x8 BSR, each only where its blocks are at least
  // half full on average (sparser blocks move several times the CSR bytes)
  for (const int b : {4, 8}) {
    const int fmt = b == 4 ? BSP_TUNE_BSR4 : BSP_TUNE_BSR8;
    if (!(formats & (1 << fmt)) || A->nnz == 0) continue;
    bsp_blockinfo bi{};
    BSP_TRY(bsp_block_info(A, b, stream, &bi));
    if (bi.avg < kMinBlockFill * b * b) continue;
    for (int kern : {BSP_BSR_VALU, BSP_BSR_MFMA, BSP_BSR_MFMA_BALANCED, BSP_BSR_VALU_BALANCED}) {
      if (b == 8 && kern > BSP_BSR_MFMA) continue;  // the balanced kernels take 4x4 blocks
      bsp_tune_record r{};
      r.format = fmt;
      r.kernel = kern;
      cands.push_back(r);
    }
  }
  if (formats & (1 << BSP_TUNE_DBSR))
    for (int kern : {BSP_DBSR_VALU, BSP_DBSR_MFMA}) {
      bsp_tune_record r{};
      r.format = BSP_TUNE_DBSR;
      r.kernel = kern;
      cands.push_back(r);
    }
  if (formats & (1 << BSP_TUNE_DASP))
    for (double t : {0.5, 0.75, 1.0})
      for (int bl : {64, 256}) {
        bsp_tune_record r{};
        r.format = BSP_TUNE_DASP;
        r.threshold = t;
        r.block_longest = bl;
        cands.push_back(r);
        if (t == 0.75 && bl == 256) {  // the reference's setting, also on MFMA
          r.kernel = BSP_DASP_MFMA;
          cands.push_back(r);
        }
      }

  if (formats & (1 << BSP_TUNE_DIA)) {
    bsp_tune_record r{};
    r.format = BSP_TUNE_DIA;
    cands.push_back(r);
  }

  DevBuf<double> x, y;
  BSP_TRY(x.alloc(A->ncols));
  BSP_TRY(y.alloc(A->nrows));
  BSP_HIP(hipMemsetAsync(x.get(), 0, sizeof(double) * std::max<int64_t>(1, A->ncols), s));
  EventPair ev;
  BSP_TRY(ev.create());

  std::unique_ptr<bsp_tuned_s> best;
  std::vector<bsp_tune_record> trials;
  bsp_status last = BSP_ERR_UNSUPPORTED;
  for (const bsp_tune_record& c : cands) {
    auto h = std::make_unique<bsp_tuned_s>();
    h->A = *A;
    h->rec = c;
    last = build(*A, c, s, h.get());
    if (last != BSP_OK) continue;  // e.g. unsorted columns for segments, too many diagonals
    double us = 0.0;
    BSP_TRY(time_spmv(h.get(), x.get(), y.get(), reps, s, ev, &us));
    h->rec.time_us = us;
    trials.push_back(h->rec);
    if (!best || us < best->rec.time_us) best = std::move(h);
  }
  if (!best) {
    if (last == BSP_ERR_UNSUPPORTED) return fail(last, "no autotune candidate selected");
    return last;  // the error of the last candidate, with its message
  }
  best->trials = std::move(trials);
  *out = best.release();
  return BSP_OK;
}

bsp_status bsp_tuned_from_record(const bsp_csr* A, const bsp_tune_record* rec, bsp_stream s,
                                 bsp_tuned* out) {
  BSP_TRY(check_csr_pattern(A));
  BSP_REQUIRE(rec && out, "null argument");
  BSP_REQUIRE(A->dtype == BSP_F64 && (A->nnz == 0 || A->val), "FP64 values required");
  auto h = std::make_unique<bsp_tuned_s>();
  h->A = *A;
  h->rec = *rec;
  h->rec.time_us = 0.0;
  BSP_TRY(build(*A, h->rec, (hipStream_t)s, h.get()));
  *out = h.release();
  return BSP_OK;
}

bsp_status bsp_tuned_record(bsp_tuned h, bsp_tune_record* rec) {
  BSP_REQUIRE(h && rec, "null argument");
  *rec = h->rec;
  return BSP_OK;
}

bsp_status bsp_tuned_trials(bsp_tuned h, bsp_tune_record* recs, int32_t cap, int32_t* count) {
  BSP_REQUIRE(h && count && cap >= 0 && (cap == 0 || recs), "bad arguments");
  const int32_t n = (int32_t)h->trials.size();
  for (int32_t i = 0; i < n && i < cap; ++i) recs[i] = h->trials[i];
  *count = n;
  return BSP_OK;
}

bsp_status bsp_tuned_spmv(bsp_tuned h, const double* x, double* y, bsp_stream s) {
  BSP_REQUIRE(h, "null handle");
  return run(h, x, y, (hipStream_t)s);
}

bsp_status bsp_tuned_destroy(bsp_tuned h) {
  delete h;
  return BSP_OK;
}

}  // extern "C"

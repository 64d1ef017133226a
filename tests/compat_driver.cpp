// A driver written against the reference's C++ call surface (the sequence of
// bench/spmvfp64.cpp:100-199 and bench/spmvfp16.cpp:60-75), compiled against
// include/benchsp_compat.hpp and linked with libbsp.so. It builds a 2-D
// 5-point Poisson matrix (side given on the command line), runs every entry
// point, checks each y against the CPU row loop as the reference's verify_res
// does (bench/spmvfp64.cpp:24-34), with the componentwise bound
// |y - y_ref| <= 1e-12 * (|A||x|)_i, and prints one line per entry point.
// Given an output directory as the second argument it also runs the
// header-defined builders (csrTodbsr, balanceDbsr, csrRecorderSelf,
// csrToBcsr<double,int32_t> 4x4 and csrToBcsr<half,int32_t> 16x16) on a
// matrix with ragged rows and writes their arrays there (raw binary, listed
// in manifest.txt as "name dtype count") for the test to compare with the
// oracle. Exit status 0 only if every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "benchsp_compat.hpp"

namespace {

struct Mat {
  int n = 0;
  std::vector<int32_t> rp, ci;
  std::vector<double> v;
};

Mat poisson5(int side) {
  Mat m;
  m.n = side * side;
  m.rp.push_back(0);
  for (int i = 0; i < side; ++i)
    for (int j = 0; j < side; ++j) {
      auto add = [&](int c, double a) {
        m.ci.push_back(c);
        m.v.push_back(a);
      };
      const int r = i * side + j;
      if (i > 0) add(r - side, -1.0);
      if (j > 0) add(r - 1, -1.0);
      add(r, 4.0);
      if (j + 1 < side) add(r + 1, -1.0);
      if (i + 1 < side) add(r + side, -1.0);
      m.rp.push_back((int32_t)m.ci.size());
    }
  return m;
}

int g_fail = 0;


void check(const char* name, const Mat& m, const std::vector<double>& x,
           const std::vector<double>& y, const int* perm = nullptr) {
  double worst = 0;
  for (int r = 0; r < m.n; ++r) {
    double s = 0, a = 0;
    for (int j = m.rp[r]; j < m.rp[r + 1]; ++j) {
      s += m.v[j] * x[m.ci[j]];
      a += std::fabs(m.v[j] * x[m.ci[j]]);
    }
    const double got = perm ? y[perm[r]] : y[r];
    const double e = std::fabs(got - s) / (a > 0 ? a : 1.0);
    if (!(e <= worst)) worst = e;  // NaN counts as failure
  }
  const bool ok = worst <= 1e-12;
  std::printf("%-28s max componentwise error %.3e %s\n", name, worst, ok ? "ok" : "FAIL");
  if (!ok) ++g_fail;
}

// Rows of lengths 0..40 with distinct ascending columns, values from a
// fixed sequence: a pattern where sorting rows by length matters.
Mat ragged(int n, int ncols) {
  Mat m;
  m.n = n;
  m.rp.push_back(0);
  std::vector<char> used(ncols, 0);
  for (int r = 0; r < n; ++r) {
    const int len = (r * 7 + r / 5) % 41;
    std::vector<int> cols;
    for (int k = 0; cols.size() < (size_t)len && k < 4 * len + 8; ++k) {
      const int c = (int)(((int64_t)r * 131 + (int64_t)k * k * 17 + k * 5) % ncols);
      if (!used[c]) used[c] = 1, cols.push_back(c);
    }
    for (int c : cols) used[c] = 0;
    std::sort(cols.begin(), cols.end());
    for (int c : cols) {
      m.ci.push_back(c);
      m.v.push_back(std::sin(0.1 * (double)m.ci.size()) + 0.5);
    }
    m.rp.push_back((int32_t)m.ci.size());
  }
  return m;
}

struct Dump {
  std::string dir;
  FILE* manifest = nullptr;
  template <class T>
  void put(const char* name, const char* dtype, const T* p, int64_t n) {
    FILE* f = std::fopen((dir + "/" + name + ".bin").c_str(), "wb");
    if (!f || (n > 0 && std::fwrite(p, sizeof(T), (size_t)n, f) != (size_t)n)) ++g_fail;
    if (f) std::fclose(f);
    std::fprintf(manifest, "%s %s %lld\n", name, dtype, (long long)n);
  }
};

void run_builders(const std::string& dir) {
  Dump d{dir, std::fopen((dir + "/manifest.txt").c_str(), "w")};
  if (!d.manifest) {
    ++g_fail;
    return;
  }
  Mat m = ragged(1203, 900);
  const int nnz = (int)m.ci.size();
  d.put("in_rowptr", "int32", m.rp.data(), m.n + 1);
  d.put("in_col", "int32", m.ci.data(), nnz);
  d.put("in_val", "float64", m.v.data(), nnz);
  std::fprintf(d.manifest, "ncols int 900\n");
  CSRFormat<double, int32_t> csr{m.rp.data(), m.ci.data(), m.v.data(), m.n, 900, nnz};

  dbsrMat db{};
  csrTodbsr(&csr, &db);
  balanceDbsr(&db);
  const int64_t nb = db.blocknnz / 16;
  d.put("dbsr_blcptr", "int32", db.blcPtr, db.blc_row + 1);
  d.put("dbsr_colidx", "int32", db.colIdx, nb * 16);
  d.put("dbsr_val", "float64", db.blcVal, nb * 16);
  d.put("dbsr_map", "uint16", db.blcMap, nb);
  d.put("dbsr_prefix", "int32", db.prefix, (int64_t)db.warpnum + 1);
  // SpMV over the unsorted DBSR: y comes back in the rows' own order
  std::vector<double> x(900), y(m.n);
  for (int i = 0; i < 900; ++i) x[i] = std::cos(0.3 * i);
  double t = 0;
  if (amgT_spmv_fp64_dbsr(&db, x.data(), y.data(), 1, 2, t)) ++g_fail;
  check("amgT_spmv_fp64_dbsr(csrTodbsr)", m, x, y);
  bsp_dbsrMat_free(reinterpret_cast<bsp_dbsrMat*>(&db));

  CSRFormat<double, int32_t> rec{};
  csrRecorderSelf(&csr, &rec);
  d.put("rec_rowptr", "int32", rec.rowPtr, m.n + 1);
  d.put("rec_col", "int32", rec.colIdx, nnz);
  d.put("rec_val", "float64", rec.values, nnz);
  bsp_free(rec.rowPtr);
  bsp_free(rec.colIdx);
  bsp_free(rec.values);

  BCSRFormat<double, int32_t> b4{};
  b4.TILE_M = 4;
  b4.TILE_N = 4;
  b4.TILE_K = 4;
  csrToBcsr(&csr, &b4);
  d.put("bcsr4_rowptr", "int32", b4.bcsrRowPtr, b4.rowRegions + 1);
  d.put("bcsr4_col", "int32", b4.bcsrcolIdx, b4.nonzeroBlocks);
  d.put("bcsr4_val", "float64", b4.values, b4.elemCount);
  d.put("bcsr4_map", "int32", b4.relativeBlockIndexMapping, b4.blockNum);
  std::free(b4.bcsrRowPtr);
  std::free(b4.bcsrcolIdx);
  std::free(b4.values);
  std::free(b4.relativeBlockIndexMapping);

  // half values, as SMaT calls it (src/smat.cu): every value here is
  // rounded to half first, so the block values must equal those halves
  std::vector<half> vh(nnz);
  for (int j = 0; j < nnz; ++j) vh[j] = __float2half((float)m.v[j]);
  CSRFormat<half, int32_t> csrh{m.rp.data(), m.ci.data(), vh.data(), m.n, 900, nnz};
  BCSRFormat<half, int32_t> b16{};
  b16.TILE_M = 16;
  b16.TILE_N = 16;
  b16.TILE_K = 16;
  csrToBcsr(&csrh, &b16);
  d.put("bcsr16_rowptr", "int32", b16.bcsrRowPtr, b16.rowRegions + 1);
  d.put("bcsr16_col", "int32", b16.bcsrcolIdx, b16.nonzeroBlocks);
  d.put("bcsr16_val", "float16", b16.values, b16.elemCount);
  std::free(b16.bcsrRowPtr);
  std::free(b16.bcsrcolIdx);
  std::free(b16.values);
  std::free(b16.relativeBlockIndexMapping);
  std::fclose(d.manifest);
}

}  // namespace

int main(int argc, char** argv) {
  const int side = argc > 1 ? std::atoi(argv[1]) : 64;
  Mat m = poisson5(side);
  const int n = m.n, nnz = (int)m.ci.size();
  std::vector<double> x(n), y(n);
  for (int i = 0; i < n; ++i) x[i] = std::sin(0.37 * i) + 0.25;
  std::vector<int64_t> rp64(m.rp.begin(), m.rp.end()), ci64(m.ci.begin(), m.ci.end());
  double t = 0;

  if (spmv_fp64_warpper(m.rp.data(), m.ci.data(), n, n, nnz, m.v.data(), x.data(), y.data(), 5, t))
    ++g_fail;
  check("spmv_fp64_warpper", m, x, y);
  if (spmv_fp64_balance_warpper(m.rp.data(), m.ci.data(), n, n, nnz, m.v.data(), x.data(),
                                y.data(), 5, t))
    ++g_fail;
  check("spmv_fp64_balance_warpper", m, x, y);
  if (amgT_spmv_fp64(m.rp.data(), m.ci.data(), m.v.data(), x.data(), y.data(), n, n, nnz, 5, t))
    ++g_fail;
  check("amgT_spmv_fp64", m, x, y);
  bcsr_spmv_fp64(rp64.data(), ci64.data(), m.v.data(), x.data(), y.data(), n, n, nnz, 2);
  check("bcsr_spmv_fp64", m, x, y);
  if (cusparse_spmv_fp64(rp64.data(), ci64.data(), m.v.data(), x.data(), y.data(), n, n, nnz, 5, t))
    ++g_fail;
  check("cusparse_spmv_fp64", m, x, y);
  MySpmvWithPrefp64(rp64.data(), ci64.data(), n, n, nnz, m.v.data(), x.data(), y.data());
  check("MySpmvWithPrefp64", m, x, y);
  ref_spmv_fp64(rp64.data(), ci64.data(), m.v.data(), x.data(), y.data(), n, n, t);
  check("ref_spmv_fp64", m, x, y);

  // DBSR: y comes back in sorted-row order (bench/spmvfp64.cpp:150-166)
  CSRFormat<double, int32_t> csr{m.rp.data(), m.ci.data(), m.v.data(), n, n, nnz};
  dbsrMat d{};
  csrTodbsrWithRecorder(&csr, &d);
  std::vector<double> ys(n);
  if (amgT_spmv_fp64_dbsr(&d, x.data(), ys.data(), 2, 5, t)) ++g_fail;
  std::vector<int> pos(n);
  for (int i = 0; i < n; ++i) pos[d.recorderMap[i]] = i;
  check("amgT_spmv_fp64_dbsr", m, x, ys, pos.data());
  if (amgT_spmv_fp64_dbsr_balance(&d, x.data(), ys.data(), 2, 5, t)) ++g_fail;
  check("amgT_spmv_fp64_dbsr_balance", m, x, ys, pos.data());
  bsp_dbsrMat_free(reinterpret_cast<bsp_dbsrMat*>(&d));

  // DASP: Y in order_rid order (bench/spmvfp64.cpp:170-182)
  std::vector<int> order(n);
  spmv_all(nullptr, m.v.data(), m.rp.data(), m.ci.data(), x.data(), ys.data(), order.data(), n, n,
           nnz, 1, 0.75, 256, t);
  for (int i = 0; i < n; ++i) pos[order[i]] = i;
  check("spmv_all", m, x, ys, pos.data());

  // FP16 surface: every input here is exact in binary16 and every row sum
  // is an integer or a short dyadic, so the FP16 results are exact too
  std::vector<float> v32(m.v.begin(), m.v.end()), x32(n), y32(n), yr(n);
  for (int i = 0; i < n; ++i) x32[i] = (float)((i % 7) - 3) * 0.5f;
  MySpmvWithPrefp16(m.rp.data(), m.ci.data(), n, n, nnz, v32.data(), x32.data(), y32.data());
  ref_spmv_fp16(m.rp.data(), m.ci.data(), v32.data(), x32.data(), yr.data(), n, n);
  int bad16 = 0;
  for (int i = 0; i < n; ++i) bad16 += y32[i] != yr[i];
  std::printf("%-28s %d rows differ from ref_spmv_fp16 %s\n", "MySpmvWithPrefp16", bad16,
              bad16 ? "FAIL" : "ok");
  g_fail += bad16 != 0;

  if (argc > 2) run_builders(argv[2]);

  std::printf("%s\n", g_fail ? "FAILED" : "ALL OK");
  return g_fail ? 1 : 0;
}

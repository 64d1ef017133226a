// A driver written against the reference's C++ call surface (the sequence of
// bench/spmvfp64.cpp:100-199 and bench/spmvfp16.cpp:60-75), compiled against
// include/benchsp_compat.hpp and linked with libbsp.so. It builds a 2-D
// 5-point Poisson matrix (side given on the command line), runs every entry
// point, checks each y against the CPU row loop as the reference's verify_res
// does (bench/spmvfp64.cpp:24-34), with the componentwise bound
// |y - y_ref| <= 1e-12 * (|A||x|)_i, and prints one line per entry point.
// Exit status 0 only if every check passes.
#include <cmath>
#include <cstdio>
#include <cstdlib>
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

  std::printf("%s\n", g_fail ? "FAILED" : "ALL OK");
  return g_fail ? 1 : 0;
}

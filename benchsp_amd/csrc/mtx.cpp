// MatrixMarket reader and COO -> CSR assembly (host side of the benchmark
// drivers, bench/spmvfp64.cpp:36-93, which read files through
// fast_matrix_market). Coordinate format only; real, integer and pattern
// fields; general, symmetric and skew-symmetric storage.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "bsp_internal.hpp"

using namespace bsp;

namespace {

std::string lower(std::string s) {
  for (auto& c : s) c = (char)std::tolower((unsigned char)c);
  return s;
}

// Rows sorted by (row, column); entries with equal (row, column) keep their
// input order and are all kept, like the reference driver's triplets.
bsp_status assemble(int64_t nrows, int64_t ncols, std::vector<int32_t>& r, std::vector<int32_t>& c,
                    std::vector<double>& v, HostCsr** out) {
  const size_t nnz = r.size();
  std::vector<int64_t> order(nnz);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
    return r[a] != r[b] ? r[a] < r[b] : c[a] < c[b];
  });
  auto h = std::make_unique<HostCsr>();
  h->nrows = nrows;
  h->ncols = ncols;
  h->rowptr.assign(nrows + 1, 0);
  h->col.resize(nnz);
  h->val.resize(nnz);
  for (size_t k = 0; k < nnz; ++k) {
    const int64_t i = order[k];
    h->rowptr[r[i] + 1] += 1;
    h->col[k] = c[i];
    h->val[k] = v[i];
  }
  for (int64_t i = 0; i < nrows; ++i) h->rowptr[i + 1] += h->rowptr[i];
  *out = h.release();
  return BSP_OK;
}

}  // namespace

extern "C" {

static bsp_status coo_to_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* rows,
                             const int32_t* cols, const double* vals, bsp_host_csr* out) {
  BSP_REQUIRE(out && nrows >= 0 && ncols >= 0 && nnz >= 0, "bad COO arguments");
  BSP_REQUIRE(nnz == 0 || (rows && cols), "null COO arrays");
  if (!sizes_fit(nrows, ncols, nnz))
    return fail(BSP_ERR_TOO_LARGE, "COO sizes must fit int32 indices (at most 2^31 - 2^20)");
  std::vector<int32_t> r(rows, rows + nnz), c(cols, cols + nnz);
  std::vector<double> v(nnz, 1.0);
  if (vals) std::copy(vals, vals + nnz, v.begin());
  for (int64_t k = 0; k < nnz; ++k)
    if (r[k] < 0 || r[k] >= nrows || c[k] < 0 || c[k] >= ncols)
      return fail(BSP_ERR_INVALID, "COO index out of range at entry " + std::to_string(k));
  HostCsr* h = nullptr;
  const bsp_status st = assemble(nrows, ncols, r, c, v, &h);
  if (st != BSP_OK) return st;
  *out = (bsp_host_csr)h;
  return BSP_OK;
}

static bsp_status mtx_read(const char* path, bsp_host_csr* out) {
  BSP_REQUIRE(path && out, "null argument");
  std::unique_ptr<FILE, int (*)(FILE*)> f(std::fopen(path, "rb"), std::fclose);
  if (!f) return fail(BSP_ERR_INVALID, std::string("cannot open ") + path + ": " + std::strerror(errno));
  std::fseek(f.get(), 0, SEEK_END);
  const long size = std::ftell(f.get());
  std::fseek(f.get(), 0, SEEK_SET);
  if (size <= 0) return fail(BSP_ERR_INVALID, "empty MatrixMarket file");
  std::string buf((size_t)size, '\0');
  if (std::fread(&buf[0], 1, (size_t)size, f.get()) != (size_t)size)
    return fail(BSP_ERR_INVALID, "short read");

  // banner: %%MatrixMarket matrix coordinate <field> <symmetry>
  size_t eol = buf.find('\n');
  std::string banner = lower(buf.substr(0, eol));
  char obj[32] = {0}, fmt[32] = {0}, field[32] = {0}, sym[32] = {0};
  if (std::sscanf(banner.c_str(), "%%%%matrixmarket %31s %31s %31s %31s", obj, fmt, field, sym) != 4 ||
      std::strcmp(obj, "matrix") != 0)
    return fail(BSP_ERR_INVALID, "not a MatrixMarket matrix banner");
  if (std::strcmp(fmt, "coordinate") != 0)
    return fail(BSP_ERR_UNSUPPORTED, "only the coordinate format is supported");
  const bool pattern = !std::strcmp(field, "pattern");
  if (!pattern && std::strcmp(field, "real") && std::strcmp(field, "integer") &&
      std::strcmp(field, "double"))
    return fail(BSP_ERR_UNSUPPORTED, std::string("unsupported field ") + field);
  const bool symm = !std::strcmp(sym, "symmetric"), skew = !std::strcmp(sym, "skew-symmetric");
  if (!symm && !skew && std::strcmp(sym, "general"))
    return fail(BSP_ERR_UNSUPPORTED, std::string("unsupported symmetry ") + sym);

  // skip comments, read the size line
  size_t pos = eol == std::string::npos ? buf.size() : eol + 1;
  while (pos < buf.size() && (buf[pos] == '%' || buf[pos] == '\n' || buf[pos] == '\r')) {
    size_t e = buf.find('\n', pos);
    pos = e == std::string::npos ? buf.size() : e + 1;
  }
  const char* p = buf.c_str() + pos;
  char* end = nullptr;
  const long long M = std::strtoll(p, &end, 10);
  const long long N = std::strtoll(end, &end, 10);
  const long long L = std::strtoll(end, &end, 10);
  if (M < 0 || N < 0 || L < 0 || end == p) return fail(BSP_ERR_INVALID, "bad size line");
  // symmetric and skew-symmetric storage mirror (i, j) to (j, i): square only
  if ((symm || skew) && M != N)
    return fail(BSP_ERR_INVALID, "symmetric storage needs a square matrix");
  if (!sizes_fit(M, N, symm || skew ? 2 * L : L))
    return fail(BSP_ERR_TOO_LARGE, "MatrixMarket sizes must fit int32 indices (at most 2^31 - 2^20)");
  // every entry takes at least four bytes of text ("1 1\n"): a header that
  // claims more entries than the file can hold is malformed (and must not
  // size the arrays below)
  if (L > (long long)(buf.size() - (size_t)(end - buf.c_str())) / 4 + 1)
    return fail(BSP_ERR_INVALID, "entry count larger than the file can hold");
  std::vector<int32_t> r, c;
  std::vector<double> v;
  const size_t cap = (size_t)(symm || skew ? 2 * L : L);
  r.reserve(cap);
  c.reserve(cap);
  v.reserve(cap);
  p = end;
  for (long long k = 0; k < L; ++k) {
    const long long i = std::strtoll(p, &end, 10);
    if (end == p) return fail(BSP_ERR_INVALID, "truncated entry list at entry " + std::to_string(k));
    p = end;
    const long long j = std::strtoll(p, &end, 10);
    if (end == p) return fail(BSP_ERR_INVALID, "truncated entry at " + std::to_string(k));
    p = end;
    double a = 1.0;
    if (!pattern) {
      a = std::strtod(p, &end);
      if (end == p) return fail(BSP_ERR_INVALID, "missing value at entry " + std::to_string(k));
      p = end;
    }
    if (i < 1 || i > M || j < 1 || j > N)
      return fail(BSP_ERR_INVALID, "index out of range at entry " + std::to_string(k));
    r.push_back((int32_t)(i - 1));
    c.push_back((int32_t)(j - 1));
    v.push_back(a);
    if ((symm || skew) && i != j) {
      r.push_back((int32_t)(j - 1));
      c.push_back((int32_t)(i - 1));
      v.push_back(skew ? -a : a);
    }
  }
  HostCsr* h = nullptr;
  const bsp_status st = assemble(M, N, r, c, v, &h);
  if (st != BSP_OK) return st;
  *out = (bsp_host_csr)h;
  return BSP_OK;
}

// Host arrays sized by the caller's (or the file's) counts: an allocation
// that fails is a status, not an exception through the C interface.
bsp_status bsp_coo_to_csr(int64_t nrows, int64_t ncols, int64_t nnz, const int32_t* rows,
                          const int32_t* cols, const double* vals, bsp_host_csr* out) {
  try {
    return coo_to_csr(nrows, ncols, nnz, rows, cols, vals, out);
  } catch (const std::bad_alloc&) {
    return fail(BSP_ERR_ALLOC, "host allocation failed");
  }
}

bsp_status bsp_mtx_read(const char* path, bsp_host_csr* out) {
  try {
    return mtx_read(path, out);
  } catch (const std::bad_alloc&) {
    return fail(BSP_ERR_ALLOC, "host allocation failed");
  }
}

}  // extern "C"

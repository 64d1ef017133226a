// Deterministic synthetic-matrix generators (host C++).
//
// The reference reads SuiteSparse .mtx files (bench/spmvfp64.cpp:36-93); the
// workloads this library is measured on are synthetic, so the generators are
// part of the library and pinned here. Every random quantity comes from
// splitmix64 (Steele, Lea & Flood 2014): element i of a stream seeded with s
// is mix64(s + (i+1)*0x9E3779B97F4A7C15). Uniform [0,1) is (r>>11)*2^-53,
// uniform [-1,1) is 2u-1, normals are Box-Muller cos branch with
// u1 = ((r1>>11)+1)*2^-53 in (0,1]. Row r of a matrix draws from its own
// stream seeded with mix64(seed ^ (0xA0761D6478BD642F*(r+1))), so rows can be
// generated in any order or in parallel with identical results.
#include "bsp_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

namespace bsp {

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() { s += 0x9E3779B97F4A7C15ull; return mix64(s); }
  double u01() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uni() { return 2.0 * u01() - 1.0; }
  double normal() {
    double u1 = (double)((next() >> 11) + 1) * 0x1.0p-53;
    double u2 = u01();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // floor(r * n / 2^64): Lemire's multiply-shift range reduction.
  uint64_t below(uint64_t n) { return (uint64_t)(((unsigned __int128)next() * n) >> 64); }
};

static inline uint64_t row_seed(uint64_t seed, uint64_t r) {
  return mix64(seed ^ (0xA0761D6478BD642Full * (r + 1)));
}

template <class F>
static void parallel_rows(int64_t n, F f) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  int64_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    int64_t a = t * chunk, b = std::min<int64_t>(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([=] { for (int64_t r = a; r < b; ++r) f(r); });
  }
  for (auto& t : th) t.join();
}

static void scan_lengths(HostCsr* h, const std::vector<int64_t>& len) {
  h->rowptr.assign(h->nrows + 1, 0);
  for (int64_t i = 0; i < h->nrows; ++i) h->rowptr[i + 1] = h->rowptr[i] + len[i];
  h->col.resize(h->rowptr[h->nrows]);
  h->val.resize(h->rowptr[h->nrows]);
}

// 2-D 5-point Poisson stencil, nx*ny grid, row = i*ny + j, diag 4, off -1.
HostCsr* gen_stencil5(int64_t nx, int64_t ny) {
  auto* h = new HostCsr;
  h->nrows = h->ncols = nx * ny;
  std::vector<int64_t> len(h->nrows);
  for (int64_t i = 0; i < nx; ++i)
    for (int64_t j = 0; j < ny; ++j)
      len[i * ny + j] = 1 + (i > 0) + (i < nx - 1) + (j > 0) + (j < ny - 1);
  scan_lengths(h, len);
  parallel_rows(h->nrows, [&](int64_t r) {
    int64_t i = r / ny, j = r % ny, p = h->rowptr[r];
    auto put = [&](int64_t c, double v) { h->col[p] = (int32_t)c; h->val[p] = v; ++p; };
    if (i > 0) put(r - ny, -1.0);
    if (j > 0) put(r - 1, -1.0);
    put(r, 4.0);
    if (j < ny - 1) put(r + 1, -1.0);
    if (i < nx - 1) put(r + ny, -1.0);
  });
  return h;
}

// 3-D stencil on nx*ny*nz, row = (i*ny + j)*nz + k, columns ascending.
// points == 7: Laplacian, diag 6, off -1.
// points == 27: off-diagonals U[-1,1) drawn in column order from the row
// stream, diagonal = 1 + sum |off-diagonal| (strictly diagonally dominant).
HostCsr* gen_stencil3d(int64_t nx, int64_t ny, int64_t nz, int points, uint64_t seed) {
  if (points != 7 && points != 27) return nullptr;
  auto* h = new HostCsr;
  h->nrows = h->ncols = nx * ny * nz;
  auto inside = [&](int64_t i, int64_t j, int64_t k) {
    return i >= 0 && i < nx && j >= 0 && j < ny && k >= 0 && k < nz;
  };
  auto nb = [&](int di, int dj, int dk) {
    int nzc = (di != 0) + (dj != 0) + (dk != 0);
    return points == 27 || nzc <= 1;
  };
  std::vector<int64_t> len(h->nrows);
  parallel_rows(h->nrows, [&](int64_t r) {
    int64_t i = r / (ny * nz), j = (r / nz) % ny, k = r % nz, c = 0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk)
          if (nb(di, dj, dk) && inside(i + di, j + dj, k + dk)) ++c;
    len[r] = c;
  });
  scan_lengths(h, len);
  parallel_rows(h->nrows, [&](int64_t r) {
    int64_t i = r / (ny * nz), j = (r / nz) % ny, k = r % nz, p = h->rowptr[r], pd = -1;
    Rng g(row_seed(seed, r));
    double offsum = 0.0;
    for (int di = -1; di <= 1; ++di)
      for (int dj = -1; dj <= 1; ++dj)
        for (int dk = -1; dk <= 1; ++dk) {
          if (!nb(di, dj, dk) || !inside(i + di, j + dj, k + dk)) continue;
          h->col[p] = (int32_t)(r + (di * ny + dj) * nz + dk);
          if (di == 0 && dj == 0 && dk == 0) {
            pd = p;
          } else if (points == 7) {
            h->val[p] = -1.0;
          } else {
            double v = g.uni();
            h->val[p] = v;
            offsum += std::fabs(v);
          }
          ++p;
        }
    h->val[pd] = points == 7 ? 6.0 : 1.0 + offsum;
  });
  return h;
}

// x rounded and clamped to [lo, hi] in floating point before the integer
// conversion (a NaN or infinite x gives lo or hi, never an undefined cast).
static int64_t clamp_count(double x, int64_t lo, int64_t hi) {
  if (!(x >= (double)lo)) return lo;  // also NaN
  if (x >= (double)hi) return hi;
  return (int64_t)x;
}

// Random matrix with lognormal row lengths:
//   L = min(cap, max(1, round(exp(mu + sigma*z)))), mu = ln(mean) - sigma^2/2,
// z the first normal of the row stream; L distinct columns uniform in
// [0, ncols) (sorted ascending); values U[-1,1) drawn after the columns.
HostCsr* gen_powerlaw(int64_t n, int64_t ncols, double mean, double sigma, int64_t cap, uint64_t seed) {
  auto* h = new HostCsr;
  h->nrows = n;
  h->ncols = ncols;
  double mu = std::log(mean) - 0.5 * sigma * sigma;
  cap = std::min<int64_t>(cap, ncols);
  std::vector<int64_t> len(n);
  parallel_rows(n, [&](int64_t r) {
    Rng g(row_seed(seed, r));
    len[r] = clamp_count(std::nearbyint(std::exp(mu + sigma * g.normal())), 1, cap);
  });
  scan_lengths(h, len);
  parallel_rows(n, [&](int64_t r) {
    Rng g(row_seed(seed, r));
    g.normal();
    int64_t L = len[r];
    std::vector<int64_t> cols;
    cols.reserve(L);
    while ((int64_t)cols.size() < L) {
      int64_t need = L - (int64_t)cols.size();
      for (int64_t t = 0; t < need; ++t) cols.push_back((int64_t)g.below((uint64_t)ncols));
      std::sort(cols.begin(), cols.end());
      cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
    }
    int64_t p = h->rowptr[r];
    for (int64_t t = 0; t < L; ++t) {
      h->col[p + t] = (int32_t)cols[t];
      h->val[p + t] = g.uni();
    }
  });
  return h;
}

// Block-sparse matrix: nbr x nbc grid of bs x bs dense blocks, exactly
// nblocks distinct positions drawn without replacement (Floyd's algorithm on
// the sequential stream `seed`), sorted row-major. Block b (in sorted order)
// takes bs*bs N(0,1) values, row-major, from row stream b. Returned as CSR
// of the expanded matrix (columns ascending within a row).
HostCsr* gen_blocksparse(int64_t nbr, int64_t nbc, int bs, int64_t nblocks, uint64_t seed,
                         std::vector<int64_t>* bpos_out) {
  int64_t total = nbr * nbc;
  if (nblocks > total || nblocks < 0) return nullptr;
  std::vector<uint8_t> taken(total, 0);
  std::vector<int64_t> pos;
  pos.reserve(nblocks);
  Rng g(seed);
  for (int64_t j = total - nblocks; j < total; ++j) {
    int64_t t = (int64_t)g.below((uint64_t)(j + 1));
    if (taken[t]) t = j;
    taken[t] = 1;
    pos.push_back(t);
  }
  std::sort(pos.begin(), pos.end());
  auto* h = new HostCsr;
  h->nrows = nbr * bs;
  h->ncols = nbc * bs;
  std::vector<int64_t> bptr(nbr + 1, 0);
  for (int64_t p : pos) bptr[p / nbc + 1]++;
  for (int64_t i = 0; i < nbr; ++i) bptr[i + 1] += bptr[i];
  std::vector<int64_t> len(h->nrows);
  for (int64_t r = 0; r < h->nrows; ++r) len[r] = (bptr[r / bs + 1] - bptr[r / bs]) * bs;
  scan_lengths(h, len);
  parallel_rows(nbr, [&](int64_t br) {
    for (int64_t b = bptr[br]; b < bptr[br + 1]; ++b) {
      Rng gb(row_seed(seed, b));
      int64_t bc = pos[b] % nbc, slot = b - bptr[br];
      for (int i = 0; i < bs; ++i)
        for (int j = 0; j < bs; ++j) {
          int64_t r = br * bs + i;
          int64_t p = h->rowptr[r] + slot * bs + j;
          h->col[p] = (int32_t)(bc * bs + j);
          h->val[p] = gb.normal();
        }
    }
  });
  if (bpos_out) *bpos_out = pos;
  return h;
}

// Power-law graph (Chung-Lu style). Node weights w_i = c*(i+1)^(-1/(gamma-1)),
// c chosen so sum(w)/n = avg_deg. Node i emits k_i = round(w_i/2) (at least 1,
// at most n/2) edges whose targets are drawn proportionally to w from the row
// stream (inverse CDF). Edges are symmetrised (unless `symmetric` is 0),
// self-loops and duplicates removed. Values: symmetric normalisation
// D^-1/2 A D^-1/2 with D the row degrees of the binary pattern. When
// symmetric == 0 the pattern is not symmetrised and values are U[-1,1).
HostCsr* gen_graph(int64_t n, double avg_deg, double gamma, int symmetric, uint64_t seed) {
  std::vector<double> w(n), cdf(n + 1, 0.0);
  double e = -1.0 / (gamma - 1.0), s = 0.0;
  for (int64_t i = 0; i < n; ++i) { w[i] = std::pow((double)(i + 1), e); s += w[i]; }
  double c = avg_deg * (double)n / s;
  for (int64_t i = 0; i < n; ++i) { w[i] *= c; cdf[i + 1] = cdf[i] + w[i]; }
  double tot = cdf[n];
  std::vector<int64_t> k(n), off(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    k[i] = clamp_count(std::nearbyint(w[i] * 0.5), 1, std::max<int64_t>(1, n / 2));
    off[i + 1] = off[i] + k[i];
  }
  int64_t m = off[n];
  std::vector<uint64_t> edges((symmetric ? 2 : 1) * m);
  parallel_rows(n, [&](int64_t i) {
    Rng g(row_seed(seed, i));
    for (int64_t t = 0; t < k[i]; ++t) {
      double u = g.u01() * tot;
      int64_t j = (int64_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin()) - 1;
      j = std::min<int64_t>(std::max<int64_t>(j, 0), n - 1);
      edges[off[i] + t] = ((uint64_t)i << 32) | (uint64_t)j;
      if (symmetric) edges[m + off[i] + t] = ((uint64_t)j << 32) | (uint64_t)i;
    }
  });
  std::sort(edges.begin(), edges.end());
  edges.erase(std::unique(edges.begin(), edges.end()), edges.end());
  edges.erase(std::remove_if(edges.begin(), edges.end(),
                             [](uint64_t x) { return (x >> 32) == (x & 0xffffffffull); }),
              edges.end());
  auto* h = new HostCsr;
  h->nrows = h->ncols = n;
  std::vector<int64_t> len(n, 0);
  for (uint64_t x : edges) len[x >> 32]++;
  scan_lengths(h, len);
  for (size_t p = 0; p < edges.size(); ++p) h->col[p] = (int32_t)(edges[p] & 0xffffffffull);
  if (symmetric) {
    parallel_rows(n, [&](int64_t r) {
      for (int64_t p = h->rowptr[r]; p < h->rowptr[r + 1]; ++p) {
        int64_t cc = h->col[p];
        h->val[p] = 1.0 / std::sqrt((double)len[r] * (double)len[cc]);
      }
    });
  } else {
    parallel_rows(n, [&](int64_t r) {
      Rng g(row_seed(seed ^ 0x5bd1e995ull, r));
      for (int64_t p = h->rowptr[r]; p < h->rowptr[r + 1]; ++p) h->val[p] = g.uni();
    });
  }
  return h;
}

// Generators allocate host arrays sized by their arguments: an allocation
// that fails is a status, not an exception through the C interface.
template <class F>
static bsp_status gen_call(bsp_host_csr* out, F&& make) {
  try {
    HostCsr* h = make();
    if (!h) return BSP_ERR_INVALID;
    *out = (bsp_host_csr)h;
    return BSP_OK;
  } catch (const std::bad_alloc&) {
    return fail(BSP_ERR_ALLOC, "host allocation failed");
  }
}

// every dimension below 2^31 before any product of them is formed
static bool dims_ok(int64_t a, int64_t b, int64_t c = 1) {
  const int64_t lim = int64_t(1) << 31;
  return a > 0 && b > 0 && c > 0 && a < lim && b < lim && c < lim;
}

}  // namespace bsp

using namespace bsp;

extern "C" {

bsp_status bsp_gen_uniform(int64_t n, uint64_t seed, double* out) {
  if (n < 0 || (n > 0 && !out)) return BSP_ERR_INVALID;
  for (int64_t i = 0; i < n; ++i) {
    uint64_t r = mix64(seed + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ull);
    out[i] = 2.0 * ((double)(r >> 11) * 0x1.0p-53) - 1.0;
  }
  return BSP_OK;
}

bsp_status bsp_gen_normal(int64_t n, uint64_t seed, double* out) {
  if (n < 0 || (n > 0 && !out)) return BSP_ERR_INVALID;
  Rng g(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = g.normal();
  return BSP_OK;
}

bsp_status bsp_gen_stencil5(int64_t nx, int64_t ny, bsp_host_csr* out) {
  if (!out || !dims_ok(nx, ny) || nx * ny >= (1ll << 31)) return BSP_ERR_INVALID;
  return gen_call(out, [&] { return gen_stencil5(nx, ny); });
}

bsp_status bsp_gen_stencil3d(int64_t nx, int64_t ny, int64_t nz, int points, uint64_t seed,
                             bsp_host_csr* out) {
  if (!out || !dims_ok(nx, ny, nz) || nx * ny >= (1ll << 31) || nx * ny * nz >= (1ll << 31))
    return BSP_ERR_INVALID;
  return gen_call(out, [&] { return gen_stencil3d(nx, ny, nz, points, seed); });
}

bsp_status bsp_gen_powerlaw(int64_t n, int64_t ncols, double mean, double sigma, int64_t cap,
                            uint64_t seed, bsp_host_csr* out) {
  if (!out || !dims_ok(n, ncols) || !(mean > 0) || cap <= 0 || !std::isfinite(sigma) ||
      !std::isfinite(mean))
    return BSP_ERR_INVALID;
  return gen_call(out, [&] { return gen_powerlaw(n, ncols, mean, sigma, cap, seed); });
}

bsp_status bsp_gen_blocksparse(int64_t nbr, int64_t nbc, int bs, int64_t nblocks, uint64_t seed,
                               bsp_host_csr* out) {
  if (!out || !dims_ok(nbr, nbc, bs) || nbc * bs >= (1ll << 31) || nbr * bs >= (1ll << 31))
    return BSP_ERR_INVALID;
  return gen_call(out, [&] { return gen_blocksparse(nbr, nbc, bs, nblocks, seed, nullptr); });
}

bsp_status bsp_gen_graph(int64_t n, double avg_deg, double gamma, int symmetric, uint64_t seed,
                         bsp_host_csr* out) {
  if (!out || n <= 1 || n >= (1ll << 31) || !(avg_deg > 0) || !std::isfinite(avg_deg) ||
      !(gamma > 1.0) || !std::isfinite(gamma))
    return BSP_ERR_INVALID;
  return gen_call(out, [&] { return gen_graph(n, avg_deg, gamma, symmetric, seed); });
}

bsp_status bsp_host_csr_info(bsp_host_csr h, int64_t* nrows, int64_t* ncols, int64_t* nnz) {
  if (!h) return BSP_ERR_INVALID;
  auto* c = (HostCsr*)h;
  if (nrows) *nrows = c->nrows;
  if (ncols) *ncols = c->ncols;
  if (nnz) *nnz = c->rowptr.empty() ? 0 : c->rowptr[c->nrows];
  return BSP_OK;
}

bsp_status bsp_host_csr_copy(bsp_host_csr h, int64_t* rowptr, int32_t* col, double* val) {
  if (!h) return BSP_ERR_INVALID;
  auto* c = (HostCsr*)h;
  if (rowptr) std::memcpy(rowptr, c->rowptr.data(), sizeof(int64_t) * (c->nrows + 1));
  if (col) std::memcpy(col, c->col.data(), sizeof(int32_t) * c->col.size());
  if (val) std::memcpy(val, c->val.data(), sizeof(double) * c->val.size());
  return BSP_OK;
}

void bsp_host_csr_free(bsp_host_csr h) { delete (HostCsr*)h; }

}  // extern "C"

"""Host-side synthetic matrices and vectors (the library's deterministic
generators, benchsp_amd/csrc/gen.cpp), returned as numpy CSR."""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._lib import i64, lib


@dataclass
class Csr:
    nrows: int
    ncols: int
    rowptr: np.ndarray  # int64, nrows+1
    col: np.ndarray     # int32
    val: np.ndarray     # float64

    @property
    def nnz(self) -> int:
        return int(self.rowptr[-1]) if self.nrows else 0

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rowptr)

    def to_dense(self) -> np.ndarray:
        d = np.zeros((self.nrows, self.ncols))
        rows = np.repeat(np.arange(self.nrows), self.row_lengths())
        d[rows, self.col] = self.val
        return d

    def csr_bytes(self) -> int:
        """Algorithmic bytes of one FP64 CSR SpMV with int32 indices:
        12*nnz + 4*(n+1) + 8*ncols + 8*n."""
        return 12 * self.nnz + 4 * (self.nrows + 1) + 8 * self.ncols + 8 * self.nrows


def _take(h: C.c_void_p) -> Csr:
    L = lib()
    nr, nc, nz = i64(), i64(), i64()
    L.bsp_host_csr_info(h, C.byref(nr), C.byref(nc), C.byref(nz))
    rp = np.empty(nr.value + 1, np.int64)
    col = np.empty(max(nz.value, 1), np.int32)
    val = np.empty(max(nz.value, 1), np.float64)
    L.bsp_host_csr_copy(h, rp.ctypes.data, col.ctypes.data, val.ctypes.data)
    L.bsp_host_csr_free(h)
    return Csr(nr.value, nc.value, rp, col[: nz.value], val[: nz.value])


def _gen(fn, *args) -> Csr:
    h = C.c_void_p()
    fn(*args, C.byref(h))
    return _take(h)


def stencil5(nx: int, ny: int) -> Csr:
    """2-D 5-point Poisson stencil, diag 4, off-diagonals -1."""
    return _gen(lib().bsp_gen_stencil5, nx, ny)


def stencil3d(nx: int, ny: int, nz: int, points: int, seed: int = 0) -> Csr:
    """3-D 7-point Laplacian or 27-point stencil with U[-1,1) off-diagonals and
    a diagonal of 1 + sum |off-diagonal|."""
    return _gen(lib().bsp_gen_stencil3d, nx, ny, nz, points, seed)


def powerlaw(n: int, mean: float = 20.0, sigma: float = 1.2, cap: int = 5000, seed: int = 0,
             ncols: int | None = None) -> Csr:
    """Lognormal row lengths (mean `mean`), distinct uniform columns, U[-1,1) values."""
    return _gen(lib().bsp_gen_powerlaw, n, ncols or n, mean, sigma, cap, seed)


def blocksparse(nbr: int, nbc: int, bs: int, nblocks: int, seed: int = 0) -> Csr:
    """nblocks distinct dense bs x bs blocks with N(0,1) values."""
    return _gen(lib().bsp_gen_blocksparse, nbr, nbc, bs, nblocks, seed)


def graph(n: int, avg_deg: float = 50.0, gamma: float = 2.1, symmetric: bool = True,
          seed: int = 0) -> Csr:
    """Power-law graph; symmetric: D^-1/2 A D^-1/2 values."""
    return _gen(lib().bsp_gen_graph, n, avg_deg, gamma, int(symmetric), seed)


def read_mtx(path: str) -> Csr:
    """MatrixMarket coordinate file -> CSR sorted by (row, column)
    (bsp_mtx_read)."""
    return _gen(lib().bsp_mtx_read, path.encode())


def coo_to_csr(nrows: int, ncols: int, rows, cols, vals=None) -> Csr:
    """0-based COO -> CSR sorted by (row, column), duplicates kept in order."""
    r = np.ascontiguousarray(rows, np.int32)
    c = np.ascontiguousarray(cols, np.int32)
    v = None if vals is None else np.ascontiguousarray(vals, np.float64)
    return _gen(lib().bsp_coo_to_csr, nrows, ncols, r.size, r.ctypes.data, c.ctypes.data,
                None if v is None else v.ctypes.data)


def uniform(n: int, seed: int = 0) -> np.ndarray:
    out = np.empty(max(n, 1), np.float64)
    lib().bsp_gen_uniform(n, seed, out.ctypes.data)
    return out[:n]


def normal(n: int, seed: int = 0) -> np.ndarray:
    out = np.empty(max(n, 1), np.float64)
    lib().bsp_gen_normal(n, seed, out.ctypes.data)
    return out[:n]


# Named workloads -------------------------------------------------------------

def poisson5_256() -> Csr:
    """5-point Poisson on a 256x256 grid: n=65,536, nnz=326,656."""
    return stencil5(256, 256)


def stencil27_64(seed: int = 1) -> Csr:
    """27-point stencil on a 64^3 grid: n=262,144, nnz=6,859,000."""
    return stencil3d(64, 64, 64, 27, seed)


def powerlaw_1m(seed: int = 2) -> Csr:
    """n=2^20, lognormal row lengths (mean 20, sigma 1.2, cap 5000): nnz ~ 20M."""
    return powerlaw(1 << 20, 20.0, 1.2, 5000, seed)


def blocksparse_16k(seed: int = 3) -> Csr:
    """16384 x 16384, 16x16 dense blocks at 5% block density (52,429 blocks)."""
    return blocksparse(1024, 1024, 16, int(round(0.05 * 1024 * 1024)), seed)


def zipf_graph_200k(symmetric: bool = True, seed: int = 4) -> Csr:
    """n=200,000, Zipf exponent 2.1, average degree ~50 after symmetrisation
    and de-duplication (sampling target 68.5: nnz = 9.98M for seed 4)."""
    return graph(200_000, 68.5, 2.1, symmetric, seed)


def batch_sizes(count: int = 512, lo: int = 1000, hi: int = 65536, seed: int = 5) -> list[int]:
    """n = round(exp(U[ln lo, ln hi))) per matrix, from the splitmix64 stream."""
    u = (uniform(count, seed) + 1.0) * 0.5
    return [int(round(math.exp(math.log(lo) + ui * (math.log(hi) - math.log(lo))))) for ui in u]


def batch_mix(count: int = 512, lo: int = 1000, hi: int = 65536, seed: int = 5) -> list[Csr]:
    """The batched workload: sizes from batch_sizes, kinds by index mod 3 --
    2-D 5-point stencil of side round(sqrt(n)), 3-D 7-point stencil of side
    round(cbrt(n)), random with lognormal row lengths (mean 12, sigma 1.2, cap
    min(n, 5000)) -- each matrix seeded by its index."""
    out = []
    for i, n in enumerate(batch_sizes(count, lo, hi, seed)):
        k = i % 3
        if k == 0:
            s = max(2, int(round(math.sqrt(n))))
            out.append(stencil5(s, s))
        elif k == 1:
            s = max(2, int(round(n ** (1.0 / 3.0))))
            out.append(stencil3d(s, s, s, 7, seed + i))
        else:
            out.append(powerlaw(n, 12.0, 1.2, min(n, 5000), seed + i))
    return out

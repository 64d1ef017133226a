"""GPU format builders bit-exact against the oracle's CPU builders, and the
blocked SpMV kernels against the oracle's SpMV (through the C-ABI)."""
import glob
import os

import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rand_csr(rng, lens, ncols):
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) if L else
                          np.zeros(0, np.int64) for L in lens]).astype(np.int32)
    return Csr(len(lens), ncols, rp, col, rng.uniform(-1, 1, int(rp[-1])))


def _mats(bsp_host):
    rng = np.random.default_rng(3)
    yield "random", _rand_csr(rng, rng.integers(0, 40, 301), 263)
    yield "stencil27-64", bsp_host.stencil27_64()
    yield "powerlaw-100k", bsp_host.powerlaw(100_000, 20.0, 1.2, 2000, seed=4)
    yield "empty-rows", _rand_csr(rng, [0] * 7 + [3] + [0] * 9, 50)
    yield from _degenerate(rng)


def _degenerate(rng):
    # no entries at all; one row; one column; rows not a multiple of any block
    yield "all-empty", Csr(37, 29, np.zeros(38, np.int64), np.zeros(0, np.int32), np.zeros(0))
    yield "one-row", _rand_csr(rng, [50], 100)
    yield "one-col", _rand_csr(rng, rng.integers(0, 2, 203), 1)
    yield "ragged-5x7", _rand_csr(rng, [3, 0, 7, 1, 2], 7)


def _same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape and np.array_equal(a, b), what


@pytest.mark.parametrize("tm,tk", [(4, 4), (3, 3), (8, 8), (16, 16), (2, 5)])
def test_bsr_builder_bit_exact(oracle, bsp_host, tm, tk):
    import torch
    from benchsp_amd.device import Bsr, DeviceCsr
    for name, A in _mats(bsp_host):
        if tm == 16 and A.nrows > 200_000:
            continue
        print(f"[bsr] {name} {tm}x{tk}", flush=True)
        g = Bsr(DeviceCsr.from_host(A), tm, tk).arrays()
        o = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, tm, tk)
        assert g["nblocks"] == o["nblocks"], name
        for k in ("rowptr", "colidx", "val"):
            _same(g[k], o[k], f"{name} {k}")
        if tm == 4 and tk == 4:
            for k in ("map", "rowptr_by_warp", "rowidx_by_warp"):
                _same(g[k], o[k], f"{name} {k}")
            assert g["warpnum"] == o["warpnum"]
            assert g["avg_nnz"] == o["avg_nnz"]
            assert abs(g["stand"] - o["stand"]) <= 1e-12 * max(1.0, o["stand"])
        torch.cuda.synchronize()


_KERNELS = {"valu": 1, "mfma": 2, "mfma_bal": 3, "valu_bal": 4}


@pytest.mark.parametrize("kernel", ["valu", "mfma", "mfma_bal", "valu_bal"])
@pytest.mark.parametrize("tile", [4, 3, 8, 16])
def test_bsr_spmv(oracle, bsp_host, kernel, tile):
    import torch
    from benchsp_amd.device import Bsr, DeviceCsr
    if (kernel == "mfma" and tile not in (4, 8)) or (kernel.endswith("_bal") and tile != 4):
        # no such kernel for this block size (MFMA: 4x4 and 8x8; balanced:
        # 4x4): the call is refused with BSP_ERR_UNSUPPORTED before any launch
        from benchsp_amd._lib import BspError
        A = bsp_host.blocksparse(30, 30, tile, 100, seed=1)
        B = Bsr(DeviceCsr.from_host(A), tile, tile)
        with pytest.raises(BspError) as e:
            B(torch.zeros(A.ncols, dtype=torch.float64, device="cuda"), kernel=_KERNELS[kernel])
        assert e.value.status == 5
        return
    mats = list(_mats(bsp_host))
    if tile in (4, 8):  # dense blocks of the kernel's size, odd block-row lengths
        mats.append((f"blocks-{tile}", bsp_host.blocksparse(300, 257, tile, 25_000, seed=tile)))
    for name, A in mats:
        if tile == 16 and A.nrows > 200_000:
            continue
        print(f"[bsr-spmv] {name} {tile} {kernel}", flush=True)
        x = bsp_host.uniform(A.ncols, 21)
        B = Bsr(DeviceCsr.from_host(A), tile, tile)
        y = B(torch.from_numpy(x).cuda(), kernel=_KERNELS[kernel])
        torch.cuda.synchronize()
        assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                          abs_spmv(oracle, A, x), what=f"{name} bsr{tile} {kernel}")
        if kernel.endswith("_bal"):  # deterministic: a second run gives the same bits
            y2 = B(torch.from_numpy(x).cuda(), kernel=_KERNELS[kernel])
            assert torch.equal(y, y2), name


@pytest.mark.parametrize("kernel,tile", [("mfma", 4), ("mfma_bal", 4), ("valu_bal", 4),
                                         ("mfma", 8)])
def test_bsr_x_not_16_byte_aligned(oracle, bsp_host, kernel, tile):
    # the MFMA and balanced bitmap kernels load x in 16-byte pieces when x is
    # 16-byte aligned and as 8-byte loads otherwise: x as a view one double
    # into a buffer (ragged last block column: 263 columns), both paths
    # against the oracle
    import torch
    from benchsp_amd.device import Bsr, DeviceCsr
    rng = np.random.default_rng(5)
    A = _rand_csr(rng, rng.integers(0, 40, 301), 263)
    x = bsp_host.uniform(A.ncols, 22)
    B = Bsr(DeviceCsr.from_host(A), tile, tile)
    buf = torch.zeros(A.ncols + 2, dtype=torch.float64, device="cuda")
    ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    for off in (0, 1):
        xd = buf[off:off + A.ncols]
        xd.copy_(torch.from_numpy(x))
        assert (xd.data_ptr() % 16 == 0) == (off == 0)
        y = B(xd, kernel=_KERNELS[kernel])
        torch.cuda.synchronize()
        assert_spmv_close(y.cpu().numpy(), ref, abs_spmv(oracle, A, x),
                          what=f"bsr{tile} {kernel} x offset {off}")


def test_bsr4_builder_bit_exact_powerlaw_1m(oracle, bsp_host):
    # the 4x4 BSR of the power-law 2^20 matrix (about one entry per block)
    from benchsp_amd.device import Bsr, DeviceCsr
    A = bsp_host.powerlaw_1m()
    g = Bsr(DeviceCsr.from_host(A), 4, 4).arrays()
    o = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 4, 4)
    assert g["nblocks"] == o["nblocks"]
    for k in ("rowptr", "colidx", "val", "map", "rowptr_by_warp", "rowidx_by_warp"):
        _same(g[k], o[k], f"powerlaw-1M {k}")
    assert g["warpnum"] == o["warpnum"] and g["avg_nnz"] == o["avg_nnz"]


@pytest.mark.parametrize("kernel", ["valu", "mfma", "mfma_bal", "valu_bal"])
def test_bsr4_spmv_powerlaw_1m(oracle, bsp_host, kernel):
    import torch
    from benchsp_amd.device import Bsr, DeviceCsr
    A = bsp_host.powerlaw_1m()
    x = bsp_host.uniform(A.ncols, 22)
    B = Bsr(DeviceCsr.from_host(A), 4, 4)
    y = B(torch.from_numpy(x).cuda(), kernel=_KERNELS[kernel])
    torch.cuda.synchronize()
    assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                      abs_spmv(oracle, A, x), what=f"powerlaw-1M bsr4 {kernel}")


def _dbsr_inputs(bsp_host):
    for p in sorted(glob.glob(os.path.join(GOLDEN, "dbsr_*.npz"))):
        g = np.load(p)
        yield os.path.basename(p), Csr(int(g["nrows"]), int(g["ncols"]),
                                       g["rowptr"].astype(np.int64), g["col"], g["val"])
    rng = np.random.default_rng(0)
    yield "overflow-case", _rand_csr(rng, [0, 1, 2, 3, 4, 5, 9, 16, 17, 33, 70, 0, 2, 64, 65, 8, 1], 90)
    yield from _degenerate(rng)
    # ~4,600 block rows (5 chunks of the parallel warp packing) whose block
    # counts sit around the 64-block warp capacity, below it and far above it
    kind = rng.choice(3, 18_500, p=[0.6, 0.3, 0.1])
    lens = np.where(kind == 0, rng.integers(0, 61, kind.size),
                    np.where(kind == 1, rng.integers(100, 261, kind.size),
                             rng.integers(256, 1101, kind.size)))
    yield "warp-capacity-mix", _rand_csr(rng, lens, 1500)
    yield "stencil27-64", bsp_host.stencil27_64()
    yield "powerlaw-1M", bsp_host.powerlaw_1m()


def test_dbsr_builder_bit_exact(oracle, bsp_host):
    from benchsp_amd.device import DeviceCsr, Dbsr
    for name, A in _dbsr_inputs(bsp_host):
        print(f"[dbsr] {name}", flush=True)
        g = Dbsr(DeviceCsr.from_host(A)).arrays()
        o = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
        assert g["warpnum"] == o["warpnum"] and g["overflow"] == o["overflow"], name
        for k in ("recorder", "blcptr", "colidx", "val", "map", "prefix"):
            _same(g[k], o[k], f"{name} {k}")


@pytest.mark.parametrize("kernel", ["valu", "mfma"])
def test_dbsr_spmv(oracle, bsp_host, kernel):
    import torch
    from benchsp_amd.device import DeviceCsr, Dbsr
    kern = Dbsr.MFMA if kernel == "mfma" else Dbsr.VALU
    for name, A in _dbsr_inputs(bsp_host):
        x = bsp_host.uniform(A.ncols, 5)
        D = Dbsr(DeviceCsr.from_host(A))
        xd = torch.from_numpy(x).cuda()
        y = D(xd, kernel=kern).cpu().numpy()
        ys = D(xd, sorted_order=True, kernel=kern).cpu().numpy()
        y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
        bound = abs_spmv(oracle, A, x)
        assert_spmv_close(y, y_ref, bound, what=f"{name} dbsr")
        rec = D.arrays()["recorder"]
        un = np.empty_like(ys)
        un[rec] = ys  # the caller's un-permute step (bench/spmvfp64.cpp:163-166)
        assert np.array_equal(un, y)


def _dasp_inputs(bsp_host):
    rng = np.random.default_rng(5)
    lens = np.concatenate([np.full(300, 1), np.full(200, 3), np.full(77, 2), np.full(50, 4),
                           np.full(9, 0), rng.integers(5, 300, 400), [256, 255, 1000, 5000]])
    rng.shuffle(lens)
    yield "mixed", _rand_csr(rng, lens, 6000)
    yield "stencil27-64", bsp_host.stencil27_64()
    yield "powerlaw-1M", bsp_host.powerlaw_1m()
    yield "poisson5-256", bsp_host.poisson5_256()
    yield from _degenerate(rng)


DASP_KEYS = ["order_rid", "short_val", "short_cid", "long_rpt", "long_val", "long_cid",
             "rid_by_warp", "block_ptr", "irreg_rpt", "reg_val", "reg_cid", "irreg_val",
             "irreg_cid"]


def test_dasp_builder_bit_exact(oracle, bsp_host):
    from benchsp_amd.device import DASP_INTS, Dasp, DeviceCsr
    for name, A in _dasp_inputs(bsp_host):
        print(f"[dasp] {name}", flush=True)
        g = Dasp(DeviceCsr.from_host(A)).arrays()
        o = oracle.dasp(A.rowptr, A.col, A.val, 0.75, 256)
        for k in DASP_INTS:
            assert g[k] == o[k], f"{name} {k}: {g[k]} vs {o[k]}"
        for k in DASP_KEYS:
            _same(g[k], o[k], f"{name} {k}")


@pytest.mark.parametrize("kernel", ["valu", "mfma"])
@pytest.mark.parametrize("threshold,longest", [(0.75, 256), (0.5, 64), (1.0, 1000)])
def test_dasp_spmv(oracle, bsp_host, threshold, longest, kernel):
    import torch
    from benchsp_amd.device import Dasp, DeviceCsr
    kern = Dasp.MFMA if kernel == "mfma" else Dasp.VALU
    for name, A in _dasp_inputs(bsp_host):
        x = bsp_host.uniform(A.ncols, 8)
        D = Dasp(DeviceCsr.from_host(A), threshold, longest)
        xd = torch.from_numpy(x).cuda()
        y = D(xd, kernel=kern).cpu().numpy()
        yp = D(xd, permuted=True, kernel=kern).cpu().numpy()
        y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
        assert_spmv_close(y, y_ref, abs_spmv(oracle, A, x), what=f"{name} dasp")
        order = D.arrays()["order_rid"]
        assert np.array_equal(yp, y[order])  # Y_val in order_rid order


@pytest.mark.parametrize("b", [3, 4, 16])
def test_block_info(oracle, bsp_host, b):
    # bench/blockinfo.cpp's two block counts, on the GPU, against the oracle
    from benchsp_amd.device import DeviceCsr, block_info
    for name, A in _mats(bsp_host):
        got = block_info(DeviceCsr.from_host(A), b)
        blocks, packed = oracle.blockinfo(A.nrows, A.ncols, A.rowptr, A.col, b)
        assert (got["nnz"], got["blocks"], got["blocks_packed"]) == (A.nnz, blocks, packed), name
        if blocks:
            assert got["avg"] == A.nnz / blocks and got["avg_packed"] == A.nnz / packed

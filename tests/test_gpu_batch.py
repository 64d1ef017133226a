"""Batched FP64 SpMV + 8-column FP16 SpMM over many matrices (bsp_batch_*):
persistent schedule (fixed, autotuned, forced BSR, forced DASP, forced
diagonal layout), plain schedule and graph replays, each matrix against the
oracle."""
import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu
COLS = 8


def _mix(bsp_host):
    rng = np.random.default_rng(8)
    mats = bsp_host.batch_mix(count=30, lo=200, hi=6000, seed=11)
    # a matrix with one very long row and empty rows
    lens = np.concatenate([[0, 3, 2000, 0], rng.integers(0, 6, 300), [700, 1]])
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(2500, L, replace=False)) for L in lens]).astype(np.int32)
    mats.append(Csr(len(lens), 2500, rp, col, rng.uniform(-1, 1, int(rp[-1]))))
    # dense 8x8 and 4x4 blocks (eligible for the BSR candidates), ragged last
    # block row and column
    for bs, nbr, nbc, nb, cut in ((8, 30, 29, 300, (3, 5)), (4, 40, 37, 600, (2, 3))):
        B = bsp_host.blocksparse(nbr, nbc, bs, nb, seed=bs)
        keep = B.nrows - cut[0], B.ncols - cut[1]
        d = B.to_dense()[: keep[0], : keep[1]]
        r, c = np.nonzero(d)
        rp = np.zeros(keep[0] + 1, np.int64)
        np.add.at(rp, r + 1, 1)
        mats.append(Csr(keep[0], keep[1], np.cumsum(rp), c.astype(np.int32), d[r, c]))
    mats.append(Csr(0, 5, np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0)))
    # 8 diagonals with holes, wider than tall (the diagonal layout's masks and
    # clamped columns; eligible for it)
    mats.append(_banded_holes(rng, 700, 900, [-300, -2, -1, 0, 1, 5, 64, 899]))
    return mats


def _banded_holes(rng, nrows, ncols, offsets, keep=0.7):
    rows, cols = [], []
    for r in range(nrows):
        for d in offsets:
            if 0 <= r + d < ncols and rng.random() < keep:
                rows.append(r)
                cols.append(r + d)
    rp = np.zeros(nrows + 1, np.int64)
    np.add.at(rp, np.array(rows, np.int64) + 1, 1)
    return Csr(nrows, ncols, np.cumsum(rp), np.array(cols, np.int32),
               rng.uniform(-1, 1, len(cols)))


def _vectors(bsp_host, mats, seed):
    xs = [bsp_host.uniform(A.ncols, seed + i) for i, A in enumerate(mats)]
    Xs = [bsp_host.normal(A.ncols * COLS, seed + 1000 + i).reshape(A.ncols, COLS).astype(np.float16)
          for i, A in enumerate(mats)]
    return xs, Xs


def _check(oracle, mats, xs, Xs, y, Y, B):
    ro = B.row_off
    for i, A in enumerate(mats):
        yi = y[ro[i]:ro[i + 1]]
        assert_spmv_close(yi, oracle.spmv_f64(A.rowptr, A.col, A.val, xs[i]),
                          abs_spmv(oracle, A, xs[i]), what=f"spmv matrix {i}")
        # FP16 A and X, FP32 accumulation: vs FP64 on the rounded operands
        a16 = A.val.astype(np.float16).astype(np.float64)
        X64 = Xs[i].astype(np.float64)
        ref = oracle.csr_spmm_f64(A.rowptr, A.col, a16, X64)
        bound = oracle.csr_spmm_f64(A.rowptr, A.col, np.abs(a16), np.abs(X64))
        Yi = Y[ro[i]:ro[i + 1]]
        assert np.all(np.abs(Yi - ref) <= 1e-5 * bound + 1e-30), f"spmm matrix {i}"


def _device(mats, xs, Xs):
    import torch
    from benchsp_amd.device import DeviceCsr
    dm = [DeviceCsr.from_host(A) for A in mats]
    x = torch.from_numpy(np.concatenate(xs)).cuda()
    X = torch.from_numpy(np.concatenate(Xs).reshape(-1, COLS)).cuda()
    return dm, x, X


@pytest.mark.parametrize("tune", [0, 1, 2, 3, 4])
def test_batch_persistent(oracle, bsp_host, tune):
    import torch
    from benchsp_amd.device import Batch
    mats = _mix(bsp_host)
    xs, Xs = _vectors(bsp_host, mats, 21)
    dm, x, X = _device(mats, xs, Xs)
    B = Batch(dm, tune=tune)
    ch = B.choice()
    assert set(np.unique(ch["spmv_window"])) <= {128, 256, 512, 1024}
    if tune == 2:  # the dense-block matrices run as 8x8 and 4x4 BSR
        assert ch["format"][-4] == 2 and ch["format"][-3] == 1
        assert np.all(ch["format"][:-4] == 0) and ch["format"][-1] == 0
    if tune == 3:  # every matrix with rows as DASP tiles
        assert np.all(ch["format"][:-2] == 3) and ch["format"][-1] == 3
    if tune == 0:
        assert np.all(ch["format"] == 0) and np.all(ch["spmv_window"] == 1024)
    if tune == 4:  # the stencils (index mod 3 = 0, 1) and the banded matrix (last) in the
        # diagonal layout, for both operations
        dia = np.array([(i < 30 and i % 3 < 2) or i == len(mats) - 1 for i in range(len(mats))])
        assert np.all((ch["format"] == Batch.DIA) == dia), ch["format"]
        assert np.all((ch["spmm_format"] == Batch.DIA) == dia), ch["spmm_format"]
    if tune != 4:
        assert np.all((ch["spmm_format"] == Batch.CSR) | (ch["spmm_format"] == Batch.DIA))
    y = torch.full((B.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    Y = torch.full((B.nrows, COLS), float("nan"), dtype=torch.float32, device="cuda")
    B.run(x, y, X, Y)
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    _check(oracle, mats, xs, Xs, yh, Y.cpu().numpy(), B)
    for i in np.nonzero(ch["format"] == Batch.DIA)[0]:  # the CPU loop's bits
        A = mats[i]
        assert np.array_equal(yh[B.row_off[i]:B.row_off[i + 1]],
                              oracle.spmv_f64(A.rowptr, A.col, A.val, xs[i])), i


def test_batch_plain_schedule_matches(oracle, bsp_host):
    import torch
    from benchsp_amd.device import PLAIN, Batch
    mats = _mix(bsp_host)
    xs, Xs = _vectors(bsp_host, mats, 31)
    dm, x, X = _device(mats, xs, Xs)
    B = Batch(dm, tune=0)
    y = torch.full((B.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    Y = torch.full((B.nrows, COLS), float("nan"), dtype=torch.float32, device="cuda")
    B.run(x, y, X, Y, schedule=PLAIN)
    torch.cuda.synchronize()
    yh = y.cpu().numpy()
    _check(oracle, mats, xs, Xs, yh, Y.cpu().numpy(), B)
    for i, A in enumerate(mats):  # thread per row, left to right: the CPU loop
        assert np.array_equal(yh[B.row_off[i]:B.row_off[i + 1]],
                              oracle.spmv_f64(A.rowptr, A.col, A.val, xs[i]))


@pytest.mark.parametrize("tune", [1, 2, 3, 4])
def test_batch_graph_replays(oracle, bsp_host, tune):
    # the graph reads x and X at replay time: change them between replays,
    # with eager work and synchronisation in between, and check every replay
    # (tune 2, 3 and 4 put BSR, DASP and diagonal tiles in the captured launch)
    import torch
    from benchsp_amd.device import Batch
    mats = _mix(bsp_host)
    xs, Xs = _vectors(bsp_host, mats, 41)
    dm, x, X = _device(mats, xs, Xs)
    B = Batch(dm, tune=tune)
    y = torch.zeros(B.nrows, dtype=torch.float64, device="cuda")
    Y = torch.zeros(B.nrows, COLS, dtype=torch.float32, device="cuda")
    B.capture(x, y, X, Y)
    for rep in range(3):
        xs, Xs = _vectors(bsp_host, mats, 50 + 7 * rep)
        x.copy_(torch.from_numpy(np.concatenate(xs)))
        X.copy_(torch.from_numpy(np.concatenate(Xs).reshape(-1, COLS)))
        y.fill_(float("nan"))
        Y.fill_(float("nan"))
        torch.cuda.synchronize()
        B.replay()
        torch.cuda.synchronize()
        _check(oracle, mats, xs, Xs, y.cpu().numpy(), Y.cpu().numpy(), B)
        _ = (x * 2).sum().item()  # eager work between replays


def test_batch_rejects_bad_input(bsp_host):
    import torch
    from benchsp_amd._lib import BspError
    from benchsp_amd.device import Batch, DeviceCsr
    A = bsp_host.stencil5(8, 8)
    d = DeviceCsr.from_host(A, dtype=torch.float32)
    with pytest.raises(BspError):
        Batch([d])


def test_batch_grid_sizes(oracle, bsp_host):
    # the launch size changes only which workgroup runs a tile, never a
    # tile's arithmetic: every grid gives the same bits
    import torch
    from benchsp_amd._lib import BspError
    from benchsp_amd.device import Batch
    mats = _mix(bsp_host)
    xs, Xs = _vectors(bsp_host, mats, 61)
    dm, x, X = _device(mats, xs, Xs)
    outs = []
    for g in (0, 1, 3, 64, 1 << 30):
        B = Batch(dm, tune=0, grid=g)
        info = B.info()
        tiles = max(info["spmv_tiles"], info["spmm_tiles"])
        assert info["grid"] == (min(g, tiles) if g else info["grid"]) and info["grid"] >= 1
        y = torch.full((B.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        Y = torch.full((B.nrows, COLS), float("nan"), dtype=torch.float32, device="cuda")
        B.run(x, y, X, Y)
        torch.cuda.synchronize()
        outs.append((y.cpu().numpy(), Y.cpu().numpy()))
    _check(oracle, mats, xs, Xs, outs[0][0], outs[0][1], B)
    for y, Y in outs[1:]:
        assert np.array_equal(y, outs[0][0]) and np.array_equal(Y, outs[0][1])
    with pytest.raises(BspError):
        Batch(dm, tune=0, grid=-1)

"""GPU parity of the CSR SpMV kernels against the CPU oracle (src/ref.cpp:26-37
restated in oracle/oracle.c), through the C-ABI."""
import numpy as np
import pytest

from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu

METHODS = ["adaptive", "scalar", "balanced", "split"]


def _run(A, x, method):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    dA = DeviceCsr.from_host(A)
    op = CsrSpmv(dA, method)
    xd = torch.from_numpy(x).cuda()
    y = op(xd)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def _check(oracle, A, x, method, what):
    print(f"[spmv] {what} method={method} n={A.nrows} nnz={A.nnz}", flush=True)
    y = _run(A, x, method)
    y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    assert_spmv_close(y, y_ref, abs_spmv(oracle, A, x), what=f"{what}/{method}")
    return y, y_ref


@pytest.mark.parametrize("method", METHODS)
def test_poisson5_256(oracle, bsp_host, method):
    A = bsp_host.poisson5_256()
    x = bsp_host.uniform(A.ncols, seed=0)
    y, y_ref = _check(oracle, A, x, method, "5pt-256")
    if method == "scalar":
        # same products, same left-to-right order, no contraction: bit-exact
        assert np.array_equal(y, y_ref)


@pytest.mark.parametrize("method", METHODS)
def test_poisson5_ones_closed_form(bsp_host, method):
    A = bsp_host.poisson5_256()
    y = _run(A, np.ones(A.ncols), method)
    n = 256
    expect = np.zeros((n, n))
    expect[0, :] += 1
    expect[-1, :] += 1
    expect[:, 0] += 1
    expect[:, -1] += 1
    assert np.array_equal(y, expect.ravel())


@pytest.mark.parametrize("method", METHODS)
def test_stencil27_64(oracle, bsp_host, method):
    A = bsp_host.stencil27_64()
    x = bsp_host.uniform(A.ncols, seed=11)
    _check(oracle, A, x, method, "27pt-64")


@pytest.mark.parametrize("method", METHODS)
def test_powerlaw_1m(oracle, bsp_host, method):
    A = bsp_host.powerlaw_1m()
    x = bsp_host.uniform(A.ncols, seed=12)
    _check(oracle, A, x, method, "powerlaw-1M")


@pytest.mark.parametrize("method", METHODS)
def test_edge_shapes(oracle, bsp_host, method):
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(7)
    # empty rows, a single huge row (> one workgroup chunk), rectangular shape
    lens = np.array([0, 3, 0, 0, 5000, 1, 129, 128, 0, 2047, 2048, 2049, 17] + [0] * 40 + [1] * 3000)
    n, ncols = lens.size, 9000
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) for L in lens]).astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1])
    A = Csr(n, ncols, rp, col, val)
    x = rng.uniform(-1, 1, ncols)
    _check(oracle, A, x, method, "edge")


@pytest.mark.parametrize("method", METHODS)
def test_all_empty_rows(method):
    from benchsp_amd.host import Csr
    A = Csr(1000, 10, np.zeros(1001, np.int64), np.zeros(0, np.int32), np.zeros(0))
    y = _run(A, np.ones(10), method)
    assert np.array_equal(y, np.zeros(1000))


@pytest.mark.parametrize("method", METHODS)
def test_degenerate_and_unsorted(oracle, method):
    # a 1x1 matrix, one column, and rows whose columns are unsorted and
    # repeated (duplicates are summed in order, as the CPU loop does); none of
    # these are column-segmented, so every method accepts them
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(12)
    cases = [Csr(1, 1, np.array([0, 1]), np.array([0], np.int32), np.array([2.5])),
             Csr(700, 1, np.arange(701, dtype=np.int64), np.zeros(700, np.int32),
                 rng.uniform(-1, 1, 700))]
    lens = rng.integers(0, 40, 500)
    rp = np.zeros(501, np.int64)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, 60, int(rp[-1])).astype(np.int32)  # unsorted, with repeats
    cases.append(Csr(500, 60, rp, col, rng.uniform(-1, 1, int(rp[-1]))))
    for A in cases:
        x = rng.uniform(-1, 1, A.ncols)
        _check(oracle, A, x, method, f"degenerate {A.nrows}x{A.ncols}")


@pytest.mark.parametrize("method", METHODS)
def test_zero_rows(method):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    A = Csr(0, 5, np.zeros(1, np.int64), np.zeros(0, np.int32), np.zeros(0))
    op = CsrSpmv(DeviceCsr.from_host(A), method)
    y = op(torch.ones(5, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    assert y.numel() == 0


def _run_seg(A, x, seg_cols, seg_mode=0):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=seg_cols, seg_mode=seg_mode)
    y = op(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    return y.cpu().numpy(), op.segments()


def test_adaptive_segments_powerlaw_auto(oracle, bsp_host):
    A = bsp_host.powerlaw_1m()
    x = bsp_host.uniform(A.ncols, seed=13)
    y, (w, k) = _run_seg(A, x, 0)
    assert w == 1 << 18 and k == 4  # scattered columns: 4 XCD-pinned segments of 2 MB of x
    assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                      what="powerlaw segmented")


def test_adaptive_stencil_not_segmented(bsp_host):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    A = bsp_host.stencil3d(80, 80, 80, 7)  # 512,000 columns, all near the diagonal
    assert CsrSpmv(DeviceCsr.from_host(A), "adaptive").segments() == (0, 1)


@pytest.mark.parametrize("seg_cols,seg_mode", [(1000, 0), (4097, 0), (4500, 0), (2250, 0),
                                               (1125, 0), (4500, 1), (2250, 1)])
def test_adaptive_forced_segments(oracle, bsp_host, seg_cols, seg_mode):
    rng = np.random.default_rng(9)
    from benchsp_amd.host import Csr
    lens = np.array([0, 3, 0, 5000, 1, 129, 2047, 2049, 17] + [0] * 40 + [1] * 3000 + [40] * 500)
    n, ncols = lens.size, 9000
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) for L in lens]).astype(np.int32)
    A = Csr(n, ncols, rp, col, rng.uniform(-1, 1, rp[-1]))
    x = rng.uniform(-1, 1, ncols)
    y, (w, k) = _run_seg(A, x, seg_cols, seg_mode)
    assert w == seg_cols and k == -(-ncols // seg_cols)
    assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                      what=f"forced segments {seg_cols}")
    S = bsp_host.stencil3d(20, 21, 22, 27, seed=5)
    xs = bsp_host.uniform(S.ncols, 3)
    y, _ = _run_seg(S, xs, seg_cols, seg_mode)
    assert_spmv_close(y, oracle.spmv_f64(S.rowptr, S.col, S.val, xs), abs_spmv(oracle, S, xs),
                      what="stencil forced segments")


@pytest.mark.parametrize("seg_cols", [20_000, 10_000, 5_000])
def test_sorted_blocks(oracle, seg_cols):
    # seg_mode 2: 2, 4 or 8 XCD-pinned segments whose blocks store their
    # entries in column order; rows of every kind: empty, short, exactly 32
    # and 33 entries (the one-lane and the wave sums), a row long enough to
    # take a block of its own in every segment, thousands of short rows
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(23)
    lens = np.array([0, 3, 0, 15_000, 1, 129, 2047, 33, 32, 31, 17] + [0] * 40 + [1] * 3000 +
                    [40] * 500 + [5] * 20_000 + [32, 33, 64, 65])
    n, ncols = lens.size, 40_000
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) for L in lens]).astype(np.int32)
    A = Csr(n, ncols, rp, col, rng.uniform(-1, 1, rp[-1]))
    x = rng.uniform(-1, 1, ncols)
    op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=seg_cols, seg_mode=2)
    assert op.segments() == (seg_cols, ncols // seg_cols)
    xt = torch.from_numpy(x).cuda()
    y, y2 = op(xt), op(xt)
    torch.cuda.synchronize()
    assert torch.equal(y, y2)  # deterministic
    assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                      abs_spmv(oracle, A, x), what=f"sorted blocks {seg_cols}")


@pytest.mark.parametrize("nseg", [2, 4])
def test_sorted_blocks_bits(oracle, nseg):
    # rows of at most 32 entries per segment are summed left to right by one
    # lane, and the segment sums added in segment order from 0.0: y equals,
    # bit for bit, the oracle's row loop over each segment's columns alone,
    # then those partial results added in order
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(41 + nseg)
    n, ncols = 20_000, 64_000
    lens = np.minimum(rng.poisson(12, n), 32)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) for L in lens]).astype(np.int32)
    A = Csr(n, ncols, rp, col, rng.uniform(-1, 1, rp[-1]))
    x = rng.uniform(-1, 1, ncols)
    w = ncols // nseg
    op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=w, seg_mode=2)
    assert op.segments() == (w, nseg)
    y = op(torch.from_numpy(x).cuda()).cpu().numpy()
    rows = np.repeat(np.arange(n), lens)
    want = np.zeros(n)
    for k in range(nseg):
        keep = (A.col >= k * w) & (A.col < (k + 1) * w)
        rpk = np.zeros(n + 1, np.int64)
        rpk[1:] = np.cumsum(np.bincount(rows[keep], minlength=n))
        want = want + oracle.spmv_f64(rpk, A.col[keep], A.val[keep], x)
    assert np.array_equal(y, want)


def test_sorted_blocks_degenerate(oracle):
    # no entries at all; entries in one column segment only (the other
    # segments empty); a single row
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(31)
    cases = [("no entries", Csr(50, 4000, np.zeros(51, np.int64), np.zeros(0, np.int32),
                                np.zeros(0))),
             ("one segment", Csr(300, 4000, np.arange(0, 901, 3, dtype=np.int64),
                                 np.tile(np.array([5, 17, 900], np.int32), 300),
                                 rng.uniform(-1, 1, 900))),
             ("one row", Csr(1, 4000, np.array([0, 4000], np.int64),
                             np.arange(4000, dtype=np.int32), rng.uniform(-1, 1, 4000)))]
    for name, A in cases:
        x = rng.uniform(-1, 1, A.ncols)
        op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=1000, seg_mode=2)
        assert op.segments() == (1000, 4), name
        y = op(torch.from_numpy(x).cuda())
        torch.cuda.synchronize()
        assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                          abs_spmv(oracle, A, x), what=f"sorted blocks, {name}")


def test_sorted_blocks_powerlaw_auto(oracle, bsp_host):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    A = bsp_host.powerlaw_1m()
    x = bsp_host.uniform(A.ncols, seed=13)
    op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_mode=2)
    assert op.segments() == (1 << 18, 4)  # 2 MB slices of x, two XCDs each
    y = op(torch.from_numpy(x).cuda())
    torch.cuda.synchronize()
    assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                      abs_spmv(oracle, A, x), what="powerlaw sorted blocks")


def test_segment_sums_y_not_16_byte_aligned(bsp_host):
    # the partial sums of an even row count are added two rows per lane
    # (16-byte stores) only into a 16-byte aligned y; an 8-byte aligned y
    # takes the row-per-lane launch, with the same bits
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    A = bsp_host.powerlaw(300_000, 20.0, 1.2, 5000, seed=1)
    op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=75_000)
    assert op.segments() == (75_000, 4)
    x = torch.from_numpy(bsp_host.uniform(A.ncols, 4)).cuda()
    buf = torch.full((A.nrows + 1,), 7.0, dtype=torch.float64, device="cuda")
    y = op(x)
    y1 = op(x, buf[1:])
    torch.cuda.synchronize()
    assert torch.equal(y, y1) and buf[0].item() == 7.0


@pytest.mark.parametrize("chunk", [1024, 2048])
def test_adaptive_variants(oracle, bsp_host, chunk):
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(17)
    lens = np.array([0, 3, 0, 5000, 1, 129, 513, 1025, 2047, 2049, 17] + [0] * 40 + [1] * 3000 +
                    [40] * 500)
    n, ncols = lens.size, 9000
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) for L in lens]).astype(np.int32)
    mats = [Csr(n, ncols, rp, col, rng.uniform(-1, 1, rp[-1])), bsp_host.stencil27_64(),
            bsp_host.powerlaw(300_000, 20.0, 1.2, 5000, seed=1)]
    for A in mats:
        x = bsp_host.uniform(A.ncols, 2)
        for seg in (0, -1):
            op = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=seg, chunk=chunk)
            y = op(torch.from_numpy(x).cuda())
            y2 = op(torch.from_numpy(x).cuda())
            torch.cuda.synchronize()
            y, y2 = y.cpu().numpy(), y2.cpu().numpy()
            assert np.array_equal(y, y2)  # deterministic
            assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                              abs_spmv(oracle, A, x), what=f"n={A.nrows} {chunk}/{seg}")


def _diag_plus(n, extra_row, extra_col):
    """n x n identity-like CSR (one entry per row on the diagonal, ragged
    values) plus one entry (extra_row, extra_col)."""
    from benchsp_amd.host import Csr
    rows = np.concatenate([np.arange(n), [extra_row]])
    cols = np.concatenate([np.arange(n), [extra_col]])
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rng = np.random.default_rng(n + extra_col)
    return Csr(n, n, np.cumsum(rp), cols.astype(np.int32), rng.uniform(-1, 1, rows.size))


@pytest.mark.parametrize("chunk", [1024, 2048])
def test_adaptive_narrow_cols(oracle, bsp_host, chunk):
    # 16-bit column offsets: the same bits as the int32 plan wherever every
    # block's columns lie within int16 of its first row, and refused past it
    import torch
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    cases = [("stencil27-64", bsp_host.stencil27_64(), True),
             ("stencil7-80", bsp_host.stencil3d(80, 80, 80, 7), True),
             ("powerlaw-300k", bsp_host.powerlaw(300_000, 20.0, 1.2, 5000, seed=1), False),
             ("edge+32767", _diag_plus(40_000, 0, 32767), True),
             ("edge+32768", _diag_plus(40_000, 0, 32768), False),
             ("edge-32768", _diag_plus(40_000, 32768, 0), True),
             ("edge-far", _diag_plus(40_000, 39_999, 0), False)]
    for name, A, expect in cases:
        d = DeviceCsr.from_host(A)
        x = bsp_host.uniform(A.ncols, 6)
        xd = torch.from_numpy(x).cuda()
        on = CsrSpmv(d, "adaptive", seg_cols=-1, chunk=chunk, narrow_cols=1)
        off = CsrSpmv(d, "adaptive", seg_cols=-1, chunk=chunk)
        if expect is not None:
            assert on.narrow() == expect, name
        assert not off.narrow()
        y_on, y_off = on(xd).cpu().numpy(), off(xd).cpu().numpy()
        assert np.array_equal(y_on, y_off), name
        assert_spmv_close(y_on, oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                          abs_spmv(oracle, A, x), what=f"{name} narrow")


def test_adaptive_segments_reject_unsorted():
    import torch
    from benchsp_amd import BspError
    from benchsp_amd.device import CsrSpmv, DeviceCsr
    from benchsp_amd.host import Csr
    A = Csr(2, 100, np.array([0, 2, 3]), np.array([50, 7, 3], np.int32), np.ones(3))
    with pytest.raises(BspError) as e:
        CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=10)
    assert e.value.status == 4
    # unsorted rows are still fine unsegmented
    y = CsrSpmv(DeviceCsr.from_host(A), "adaptive", seg_cols=-1)(torch.ones(100, dtype=torch.float64,
                                                                             device="cuda"))
    assert y.cpu().tolist() == [2.0, 1.0]


def test_reference_surface_wrappers(oracle, bsp_host):
    import ctypes as C
    from benchsp_amd._lib import lib
    A = bsp_host.stencil3d(12, 13, 14, 27, seed=3)
    x = bsp_host.uniform(A.ncols, seed=4)
    rp = A.rowptr.astype(np.int32)
    y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    for fn in ("bsp_spmv_fp64_warpper", "bsp_spmv_fp64_balance_warpper"):
        out = np.full(A.nrows, 123.0)
        t = C.c_double(-1)
        rc = getattr(lib(), fn)(rp.ctypes.data, A.col.ctypes.data, A.nrows, A.ncols, A.nnz,
                                A.val.ctypes.data, x.ctypes.data, out.ctypes.data, 3, C.byref(t))
        assert rc == 0 and t.value >= 0
        assert_spmv_close(out, y_ref, abs_spmv(oracle, A, x), what=fn)


@pytest.mark.parametrize("rp,col,ncols,ok,sorted_", [
    ([0, 2, 3, 5], [0, 3, 1, 2, 4], 5, True, True),
    ([0, 2, 3, 5], [3, 0, 1, 2, 4], 5, True, False),   # unsorted: valid, reported
    ([0, 2, 2, 5], [0, 3, 1, 2, 2], 5, True, False),   # repeated column: not strictly ascending
    ([0, 3, 2, 5], [0, 3, 1, 2, 4], 5, False, None),   # rowptr decreases
    ([0, 2, 3, 4], [0, 3, 1, 2, 4], 5, False, None),   # rowptr ends before nnz
    ([1, 2, 3, 5], [0, 3, 1, 2, 4], 5, False, None),   # rowptr starts past 0
    ([0, 2, 3, 5], [0, 5, 1, 2, 4], 5, False, None),   # column == ncols
    ([0, 2, 3, 5], [0, -1, 1, 2, 4], 5, False, None),  # negative column
])
def test_device_csr_check(rp, col, ncols, ok, sorted_):
    import ctypes as C
    import torch
    from benchsp_amd._lib import BspError, CsrDesc, lib
    from benchsp_amd.device import ptr, stream_handle
    rpd = torch.tensor(rp, dtype=torch.int32, device="cuda")
    cd = torch.tensor(col, dtype=torch.int32, device="cuda")
    vd = torch.ones(len(col), dtype=torch.float64, device="cuda")
    d = CsrDesc(len(rp) - 1, ncols, len(col), ptr(rpd), ptr(cd), ptr(vd), 0)
    s = C.c_int32(-1)
    if ok:
        lib().bsp_csr_check(C.byref(d), stream_handle(), C.byref(s))
        assert s.value == int(sorted_)
    else:
        with pytest.raises(BspError) as e:
            lib().bsp_csr_check(C.byref(d), stream_handle(), C.byref(s))
        assert e.value.status == 1


def test_line_probe_sums():
    # x of ones: each workgroup's stored sum is its 256 threads x 32 runs x
    # two doubles per 16-byte load
    import ctypes as C
    import torch
    from benchsp_amd._lib import lib
    from benchsp_amd.device import ptr, stream_handle
    x = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
    g = C.c_int64()
    lib().bsp_line_probe(None, x.numel(), 4, 16 << 20, None, 0, C.byref(g), None)
    sink = torch.zeros(g.value, dtype=torch.float64, device="cuda")
    lib().bsp_line_probe(ptr(x), x.numel(), 4, 16 << 20, ptr(sink), sink.numel(), None,
                         stream_handle())
    torch.cuda.synchronize()
    assert torch.all(sink == 256 * 32 * 2)


def test_gather_probe_sums():
    # x of ones: each workgroup's stored sum is its 256 threads x 32 gathers
    import ctypes as C
    import torch
    from benchsp_amd._lib import lib
    from benchsp_amd.device import ptr, stream_handle
    x = torch.ones(1 << 20, dtype=torch.float64, device="cuda")
    g = C.c_int64()
    lib().bsp_gather_probe(None, x.numel(), 4, 1_000_000, None, 0, C.byref(g), None)
    sink = torch.zeros(g.value, dtype=torch.float64, device="cuda")
    lib().bsp_gather_probe(ptr(x), x.numel(), 4, 1_000_000, ptr(sink), sink.numel(), None,
                           stream_handle())
    torch.cuda.synchronize()
    assert torch.all(sink == 256 * 32)

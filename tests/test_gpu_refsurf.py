"""The reference's host-pointer call surface (bsp_<reference name> in
include/bsp.h), called through the C-ABI exactly as the reference's benches
call the originals, against the oracle."""
import ctypes as C
import glob
import os

import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close, rel_fro

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _L():
    from benchsp_amd._lib import lib
    return lib()


def _p(a):
    return a.ctypes.data


def _rand_csr(rng, lens, ncols):
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) if L else
                          np.zeros(0, np.int64) for L in lens]).astype(np.int32)
    return Csr(len(lens), ncols, rp, col, rng.uniform(-1, 1, int(rp[-1])))


def _inputs(bsp_host):
    rng = np.random.default_rng(11)
    yield "stencil27-12x13x14", bsp_host.stencil3d(12, 13, 14, 27, seed=3)
    yield "powerlaw-20k", bsp_host.powerlaw(20_000, 15.0, 1.3, 3000, seed=8)
    yield "ragged", _rand_csr(rng, np.concatenate([rng.integers(0, 9, 500), [400, 0, 129, 128]]),
                              900)


def _dbsr_mat(A):
    from benchsp_amd._lib import CSRFormatF64, DbsrMat
    rp, col, val = A.rowptr.astype(np.int32), A.col.astype(np.int32), np.ascontiguousarray(A.val)
    csr = CSRFormatF64(_p(rp), _p(col), _p(val), A.nrows, A.ncols, A.nnz)
    m = DbsrMat()
    assert _L().bsp_csrTodbsrWithRecorder(C.byref(csr), C.byref(m)) == 0
    return m, (rp, col, val)


def _np(ptr, n, ct, dt):
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ct)), shape=(n,)).astype(dt).copy() \
        if n else np.zeros(0, dt)


def _mat_arrays(m):
    nb = m.blc_num
    return dict(recorder=_np(m.recorderMap, m.row, C.c_int32, np.int32),
                blcptr=_np(m.blcPtr, m.blc_row + 1, C.c_int32, np.int32),
                colidx=_np(m.colIdx, nb * 16, C.c_int32, np.int32),
                val=_np(m.blcVal, nb * 16, C.c_double, np.float64),
                map=_np(m.blcMap, nb, C.c_uint16, np.uint16),
                prefix=_np(m.prefix, m.warpnum + 1, C.c_int32, np.int32))


def test_csrTodbsrWithRecorder_matches_oracle(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        m, keep = _dbsr_mat(A)
        try:
            ref = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
            got = _mat_arrays(m)
            assert m.blc_row == ref["brows"] and m.blc_num == ref["nblocks"]
            assert m.warpnum == ref["warpnum"] and m.blocknnz == 16 * ref["nblocks"]
            for k in got:
                assert np.array_equal(got[k], ref[k]), (name, k)
        finally:
            _L().bsp_dbsrMat_free(C.byref(m))
        assert not m.blcPtr and not m.prefix


def test_csrTodbsrWithRecorder_golden():
    files = sorted(glob.glob(os.path.join(GOLDEN, "dbsr_*.npz")))
    assert files
    for f in files:
        g = np.load(f)
        A = Csr(int(g["nrows"]), int(g["ncols"]), g["rowptr"].astype(np.int64), g["col"],
                g["val"])
        m, keep = _dbsr_mat(A)
        try:
            got = _mat_arrays(m)
            for k, gk in (("recorder", "recorder"), ("blcptr", "blcptr"), ("colidx", "colidx"),
                          ("val", "blcval"), ("map", "map"), ("prefix", "prefix")):
                assert np.array_equal(got[k], g[gk]), (os.path.basename(f), k)
        finally:
            _L().bsp_dbsrMat_free(C.byref(m))


@pytest.mark.parametrize("fn", ["bsp_amgT_spmv_fp64_dbsr", "bsp_amgT_spmv_fp64_dbsr_balance"])
def test_amgT_spmv_fp64_dbsr(oracle, bsp_host, fn):
    for name, A in _inputs(bsp_host):
        m, keep = _dbsr_mat(A)
        try:
            x = bsp_host.uniform(A.ncols, 7)
            ys = np.full(A.nrows, np.nan)
            t = C.c_double(-1)
            assert getattr(_L(), fn)(C.byref(m), _p(x), _p(ys), 2, 3, C.byref(t)) == 0
            assert t.value >= 0
            rec = _mat_arrays(m)["recorder"]
            y = np.empty_like(ys)
            y[rec] = ys  # bench/spmvfp64.cpp:163-166 un-permutes the same way
            assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                              abs_spmv(oracle, A, x), what=f"{name} {fn}")
        finally:
            _L().bsp_dbsrMat_free(C.byref(m))


def test_amgT_spmv_fp64_both_kernels(oracle, bsp_host):
    # the 27-point stencil packs > 10 entries per 4x4 block (MFMA kernel);
    # the ragged random matrix packs few (VALU kernel)
    for name, A in _inputs(bsp_host):
        rp = A.rowptr.astype(np.int32)
        x = bsp_host.uniform(A.ncols, 9)
        y = np.full(A.nrows, np.nan)
        t = C.c_double(-1)
        rc = _L().bsp_amgT_spmv_fp64(_p(rp), _p(A.col), _p(A.val), _p(x), _p(y), A.nrows, A.ncols,
                                     A.nnz, 3, C.byref(t))
        assert rc == 0
        assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=name)


def test_bcsr_spmv_fp64(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp, col = A.rowptr.astype(np.int64), A.col.astype(np.int64)
        x = bsp_host.uniform(A.ncols, 10)
        y = np.full(A.nrows, np.nan)
        _L().bsp_bcsr_spmv_fp64(_p(rp), _p(col), _p(A.val), _p(x), _p(y), A.nrows, A.ncols, A.nnz, 2)
        assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=name)


def test_spmv_all_dasp(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp = A.rowptr.astype(np.int32)
        x = bsp_host.uniform(A.ncols, 12)
        yp = np.full(A.nrows, np.nan)
        order = np.full(A.nrows, -1, np.int32)
        t = C.c_double(-1)
        _L().bsp_spmv_all(b"unused", _p(A.val), _p(rp), _p(A.col), _p(x), _p(yp), _p(order),
                          A.nrows, A.ncols, A.nnz, 1, 0.75, 256, C.byref(t))
        assert t.value >= 0
        lay = oracle.dasp(A.rowptr, A.col, A.val, 0.75, 256)
        assert np.array_equal(order, lay["order_rid"]), name
        y = np.empty_like(yp)
        y[order] = yp
        assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=name)


def test_MySpmvWithPrefp64(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp, col = A.rowptr.astype(np.int64), A.col.astype(np.int64)
        x = bsp_host.uniform(A.ncols, 13)
        y = np.full(A.nrows, np.nan)
        _L().bsp_MySpmvWithPrefp64(_p(rp), _p(col), A.nrows, A.ncols, A.nnz, _p(A.val), _p(x), _p(y))
        assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=name)


def test_MySpmvWithPrefp16(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp = A.rowptr.astype(np.int32)
        v = A.val.astype(np.float32)
        x = bsp_host.uniform(A.ncols, 14).astype(np.float32)
        y = np.full(A.nrows, np.nan, np.float32)
        _L().bsp_MySpmvWithPrefp16(_p(rp), _p(A.col), A.nrows, A.ncols, A.nnz, _p(v), _p(x), _p(y))
        y_ref = oracle.spmv_f16(A.rowptr, A.col, v, x)
        lens = np.diff(A.rowptr)
        short = lens <= 128
        # rows up to 128 entries: one thread, the oracle's order -> bit-exact
        assert np.array_equal(y[short], y_ref[short]), name
        # longer rows: a wave-wide FP32 sum, then one rounding to half: at most
        # one half ulp apart plus the FP32 reordering error
        absx = np.abs(v.astype(np.float64))
        bound = oracle.spmv_f64(A.rowptr, A.col, absx, np.abs(x.astype(np.float64)))
        tol = 2.0 ** -10 * np.abs(y_ref) + 2.0 ** -18 * bound
        assert np.all(np.abs(y - y_ref)[~short] <= tol[~short]), name


def test_smatSpmmfp16(oracle, bsp_host):
    A = bsp_host.blocksparse(32, 24, 16, 90, seed=5)
    rp = A.rowptr.astype(np.int32)
    v = A.val.astype(np.float32)
    for n in (32, 96, 256):
        out = np.full((A.nrows, n), np.nan, np.float32)
        _L().bsp_smatSpmmfp16(_p(rp), _p(A.col), A.nrows, A.ncols, n, A.nnz, _p(v), _p(out))
        # B = 1: every column of C is the row sums of A
        ref = oracle.csr_spmm_f64(A.rowptr, A.col, v.astype(np.float64), np.ones((A.ncols, n)))
        assert rel_fro(out, ref) <= 1e-2
        assert np.all(out == out[:, :1])


def test_cusparse_spmv_fp64_is_rocsparse(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp, col = A.rowptr.astype(np.int64), A.col.astype(np.int64)
        x = bsp_host.uniform(A.ncols, 15)
        y = np.full(A.nrows, np.nan)
        t = C.c_double(-1)
        rc = _L().bsp_cusparse_spmv_fp64(_p(rp), _p(col), _p(A.val), _p(x), _p(y), A.nrows,
                                         A.ncols, A.nnz, 3, C.byref(t))
        assert rc == 0 and t.value >= 0
        assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=name)


@pytest.mark.parametrize("alg", [0, 2, 3, 7])
def test_rocsparse_plan_on_device_arrays(oracle, bsp_host, alg):
    # every CSR algorithm of the comparator (default, adaptive, rowsplit,
    # LRB) against the oracle
    import torch
    from benchsp_amd.device import DeviceCsr
    from benchsp_amd._lib import CsrDesc
    A = bsp_host.powerlaw(50_000, 12.0, 1.2, 5000, seed=2)
    d = DeviceCsr.from_host(A)
    desc = d.desc()
    h = C.c_void_p()
    _L().bsp_rocsparse_plan_create(C.byref(desc), alg, None, C.byref(h))
    try:
        x = bsp_host.uniform(A.ncols, 16)
        xd = torch.from_numpy(x).cuda()
        yd = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        for _ in range(2):
            _L().bsp_rocsparse_spmv(h, xd.data_ptr(), yd.data_ptr(), None)
        torch.cuda.synchronize()
        assert_spmv_close(yd.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                          abs_spmv(oracle, A, x), what=f"rocsparse alg {alg}")
    finally:
        _L().bsp_rocsparse_plan_destroy(h)
    assert isinstance(desc, CsrDesc)


def test_ref_spmv_gpu_bit_identical_to_cpu_loop(oracle, bsp_host):
    # include/ref.h on the GPU: thread per row, left to right, no FMA
    for name, A in _inputs(bsp_host):
        rp64, col64 = A.rowptr.astype(np.int64), A.col.astype(np.int64)
        x = bsp_host.uniform(A.ncols, 17)
        y = np.full(A.nrows, np.nan)
        t = C.c_double(-1)
        _L().bsp_ref_spmv_fp64(_p(rp64), _p(col64), _p(A.val), _p(x), _p(y), A.nrows, A.ncols,
                               C.byref(t))
        assert np.array_equal(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x)), name
        assert t.value >= 0
        rp = A.rowptr.astype(np.int32)
        v32, x32 = A.val.astype(np.float32), x.astype(np.float32)
        for fn, ref in (("bsp_ref_spmv_fp32", oracle.spmv_f32), ("bsp_ref_spmv_fp16", oracle.spmv_f16)):
            y32 = np.full(A.nrows, np.nan, np.float32)
            getattr(_L(), fn)(_p(rp), _p(A.col), _p(v32), _p(x32), _p(y32), A.nrows, A.ncols)
            assert np.array_equal(y32, ref(A.rowptr, A.col, v32, x32)), (name, fn)


def test_cusparse_fp32_fp16_surfaces_are_rocsparse(oracle, bsp_host):
    for name, A in _inputs(bsp_host):
        rp = A.rowptr.astype(np.int32)
        v32 = A.val.astype(np.float32)
        x32 = bsp_host.uniform(A.ncols, 18).astype(np.float32)
        absb = oracle.spmv_f64(A.rowptr, A.col, np.abs(v32.astype(np.float64)),
                               np.abs(x32.astype(np.float64)))
        # FP32: any summation order, FP32 rounding per step
        y = np.full(A.nrows, np.nan, np.float32)
        assert _L().bsp_cusparse_spmv(_p(rp), _p(A.col.astype(np.int64)), _p(v32), _p(x32), _p(y),
                                      A.nrows, A.ncols, A.nnz) == 0
        y64 = oracle.spmv_f64(A.rowptr, A.col, v32.astype(np.float64), x32.astype(np.float64))
        assert np.all(np.abs(y - y64) <= 2.0 ** -20 * absb + 1e-30), name
        # FP16 in, FP32 compute, FP16 out: within one half ulp of the oracle's
        # FP16 result plus the FP32 reordering error
        ref16 = oracle.spmv_f16(A.rowptr, A.col, v32, x32)
        vh = v32.astype(np.float16)
        xh = x32.astype(np.float16)
        yh = np.zeros(A.nrows, np.float16)
        assert _L().bsp_cusparse_spmv_fp16(_p(rp), _p(A.col), _p(vh), _p(xh), _p(yh), A.nrows,
                                           A.ncols, A.nnz) == 0
        tol = 2.0 ** -10 * np.abs(ref16) + 2.0 ** -18 * absb + 2.0 ** -24
        assert np.all(np.abs(yh.astype(np.float32) - ref16) <= tol), name
        yw = np.full(A.nrows, np.nan, np.float32)
        _L().bsp_cusparse_spmv_fp16_warpper(_p(rp), _p(A.col), _p(v32), _p(x32), _p(yw), A.nrows,
                                            A.ncols, A.nnz)
        assert np.array_equal(yw, yh.astype(np.float32)), name


def test_compat_driver_runs():
    # the C++ driver over include/benchsp_compat.hpp, built by build()
    import subprocess
    from benchsp_amd import build as b
    assert os.path.exists(b.COMPAT_BIN), "run __graft_entry__.build() first"
    r = subprocess.run([b.COMPAT_BIN, "96"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr


def test_compat_header_builders_match_oracle(oracle, tmp_path):
    # csrTodbsr + balanceDbsr, csrRecorderSelf and csrToBcsr (double 4x4,
    # half 16x16) called through include/benchsp_compat.hpp by the C++
    # driver, each array against the oracle's restatement
    import subprocess
    from benchsp_amd import build as b
    r = subprocess.run([b.COMPAT_BIN, "32", str(tmp_path)], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    got = {}
    for line in open(tmp_path / "manifest.txt"):
        name, dt, n = line.split()
        if dt == "int":
            got[name] = int(n)
        else:
            got[name] = np.fromfile(tmp_path / f"{name}.bin", dtype=dt, count=int(n))
    rp, col, val = got["in_rowptr"].astype(np.int64), got["in_col"], got["in_val"]
    n, ncols = rp.size - 1, got["ncols"]
    assert np.ptp(np.diff(rp)) >= 30  # ragged rows
    o = oracle.dbsr(n, ncols, rp, col, val, sort_rows=False)
    for k in ("blcptr", "colidx", "val", "map", "prefix"):
        assert np.array_equal(got[f"dbsr_{k}"], o[k]), k
    rec = oracle.dbsr(n, ncols, rp, col, val)["recorder"]
    lens = np.diff(rp)
    assert np.array_equal(np.diff(got["rec_rowptr"]), lens[rec])
    perm = np.concatenate([np.arange(rp[i], rp[i + 1]) for i in rec])
    assert np.array_equal(got["rec_col"], col[perm]) and np.array_equal(got["rec_val"], val[perm])
    b4 = oracle.bcsr(n, ncols, rp, col, val, 4, 4)
    assert np.array_equal(got["bcsr4_rowptr"], b4["rowptr"])
    assert np.array_equal(got["bcsr4_col"], b4["colidx"])
    assert np.array_equal(got["bcsr4_val"], b4["val"])
    # the dense block map: block index per (block row, block column), -1 if empty
    creg = -(-ncols // 4)
    dense = np.full((b4["brows"], creg), -1, np.int64)
    for br in range(b4["brows"]):
        for k in range(b4["rowptr"][br], b4["rowptr"][br + 1]):
            dense[br, b4["colidx"][k]] = k
    assert np.array_equal(got["bcsr4_map"], dense.ravel())
    vh = val.astype(np.float32).astype(np.float16)  # the driver's __float2half((float)v)
    b16 = oracle.bcsr(n, ncols, rp, col, vh.astype(np.float64), 16, 16)
    assert np.array_equal(got["bcsr16_rowptr"], b16["rowptr"])
    assert np.array_equal(got["bcsr16_col"], b16["colidx"])
    assert np.array_equal(got["bcsr16_val"].astype(np.float64), b16["val"])


@pytest.mark.parametrize("dtype_name", ["float32", "float16"])
@pytest.mark.parametrize("transpose", [False, True])
@pytest.mark.parametrize("alg", [0, 1, 4, 5, 9])
def test_rocsparse_spmm_comparator(oracle, bsp_host, dtype_name, transpose, alg):
    # every CSR SpMM algorithm the comparator may report; a combination that
    # rocSPARSE itself refuses comes back as BSP_ERR_HIP (2), never a wrong C
    import torch
    from benchsp_amd._lib import BspError
    from benchsp_amd.device import DeviceCsr, RocsparseSpmm
    dt = getattr(torch, dtype_name)
    A = bsp_host.powerlaw(3000, 12.0, 1.2, 400, seed=6, ncols=2500)
    N = 64
    rows_b = A.nrows if transpose else A.ncols
    B = bsp_host.normal(rows_b * N, 19).reshape(rows_b, N)
    d = DeviceCsr.from_host(A, dtype=dt)
    Bd = torch.from_numpy(B).to("cuda", dt)
    try:
        op = RocsparseSpmm(d, N, transpose, alg)
        Cm = op(Bd)
    except BspError as e:
        assert alg != 0 and e.status == 2, str(e)
        return
    torch.cuda.synchronize()
    a = d.val.double().cpu().numpy()  # the operands as stored
    Bq = Bd.double().cpu().numpy()
    if transpose:
        ref = oracle.csr_spmm_t_f64(A.rowptr, A.col, a, Bq, A.ncols)
    else:
        ref = oracle.csr_spmm_f64(A.rowptr, A.col, a, Bq)
    assert rel_fro(Cm.cpu().numpy(), ref) <= 1e-5

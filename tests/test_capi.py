"""CPU-side checks of the C-ABI library: it loads, exports every function the
public headers declare, reports errors through status codes, and its
deterministic generators produce the documented matrices."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(bsp_\w+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_function():
    from benchsp_amd._lib import lib
    L = lib().cdll
    missing = [n for n in _declared() if not hasattr(L, n)]
    assert not missing, missing
    assert len(_declared()) >= 15


def test_python_binding_declares_only_real_symbols():
    from benchsp_amd._lib import exported_symbols, lib
    for n in exported_symbols():
        assert hasattr(lib().cdll, n), n


def test_python_binding_covers_every_declared_function():
    from benchsp_amd._lib import exported_symbols
    missing = sorted(set(_declared()) - set(exported_symbols()))
    assert not missing, missing


def test_struct_mirrors_match_header_layout():
    # the reference structs are mirrored field for field; on LP64 the dbsrMat
    # layout is 6 ints, 2 doubles, 4 pointers, 2 int32, 4 pointers
    from benchsp_amd._lib import CSRFormatF64, DbsrMat
    assert C.sizeof(CSRFormatF64) == 3 * 8 + 3 * 4 + 4
    assert C.sizeof(DbsrMat) == 6 * 4 + 2 * 8 + 4 * 8 + 2 * 4 + 4 * 8
    assert DbsrMat.blocknnz.offset == 6 * 4 + 2 * 8 + 4 * 8


def test_status_and_last_error():
    from benchsp_amd._lib import BspError, CsrDesc, lib
    L = lib()
    assert L.bsp_status_string(0) == b"ok"
    assert L.bsp_version().startswith(b"bsp")
    d = CsrDesc(-1, 4, 0, None, None, None, 0)
    h = C.c_void_p()
    with pytest.raises(BspError) as e:
        L.bsp_csr_plan_create(C.byref(d), 0, None, C.byref(h))
    assert e.value.status == 1 and "negative" in str(e.value)
    d = CsrDesc(10, 10, 1 << 31, None, None, None, 0)
    with pytest.raises(BspError) as e:
        L.bsp_csr_plan_create(C.byref(d), 0, None, C.byref(h))
    assert e.value.status == 6
    with pytest.raises(BspError):
        L.bsp_gen_stencil3d(4, 4, 4, 9, 0, C.byref(h))  # only 7 or 27 points


def test_reference_wrapper_rejects_null_without_exiting():
    from benchsp_amd._lib import lib
    t = C.c_double(0)
    assert lib().bsp_spmv_fp64_warpper(None, None, 4, 4, 4, None, None, None, 1, C.byref(t)) == 1


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


# Malformed host patterns: (rowptr, col, ncols, message fragment). nrows is 3
# and nnz 4 in every case; each must be refused before any upload.
_BAD_CSR = [
    ([1, 2, 3, 4], [0, 1, 2, 3], 4, "rowptr[0]"),
    ([0, 3, 2, 4], [0, 1, 2, 3], 4, "decreases"),
    ([0, 1, 2, 3], [0, 1, 2, 3], 4, "equal nnz"),
    ([0, 1, 2, 4], [0, 1, 4, 3], 4, "out of range"),
    ([0, 1, 2, 4], [0, -1, 2, 3], 4, "out of range"),
]


def _last_error():
    from benchsp_amd._lib import lib
    return lib().bsp_last_error().decode()


@pytest.mark.parametrize("rp,col,ncols,msg", _BAD_CSR)
def test_host_csr_rejected_before_upload_int32(rp, col, ncols, msg):
    from benchsp_amd._lib import lib
    L = lib()
    rp32, c32 = np.array(rp, np.int32), np.array(col, np.int32)
    v, x, y = np.ones(4), np.ones(ncols), np.zeros(3)
    t = C.c_double(0)
    for fn in (L.bsp_spmv_fp64_warpper, L.bsp_spmv_fp64_balance_warpper):
        assert fn(_p(rp32), _p(c32), 3, ncols, 4, _p(v), _p(x), _p(y), 1, C.byref(t)) == 1
        assert msg in _last_error()
    assert L.bsp_amgT_spmv_fp64(_p(rp32), _p(c32), _p(v), _p(x), _p(y), 3, ncols, 4, 1,
                                C.byref(t)) == 1
    assert msg in _last_error()
    v32, x32, y32 = np.ones(4, np.float32), np.ones(ncols, np.float32), np.zeros(3, np.float32)
    L.bsp_ref_spmv_fp32(_p(rp32), _p(c32), _p(v32), _p(x32), _p(y32), 3, ncols)
    assert not y32.any()
    L.bsp_MySpmvWithPrefp16(_p(rp32), _p(c32), 3, ncols, 4, _p(v32), _p(x32), _p(y32))
    assert not y32.any() and msg in _last_error()
    from benchsp_amd._lib import CSRFormatF64, DbsrMat
    csr = CSRFormatF64(_p(rp32), _p(c32), _p(v), 3, ncols, 4)
    out = DbsrMat()
    assert L.bsp_csrTodbsrWithRecorder(C.byref(csr), C.byref(out)) == 1
    assert msg in _last_error()


@pytest.mark.parametrize("rp,col,ncols,msg", _BAD_CSR)
def test_host_csr_rejected_before_upload_int64(rp, col, ncols, msg):
    from benchsp_amd._lib import lib
    L = lib()
    rp64, c64 = np.array(rp, np.int64), np.array(col, np.int64)
    v, x, y = np.ones(4), np.ones(ncols), np.zeros(3)
    t = C.c_double(0)
    assert L.bsp_cusparse_spmv_fp64(_p(rp64), _p(c64), _p(v), _p(x), _p(y), 3, ncols, 4, 1,
                                    C.byref(t)) == 1
    assert msg in _last_error()
    L.bsp_MySpmvWithPrefp64(_p(rp64), _p(c64), 3, ncols, 4, _p(v), _p(x), _p(y))
    assert not y.any() and msg in _last_error()
    L.bsp_bcsr_spmv_fp64(_p(rp64), _p(c64), _p(v), _p(x), _p(y), 3, ncols, 4, 1)
    assert not y.any() and msg in _last_error()


def test_int64_index_that_does_not_fit_int32_is_rejected():
    from benchsp_amd._lib import lib
    L = lib()
    rp = np.array([0, 1, 2], np.int64)
    col = np.array([0, (1 << 32) + 1], np.int64)  # would narrow to 1
    v, x, y = np.ones(2), np.ones(4), np.zeros(2)
    t = C.c_double(0)
    assert L.bsp_cusparse_spmv_fp64(_p(rp), _p(col), _p(v), _p(x), _p(y), 2, 4, 2, 1,
                                    C.byref(t)) == 1
    assert "out of range" in _last_error()


def _dbsr_host(recorder, blcptr, colidx, row=5, col=6):
    from benchsp_amd._lib import DbsrMat
    keep = [np.array(recorder, np.int32), np.array(blcptr, np.int32),
            np.array(colidx, np.int32)]
    nb = len(colidx) // 16
    keep += [np.zeros(16 * nb), np.zeros(nb, np.uint16), np.zeros(2, np.int32)]
    m = DbsrMat()
    m.row, m.col, m.nnz, m.blc_row, m.blc_num = row, col, 0, (row + 3) // 4, nb
    m.blocknnz, m.warpnum = 16 * nb, 1
    m.recorderMap, m.blcPtr, m.colIdx, m.blcVal, m.blcMap, m.prefix = map(_p, keep)
    return m, keep


@pytest.mark.parametrize("recorder,blcptr,colidx,msg", [
    ([0, 1, 2, 3, 3], [0, 1, 2], [0] * 32, "permutation"),
    ([0, 1, 2, 3, 5], [0, 1, 2], [0] * 32, "permutation"),
    ([4, 3, 2, 1, 0], [0, 2, 1], [0] * 32, "decreases"),
    ([4, 3, 2, 1, 0], [0, 1, 1], [0] * 32, "blocknnz"),
    ([4, 3, 2, 1, 0], [0, 1, 2], [0] * 31 + [6], "colIdx"),
    ([4, 3, 2, 1, 0], [1, 1, 2], [0] * 32, "blcPtr[0]"),
])
def test_host_dbsr_rejected_before_upload(recorder, blcptr, colidx, msg):
    from benchsp_amd._lib import BspError, lib
    L = lib()
    m, keep = _dbsr_host(recorder, blcptr, colidx)
    h = C.c_void_p()
    with pytest.raises(BspError) as e:
        L.bsp_dbsr_from_host(C.byref(m), None, C.byref(h))
    assert e.value.status == 1 and msg in str(e.value)
    x, y, t = np.ones(6), np.zeros(5), C.c_double(0)
    assert L.bsp_amgT_spmv_fp64_dbsr(C.byref(m), _p(x), _p(y), 1, 1, C.byref(t)) == 1
    assert msg in _last_error()


def test_generator_sizes(bsp_host):
    A = bsp_host.poisson5_256()
    assert (A.nrows, A.nnz) == (65536, 326656)
    assert np.all(np.diff(A.col[A.rowptr[5]:A.rowptr[6]]) > 0)
    B = bsp_host.stencil3d(6, 7, 8, 27, seed=1)
    assert B.nnz == (3 * 6 - 2) * (3 * 7 - 2) * (3 * 8 - 2)
    # diagonal = 1 + sum |off-diagonal| in every row
    for r in range(0, B.nrows, 13):
        s, e = B.rowptr[r], B.rowptr[r + 1]
        c, v = B.col[s:e], B.val[s:e]
        off = 0.0
        for a in np.abs(v[c != r]):  # left to right, as generated
            off += a
        assert v[c == r][0] == 1.0 + off
    S = bsp_host.stencil3d(5, 5, 5, 7)
    assert S.nnz == 7 * 125 - 6 * 25
    assert np.array_equal(np.unique(S.val), [-1.0, 6.0])


def test_generators_reject_bad_arguments_and_clamp_row_lengths():
    # non-finite distribution parameters and dimensions whose products would
    # overflow are argument errors; an extreme but finite sigma or mean gives
    # row lengths clamped to [1, cap] (never an undefined float-to-int cast)
    from benchsp_amd._lib import BspError, lib
    h = C.c_void_p()
    inf, nan = float("inf"), float("nan")
    bad = [("bsp_gen_powerlaw", (100, 50, 20.0, inf, 40, 1)),
           ("bsp_gen_powerlaw", (100, 50, 20.0, nan, 40, 1)),
           ("bsp_gen_stencil5", (1 << 40, 1 << 40)),
           ("bsp_gen_stencil3d", (1 << 30, 1 << 30, 1 << 30, 7, 1)),
           ("bsp_gen_blocksparse", (1 << 30, 1 << 26, 16, 10, 1)),
           ("bsp_gen_graph", (100, inf, 2.1, 1, 1))]
    for name, args in bad:
        with pytest.raises(BspError) as e:
            getattr(lib(), name)(*args, C.byref(h))
        assert e.value.status == 1, name
    for sigma, mean in ((50.0, 20.0), (1.0, 1e300)):
        lib().bsp_gen_powerlaw(100, 50, mean, sigma, 40, 1, C.byref(h))
        nr, nc, nz = C.c_int64(), C.c_int64(), C.c_int64()
        lib().bsp_host_csr_info(h, C.byref(nr), C.byref(nc), C.byref(nz))
        rp = np.zeros(101, np.int64)
        lib().bsp_host_csr_copy(h, rp.ctypes.data_as(C.c_void_p), None, None)
        lib().bsp_host_csr_free(h)
        lens = np.diff(rp)
        assert lens.min() >= 1 and lens.max() <= 40 and nz.value == rp[-1]


def test_generator_determinism_and_rows(bsp_host):
    P1 = bsp_host.powerlaw(5000, 20.0, 1.2, 300, seed=9)
    P2 = bsp_host.powerlaw(5000, 20.0, 1.2, 300, seed=9)
    assert np.array_equal(P1.col, P2.col) and np.array_equal(P1.val, P2.val)
    L = P1.row_lengths()
    assert L.min() >= 1 and L.max() <= 300 and 15 < L.mean() < 25
    for r in range(0, 5000, 97):
        c = P1.col[P1.rowptr[r]:P1.rowptr[r + 1]]
        assert np.all(np.diff(c) > 0)  # distinct, ascending
    x = bsp_host.uniform(10, 0)
    # element i of the stream is mix64(seed + (i+1)*golden)
    def mix(z):
        m = (1 << 64) - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
        return z ^ (z >> 31)
    expect = [2.0 * ((mix((i + 1) * 0x9E3779B97F4A7C15 & ((1 << 64) - 1)) >> 11) * 2.0**-53) - 1.0
              for i in range(10)]
    assert np.array_equal(x, expect)


def test_blocksparse_generator(bsp_host):
    B = bsp_host.blocksparse(32, 40, 16, 100, seed=3)
    assert B.nrows == 512 and B.ncols == 640 and B.nnz == 100 * 256
    D = B.to_dense().reshape(32, 16, 40, 16).transpose(0, 2, 1, 3)
    occ = np.abs(D).sum(axis=(2, 3)) > 0
    assert occ.sum() == 100


def test_graph_generator(bsp_host):
    G = bsp_host.graph(3000, 20.0, 2.1, True, seed=1)
    D = G.to_dense()
    assert np.array_equal(D != 0, (D != 0).T)  # symmetric pattern
    assert not np.any(np.diag(D))
    deg = (D != 0).sum(1)
    np.testing.assert_allclose(D[D != 0], (1 / np.sqrt(np.outer(deg, deg)))[D != 0], rtol=1e-15)


def test_compat_header_driver_builds(tmp_path):
    # include/benchsp_compat.hpp: the reference's C++ names and signatures
    # (double& avg_time, CSRFormat<double,int32_t>, dbsrMat) compile against
    # libbsp.so, and its static_asserts pin the struct layouts
    import subprocess
    from benchsp_amd import build as b
    out = tmp_path / "driver"
    r = subprocess.run([b.HIPCC, "-O0", "-std=c++17", "-I", os.path.join(ROOT, "include"),
                        b.COMPAT_SRC, "-L", b.HERE, "-lbsp", "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("which", ["bsr", "dbsr", "dasp", "dia", "spmm", "blockinfo",
                                   "rocsparse", "rocsparse_spmm", "batch", "autotune",
                                   "tuned_from_record"])
@pytest.mark.parametrize("size", ["nnz", "nrows", "ncols", "negative"])
def test_builders_reject_bad_sizes_before_any_gpu_work(which, size):
    # every builder validates the descriptor first: sizes past 2^31 - 2^20
    # (int32 indices with room for in-kernel offsets) give BSP_ERR_TOO_LARGE
    # (6), negative sizes BSP_ERR_INVALID (1), and no device call is made
    # (this runs without a GPU)
    from benchsp_amd._lib import BlockInfo, BspError, CsrDesc, TuneRecord, lib
    from benchsp_amd.device import BatchOptions
    big = (1 << 31) - (1 << 20) + 1
    n, m, z = {"nnz": (10, 10, big), "nrows": (big, 10, 5), "ncols": (10, big, 5),
               "negative": (10, -1, 5)}[size]
    d = CsrDesc(n, m, z, None, None, None, 0)  # FP64 values
    if which in ("rocsparse_spmm",):
        d = CsrDesc(n, m, z, None, None, None, 1)  # FP32 values
    h = C.c_void_p()
    L = lib()
    calls = {
        "bsr": lambda: L.bsp_bsr_build(C.byref(d), 4, 4, None, C.byref(h)),
        "dbsr": lambda: L.bsp_dbsr_build(C.byref(d), None, C.byref(h)),
        "dasp": lambda: L.bsp_dasp_build(C.byref(d), 0.75, 256, None, C.byref(h)),
        "dia": lambda: L.bsp_dia_build(C.byref(d), None, C.byref(h)),
        "spmm": lambda: L.bsp_spmm_plan_create(C.byref(d), None, C.byref(h)),
        "blockinfo": lambda: L.bsp_block_info(C.byref(d), 4, None, C.byref(BlockInfo())),
        "rocsparse": lambda: L.bsp_rocsparse_plan_create(C.byref(d), 0, None, C.byref(h)),
        "rocsparse_spmm": lambda: L.bsp_rocsparse_spmm_plan_create(C.byref(d), 64, 0, 0, None,
                                                                    C.byref(h)),
        "batch": lambda: L.bsp_batch_create(C.byref(d), 1, C.byref(BatchOptions(0, 0)), None,
                                            C.byref(h)),
        "autotune": lambda: L.bsp_autotune(C.byref(d), None, None, C.byref(h)),
        "tuned_from_record": lambda: L.bsp_tuned_from_record(C.byref(d), C.byref(TuneRecord()),
                                                             None, C.byref(h)),
    }
    with pytest.raises(BspError) as e:
        calls[which]()
    assert e.value.status == (1 if size == "negative" else 6), (which, size, str(e.value))


@pytest.mark.parametrize("rec", [dict(format=7), dict(format=0, chunk=64),
                                 dict(format=1, kernel=0), dict(format=3, threshold=0.0,
                                                                block_longest=256),
                                 dict(format=3, threshold=0.75, block_longest=4)])
def test_tune_record_validated_before_any_gpu_work(rec):
    # an empty matrix and a malformed record: BSP_ERR_INVALID (1), no device call
    from benchsp_amd._lib import BspError, CsrDesc, TuneRecord, lib
    d = CsrDesc(0, 0, 0, None, None, None, 0)
    r = TuneRecord()
    for k, v in rec.items():
        setattr(r, k, v)
    h = C.c_void_p()
    with pytest.raises(BspError) as e:
        lib().bsp_tuned_from_record(C.byref(d), C.byref(r), None, C.byref(h))
    assert e.value.status == 1, str(e.value)


def test_noalloc_spmv_rejects_sizes_beyond_its_schedule_array():
    # rows + entries above 2^31 do not fit the fixed 4 MB window table; the
    # check comes before any device access (the pointers are never read)
    from benchsp_amd._lib import BspError, CsrDesc, lib
    d = CsrDesc(1_500_000_000, 10, 1_000_000_000, 8, 8, 8, 0)
    with pytest.raises(BspError) as e:
        lib().bsp_csr_spmv_noalloc(C.byref(d), 8, 8, None)
    assert e.value.status == 6 and "2^31" in str(e.value)
    d = CsrDesc(10, 10, 5, 8, 8, 8, 1)  # FP32 values: FP64 only
    with pytest.raises(BspError) as e:
        lib().bsp_csr_spmv_noalloc(C.byref(d), 8, 8, None)
    assert e.value.status == 1


def test_gather_probe_grid_query_and_argument_checks():
    # sink == NULL: the workgroup count only, no device access; bad slice
    # counts and non-power-of-two slices are argument errors
    from benchsp_amd._lib import BspError, lib
    g = C.c_int64()
    lib().bsp_gather_probe(None, 1 << 20, 4, 20_985_088, None, 0, C.byref(g), None)
    assert g.value == -(-(-(-20_985_088 // 32)) // 256)
    for n, slices in ((1 << 20, 3), (3 << 18, 4), (0, 1)):
        with pytest.raises(BspError) as e:
            lib().bsp_gather_probe(None, n, slices, 1000, None, 0, C.byref(g), None)
        assert e.value.status == 1


def test_line_probe_grid_query_and_argument_checks():
    # 16 bytes per lane, 32 runs per lane, 256 lanes per workgroup; slices
    # shorter than one 1 KB run are argument errors
    from benchsp_amd._lib import BspError, lib
    g = C.c_int64()
    lib().bsp_line_probe(None, 1 << 20, 4, 256 << 20, None, 0, C.byref(g), None)
    assert g.value == (256 << 20) // (16 * 32 * 256)
    for n, slices in ((1 << 20, 3), (3 << 18, 4), (0, 1), (256, 4)):
        with pytest.raises(BspError) as e:
            lib().bsp_line_probe(None, n, slices, 1000, None, 0, C.byref(g), None)
        assert e.value.status == 1


def test_rocsparse_comparator_rejects_non_csr_algorithms():
    # rocsparse_spmv_alg codes other than the CSR ones (1 COO, 4 ELL, ...)
    # are argument errors, checked before any device work
    from benchsp_amd._lib import BspError, CsrDesc, lib
    d = CsrDesc(0, 0, 0, None, None, None, 0)
    h = C.c_void_p()
    for alg in (1, 4, 5, 6, 8, 9, -1):
        with pytest.raises(BspError) as e:
            lib().bsp_rocsparse_plan_create(C.byref(d), alg, None, C.byref(h))
        assert e.value.status == 1, alg

"""ctypes front end of the CPU oracle (oracle.c) and of the reference's own
DBSR builder compiled from its header (_ref/libdbsr_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker. The library never uses it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(HERE, "liboracle.so")
_REF = os.path.join(HERE, "_ref", "libdbsr_ref.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")

_lib = None
_ref = None


def build():
    """Compile liboracle.so (and _ref when the reference tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_spmv_f64.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p]
        L.orc_spmv_f64_par.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.orc_spmv_f32.argtypes = [C.c_int64, _i64p, _i32p, _f32p, _f32p, _f32p]
        L.orc_spmv_f16.argtypes = [C.c_int64, _i64p, _i32p, _f32p, _f32p, _f32p]
        L.orc_f32_to_f16.argtypes = [C.c_float]
        L.orc_f32_to_f16.restype = C.c_uint16
        L.orc_f16_to_f32.argtypes = [C.c_uint16]
        L.orc_f16_to_f32.restype = C.c_float
        L.orc_split.argtypes = [C.c_int64, _i64p, C.c_int64, _i32p, C.POINTER(C.c_int64), _i32p,
                                C.POINTER(C.c_int64)]
        L.orc_bcsr_build.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p, C.c_int, C.c_int]
        L.orc_bcsr_build.restype = C.POINTER(_Bcsr)
        L.orc_bcsr_free.argtypes = [C.POINTER(_Bcsr)]
        L.orc_dbsr_build.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p]
        L.orc_dbsr_build.restype = C.POINTER(_Dbsr)
        L.orc_dbsr_build_ex.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p, C.c_int]
        L.orc_dbsr_build_ex.restype = C.POINTER(_Dbsr)
        L.orc_dbsr_free.argtypes = [C.POINTER(_Dbsr)]
        L.orc_dasp_build.argtypes = [C.c_int32, _i32p, _i32p, _f64p, C.c_double, C.c_int32]
        L.orc_dasp_build.restype = C.POINTER(_Dasp)
        L.orc_dasp_free.argtypes = [C.POINTER(_Dasp)]
        L.orc_blockinfo.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, C.c_int32,
                                    C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        _lib = L
    return _lib


class _Bcsr(C.Structure):
    _fields_ = [("brows", C.c_int64), ("bcols", C.c_int64), ("nblocks", C.c_int64),
                ("tm", C.c_int32), ("tk", C.c_int32),
                ("rowptr", C.POINTER(C.c_int32)), ("colidx", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double)), ("map", C.POINTER(C.c_uint16)),
                ("warpnum", C.c_int32), ("rowptr_by_warp", C.POINTER(C.c_int32)),
                ("rowidx_by_warp", C.POINTER(C.c_int32)),
                ("avg_nnz", C.c_double), ("stand", C.c_double)]


class _Dbsr(C.Structure):
    _fields_ = [("nrows", C.c_int64), ("ncols", C.c_int64), ("nnz", C.c_int64),
                ("brows", C.c_int64), ("nblocks", C.c_int64),
                ("recorder", C.POINTER(C.c_int32)), ("blcptr", C.POINTER(C.c_int32)),
                ("colidx", C.POINTER(C.c_int32)), ("val", C.POINTER(C.c_double)),
                ("map", C.POINTER(C.c_uint16)), ("warpnum", C.c_int32),
                ("prefix", C.POINTER(C.c_int32)), ("overflow", C.c_int32)]


_DASP_INTS = ["row_long", "row_block", "row_zero", "short_row_1", "short_row_2", "short_row_3",
              "short_row_4", "common_13", "fill0_nnz_short13", "fill0_nnz_short34",
              "fill0_nnz_short22", "fill0_nnz_short", "threadblock13", "threadblock34",
              "threadblock22", "blocknum", "fill0_nnz_reg", "nnz_irreg", "warp_number",
              "fill0_nnz_long", "nnz_long"]


class _Dasp(C.Structure):
    _fields_ = [(k, C.c_int32) for k in _DASP_INTS] + [
        ("order_rid", C.POINTER(C.c_int32)), ("short_val", C.POINTER(C.c_double)),
        ("short_cid", C.POINTER(C.c_int32)), ("long_rpt", C.POINTER(C.c_int32)),
        ("long_val", C.POINTER(C.c_double)), ("long_cid", C.POINTER(C.c_int32)),
        ("rid_by_warp", C.POINTER(C.c_int32)), ("block_ptr", C.POINTER(C.c_int32)),
        ("irreg_rpt", C.POINTER(C.c_int32)), ("reg_val", C.POINTER(C.c_double)),
        ("reg_cid", C.POINTER(C.c_int32)), ("irreg_val", C.POINTER(C.c_double)),
        ("irreg_cid", C.POINTER(C.c_int32))]


def _arr(ptr, n, dtype):
    if n <= 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(int(n),)).astype(dtype, copy=True)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _cf64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# SpMV (src/ref.cpp)

def spmv_f64(rowptr, col, val, x, threads: int = 1):
    """y = A*x by the reference CPU loop (ref.cpp:26-37)."""
    rp = _c64(rowptr)
    n = rp.shape[0] - 1
    y = np.zeros(n, dtype=np.float64)
    if threads > 1:
        lib().orc_spmv_f64_par(n, rp, _c32(col), _cf64(val), _cf64(x), y, int(threads))
    else:
        lib().orc_spmv_f64(n, rp, _c32(col), _cf64(val), _cf64(x), y)
    return y


def spmv_f32(rowptr, col, val, x):
    rp = _c64(rowptr)
    y = np.zeros(rp.shape[0] - 1, dtype=np.float32)
    lib().orc_spmv_f32(rp.shape[0] - 1, rp, _c32(col), np.ascontiguousarray(val, np.float32),
                       np.ascontiguousarray(x, np.float32), y)
    return y


def spmv_f16(rowptr, col, val, x):
    rp = _c64(rowptr)
    y = np.zeros(rp.shape[0] - 1, dtype=np.float32)
    lib().orc_spmv_f16(rp.shape[0] - 1, rp, _c32(col), np.ascontiguousarray(val, np.float32),
                       np.ascontiguousarray(x, np.float32), y)
    return y


def f32_to_f16_bits(v: float) -> int:
    return int(lib().orc_f32_to_f16(float(v)))


def f16_bits_to_f32(h: int) -> float:
    return float(lib().orc_f16_to_f32(int(h)))


# --------------------------------------------------------------------------
# Builders

def blockinfo(nrows, ncols, rowptr, col, b: int = 4):
    """(blocks, packed blocks) of bench/blockinfo.cpp:96-119."""
    bl, pk = C.c_int64(), C.c_int64()
    lib().orc_blockinfo(int(nrows), int(ncols), _c64(rowptr), _c32(col), int(b), C.byref(bl),
                        C.byref(pk))
    return bl.value, pk.value


def split(rowptr, threshold: int = 128):
    """Short/long row maps (src/spmv.cu:131-196)."""
    rp = _c64(rowptr)
    n = rp.shape[0] - 1
    sm = np.zeros(max(n, 1), np.int32)
    lm = np.zeros(max(n, 1), np.int32)
    ns, nl = C.c_int64(), C.c_int64()
    lib().orc_split(n, rp, int(threshold), sm, C.byref(ns), lm, C.byref(nl))
    return sm[: ns.value].copy(), lm[: nl.value].copy()


def bcsr(nrows, ncols, rowptr, col, val, tm: int, tk: int):
    """BCSR tm x tk (utils/utils.hpp:9-124) + AmgT extras for 4x4."""
    p = lib().orc_bcsr_build(int(nrows), int(ncols), _c64(rowptr), _c32(col), _cf64(val), tm, tk)
    b = p.contents
    nb = b.nblocks
    out = dict(brows=b.brows, bcols=b.bcols, nblocks=nb, tm=tm, tk=tk,
               rowptr=_arr(b.rowptr, b.brows + 1, np.int32),
               colidx=_arr(b.colidx, nb, np.int32),
               val=_arr(b.val, nb * tm * tk, np.float64),
               warpnum=b.warpnum,
               rowptr_by_warp=_arr(b.rowptr_by_warp, b.brows + 1, np.int32),
               rowidx_by_warp=_arr(b.rowidx_by_warp, b.warpnum, np.int32),
               avg_nnz=b.avg_nnz, stand=b.stand)
    if tm == 4 and tk == 4:
        out["map"] = _arr(b.map, nb + 1, np.uint16)
    lib().orc_bcsr_free(p)
    return out


def dbsr(nrows, ncols, rowptr, col, val, sort_rows: bool = True):
    """DBSR with recorder and balance (include/dbsr.h:10-253). sort_rows=False
    is csrTodbsr + balanceDbsr on the rows as given (recorder = identity)."""
    p = lib().orc_dbsr_build_ex(int(nrows), int(ncols), _c64(rowptr), _c32(col), _cf64(val),
                                int(bool(sort_rows)))
    d = p.contents
    nb = d.nblocks
    out = dict(brows=d.brows, nblocks=nb,
               recorder=_arr(d.recorder, nrows, np.int32),
               blcptr=_arr(d.blcptr, d.brows + 1, np.int32),
               colidx=_arr(d.colidx, nb * 16, np.int32),
               val=_arr(d.val, nb * 16, np.float64),
               map=_arr(d.map, nb, np.uint16),
               warpnum=d.warpnum, overflow=bool(d.overflow),
               prefix=_arr(d.prefix, d.warpnum + 1, np.int32))
    lib().orc_dbsr_free(p)
    return out


def dasp(rowptr, col, val, threshold: float = 0.75, block_longest: int = 256):
    """DASP layout (src/dasp.cu:814-1503)."""
    rp = _c32(rowptr)
    n = rp.shape[0] - 1
    p = lib().orc_dasp_build(n, rp, _c32(col), _cf64(val), float(threshold), int(block_longest))
    D = p.contents
    out = {k: getattr(D, k) for k in _DASP_INTS}
    out.update(
        order_rid=_arr(D.order_rid, n, np.int32),
        short_val=_arr(D.short_val, D.fill0_nnz_short, np.float64),
        short_cid=_arr(D.short_cid, D.fill0_nnz_short, np.int32),
        long_rpt=_arr(D.long_rpt, D.row_long + 1, np.int32),
        long_val=_arr(D.long_val, D.fill0_nnz_long, np.float64),
        long_cid=_arr(D.long_cid, D.fill0_nnz_long, np.int32),
        rid_by_warp=_arr(D.rid_by_warp, D.warp_number, np.int32),
        block_ptr=_arr(D.block_ptr, D.blocknum + 1, np.int32),
        irreg_rpt=_arr(D.irreg_rpt, D.row_block + 1, np.int32),
        reg_val=_arr(D.reg_val, D.fill0_nnz_reg, np.float64),
        reg_cid=_arr(D.reg_cid, D.fill0_nnz_reg, np.int32),
        irreg_val=_arr(D.irreg_val, D.nnz_irreg, np.float64),
        irreg_cid=_arr(D.irreg_cid, D.nnz_irreg, np.int32))
    lib().orc_dasp_free(p)
    return out


def build_seconds(kind: str, nrows, ncols, rowptr, col, val) -> float:
    """Wall time of one restated host builder ("bsr4", "dbsr" or "dasp"),
    the C call alone (inputs converted beforehand, nothing copied out)."""
    import time
    rp64, c32, v = _c64(rowptr), _c32(col), _cf64(val)
    rp32 = _c32(rowptr)
    L = lib()
    t0 = time.perf_counter()
    if kind == "bsr4":
        L.orc_bcsr_free(L.orc_bcsr_build(int(nrows), int(ncols), rp64, c32, v, 4, 4))
    elif kind == "dbsr":
        L.orc_dbsr_free(L.orc_dbsr_build(int(nrows), int(ncols), rp64, c32, v))
    elif kind == "dasp":
        L.orc_dasp_free(L.orc_dasp_build(int(nrows), rp32, c32, v, 0.75, 256))
    else:
        raise ValueError(kind)
    return time.perf_counter() - t0


# --------------------------------------------------------------------------
# The reference's own DBSR builder (compiled from include/dbsr.h)

class _RefDbsr(C.Structure):
    _fields_ = [("brows", C.c_int32), ("nblocks", C.c_int32), ("warpnum", C.c_int32),
                ("blocknnz", C.c_int32),
                ("blcPtr", C.POINTER(C.c_int32)), ("colIdx", C.POINTER(C.c_int32)),
                ("recorderMap", C.POINTER(C.c_int32)), ("prefix", C.POINTER(C.c_int32)),
                ("blcVal", C.POINTER(C.c_double)), ("blcMap", C.POINTER(C.c_uint16))]


def ref_available() -> bool:
    return os.path.exists(_REF)


def _ref_lib():
    global _ref
    if _ref is None:
        _ref = C.CDLL(_REF)
        for name in ("ref_csr_to_dbsr", "ref_csr_to_dbsr_plain"):
            getattr(_ref, name).argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p, _f64p,
                                            C.POINTER(_RefDbsr)]
        pi, pd = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        _ref.ref_csr_recorder_self.argtypes = [C.c_int32, C.c_int32, C.c_int32, _i32p, _i32p,
                                               _f64p, C.POINTER(pi), C.POINTER(pi),
                                               C.POINTER(pd)]
    return _ref


def ref_dbsr(nrows, ncols, rowptr, col, val, sort_rows: bool = True):
    """csrTodbsrWithRecorder from the reference header, or (sort_rows=False)
    its csrTodbsr followed by balanceDbsr (small inputs only: the arrays it
    allocates are not freed)."""
    rp, ci, v = _c32(rowptr), _c32(col), _cf64(val)
    o = _RefDbsr()
    fn = _ref_lib().ref_csr_to_dbsr if sort_rows else _ref_lib().ref_csr_to_dbsr_plain
    fn(int(nrows), int(ncols), int(rp[-1]), rp, ci, v, C.byref(o))
    nb = o.nblocks
    rec = (_arr(o.recorderMap, nrows, np.int32) if sort_rows
           else np.arange(nrows, dtype=np.int32))
    return dict(brows=o.brows, nblocks=nb, warpnum=o.warpnum,
                recorder=rec,
                blcptr=_arr(o.blcPtr, o.brows + 1, np.int32),
                colidx=_arr(o.colIdx, nb * 16, np.int32),
                val=_arr(o.blcVal, nb * 16, np.float64),
                map=_arr(o.blcMap, nb, np.uint16),
                prefix=_arr(o.prefix, o.warpnum + 1, np.int32))


def ref_recorder_self(nrows, ncols, rowptr, col, val):
    """csrRecorderSelf from the reference header: (rowptr, col, val) of the
    CSR with rows sorted by (length, index)."""
    rp, ci, v = _c32(rowptr), _c32(col), _cf64(val)
    orp, oci = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
    ov = C.POINTER(C.c_double)()
    _ref_lib().ref_csr_recorder_self(int(nrows), int(ncols), int(rp[-1]), rp, ci, v,
                                     C.byref(orp), C.byref(oci), C.byref(ov))
    nnz = int(rp[-1])
    return (_arr(orp, nrows + 1, np.int32), _arr(oci, nnz, np.int32), _arr(ov, nnz, np.float64))


# --------------------------------------------------------------------------
# FP64 SpMM / SDDMM / block SpMM

def _spmm_lib():
    L = lib()
    if not getattr(L, "_spmm_declared", False):
        L.orc_csr_spmm_f64.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int64, _f64p]
        L.orc_csr_spmm_t_f64.argtypes = [C.c_int64, C.c_int64, _i64p, _i32p, _f64p, _f64p,
                                         C.c_int64, _f64p]
        L.orc_sddmm_f64.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int64, _f64p]
        L.orc_bsr_spmm_f64.argtypes = [C.c_int64, C.c_int32, _i32p, _i32p, _f64p, _f64p,
                                       C.c_int64, _f64p]
        L._spmm_declared = True
    return L


def csr_spmm_f64(rowptr, col, val, X):
    """Y = A X in FP64 (X row-major, nrows(X) = ncols(A))."""
    rp = _c64(rowptr)
    X = np.ascontiguousarray(X, np.float64)
    Y = np.zeros((rp.shape[0] - 1, X.shape[1]))
    _spmm_lib().orc_csr_spmm_f64(rp.shape[0] - 1, rp, _c32(col), _cf64(val), X, X.shape[1], Y)
    return Y


def csr_spmm_t_f64(rowptr, col, val, dY, ncols):
    """dX = A^T dY in FP64."""
    rp = _c64(rowptr)
    dY = np.ascontiguousarray(dY, np.float64)
    dX = np.zeros((int(ncols), dY.shape[1]))
    _spmm_lib().orc_csr_spmm_t_f64(rp.shape[0] - 1, int(ncols), rp, _c32(col), _cf64(val), dY,
                                   dY.shape[1], dX)
    return dX


def sddmm_f64(rowptr, col, dY, X):
    """dvals[j] = dY[row_j] . X[col_j] in FP64."""
    rp = _c64(rowptr)
    dY = np.ascontiguousarray(dY, np.float64)
    X = np.ascontiguousarray(X, np.float64)
    out = np.zeros(int(rp[-1]))
    _spmm_lib().orc_sddmm_f64(rp.shape[0] - 1, rp, _c32(col), dY, X, X.shape[1], out)
    return out


def bsr_spmm_f64(browptr, bcol, bval, bs, B):
    """C = A B for A in BSR (bs x bs row-major blocks), FP64."""
    bp = _c32(browptr)
    B = np.ascontiguousarray(B, np.float64)
    Cm = np.zeros(((bp.shape[0] - 1) * bs, B.shape[1]))
    _spmm_lib().orc_bsr_spmm_f64(bp.shape[0] - 1, int(bs), bp, _c32(bcol), _cf64(bval), B,
                                 B.shape[1], Cm)
    return Cm

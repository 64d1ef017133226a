"""Loads libbsp.so (the HIP library) and declares its C-ABI for ctypes.

There is no fallback: if the library is missing or fails to load, importing
the package's compute modules raises, naming the build command.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbsp.so")

vp = C.c_void_p
i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
i32p, i64p, f64p = C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_double)


class BspError(RuntimeError):
    """A libbsp call returned a non-zero bsp_status."""

    def __init__(self, fn: str, status: int, msg: str):
        super().__init__(f"{fn} failed: status {status} ({msg})")
        self.status = status


class CsrOptions(C.Structure):
    """struct bsp_csr_options (include/bsp.h)."""
    _fields_ = [("seg_cols", i64), ("chunk", i32), ("seg_mode", i32), ("narrow_cols", i32),
                ("reserved32", i32), ("reserved", i64 * 5)]


class CsrDesc(C.Structure):
    """struct bsp_csr (include/bsp.h)."""
    _fields_ = [("nrows", i64), ("ncols", i64), ("nnz", i64), ("rowptr", vp), ("colidx", vp),
                ("val", vp), ("dtype", C.c_int)]


class CSRFormatF64(C.Structure):
    """struct bsp_CSRFormat_f64: CSRFormat<double, int32_t> (include/matrixFormat.hpp)."""
    _fields_ = [("rowPtr", vp), ("colIdx", vp), ("values", vp), ("row", i32), ("col", i32),
                ("nnz", i32)]


class DbsrMat(C.Structure):
    """struct bsp_dbsrMat: dbsrMat (include/matrixFormat.hpp:86-106)."""
    _fields_ = [("row", C.c_int), ("col", C.c_int), ("nnz", C.c_int), ("blc_row", C.c_int),
                ("blc_col", C.c_int), ("blc_num", C.c_int), ("stand", f64), ("avg_nnz", f64),
                ("blcPtr", vp), ("colIdx", vp), ("blcVal", vp), ("blcMap", vp),
                ("blocknnz", i32), ("warpnum", i32), ("rowPtrbyWarp", vp), ("rowIdxbyWarp", vp),
                ("recorderMap", vp), ("prefix", vp)]


class BsrView(C.Structure):
    """struct bsp_bsr_view."""
    _fields_ = [("nrows", i64), ("ncols", i64), ("nnz", i64), ("brows", i64), ("bcols", i64),
                ("nblocks", i64), ("tm", i32), ("tk", i32), ("rowptr", vp), ("colidx", vp),
                ("val", vp), ("map", vp), ("warpnum", i32), ("rowptr_by_warp", vp),
                ("rowidx_by_warp", vp), ("avg_nnz", f64), ("stand", f64)]


class DbsrView(C.Structure):
    """struct bsp_dbsr_view."""
    _fields_ = [("nrows", i64), ("ncols", i64), ("nnz", i64), ("brows", i64), ("nblocks", i64),
                ("recorder", vp), ("blcptr", vp), ("colidx", vp), ("val", vp), ("map", vp),
                ("warpnum", i32), ("prefix", vp), ("overflow", i32)]


_DASP_I32 = ["row_long", "row_block", "row_zero", "short_row_1", "short_row_2", "short_row_3",
             "short_row_4", "common_13", "fill0_nnz_short13", "fill0_nnz_short34",
             "fill0_nnz_short22", "fill0_nnz_short", "threadblock13", "threadblock34",
             "threadblock22", "blocknum", "fill0_nnz_reg", "nnz_irreg", "warp_number",
             "fill0_nnz_long", "nnz_long"]


class DaspView(C.Structure):
    """struct bsp_dasp_view."""
    _fields_ = ([("nrows", i64), ("ncols", i64), ("nnz", i64)] + [(k, i32) for k in _DASP_I32] +
                [(k, vp) for k in ("order_rid", "short_val", "short_cid", "long_rpt", "long_val",
                                   "long_cid", "rid_by_warp", "block_ptr", "irreg_rpt", "reg_val",
                                   "reg_cid", "irreg_val", "irreg_cid")])


class BlockInfo(C.Structure):
    """struct bsp_blockinfo."""
    _fields_ = [("nnz", i64), ("blocks", i64), ("blocks_packed", i64), ("avg", f64),
                ("avg_packed", f64)]


class TuneRecord(C.Structure):
    """struct bsp_tune_record."""
    _fields_ = [("format", i32), ("kernel", i32), ("chunk", i32), ("block_longest", i32),
                ("seg_cols", i64), ("threshold", f64), ("time_us", f64), ("reserved", i64 * 4)]


class TuneOptions(C.Structure):
    """struct bsp_tune_options."""
    _fields_ = [("formats", i32), ("reps", i32), ("reserved", i64 * 7)]


VOID = "void"  # restype marker for the reference's void functions

# name -> (restype, argtypes); restype None means bsp_status (checked).
_SIGS: dict[str, tuple] = {
    "bsp_csr_spmv_f16": (None, [vp, vp, vp, vp]),
    "bsp_rocsparse_spmm_plan_create": (None, [C.POINTER(CsrDesc), i64, C.c_int, C.c_int, vp,
                                              C.POINTER(vp)]),
    "bsp_rocsparse_spmm": (None, [vp, vp, vp, vp]),
    "bsp_rocsparse_spmm_plan_destroy": (None, [vp]),
    "bsp_mtx_read": (None, [C.c_char_p, C.POINTER(vp)]),
    "bsp_coo_to_csr": (None, [i64, i64, i64, vp, vp, vp, C.POINTER(vp)]),
    "bsp_block_info": (None, [C.POINTER(CsrDesc), C.c_int, vp, C.POINTER(BlockInfo)]),
    "bsp_batch_create": (None, [vp, i64, vp, vp, C.POINTER(vp)]),
    "bsp_batch_layout": (None, [vp, vp, vp]),
    "bsp_batch_choice": (None, [vp, vp, vp, vp]),
    "bsp_batch_spmm_format": (None, [vp, vp]),
    "bsp_batch_info": (None, [vp, i64p, i64p, i64p]),
    "bsp_batch_run": (None, [vp, C.c_int, vp, vp, vp, vp, vp]),
    "bsp_batch_graph_capture": (None, [vp, C.c_int, vp, vp, vp, vp]),
    "bsp_batch_graph_launch": (None, [vp, vp]),
    "bsp_batch_destroy": (None, [vp]),
    "bsp_rocsparse_plan_create": (None, [C.POINTER(CsrDesc), C.c_int, vp, C.POINTER(vp)]),
    "bsp_rocsparse_spmv": (None, [vp, vp, vp, vp]),
    "bsp_cusparse_spmv": (C.c_int, [vp, vp, vp, vp, vp, i64, i64, i64]),
    "bsp_cusparse_spmv_fp16": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32]),
    "bsp_cusparse_spmv_fp16_warpper": (VOID, [vp, vp, vp, vp, vp, i32, i32, i32]),
    "bsp_ref_spmv_fp64": (VOID, [vp, vp, vp, vp, vp, i64, i64, f64p]),
    "bsp_ref_spmv_fp32": (VOID, [vp, vp, vp, vp, vp, i32, i32]),
    "bsp_ref_spmv_fp16": (VOID, [vp, vp, vp, vp, vp, i32, i32]),
    "bsp_rocsparse_plan_destroy": (None, [vp]),
    "bsp_dbsr_from_host": (None, [C.POINTER(DbsrMat), vp, C.POINTER(vp)]),
    "bsp_csrTodbsrWithRecorder": (C.c_int, [C.POINTER(CSRFormatF64), C.POINTER(DbsrMat)]),
    "bsp_dbsrMat_free": (VOID, [C.POINTER(DbsrMat)]),
    "bsp_csrTodbsr": (C.c_int, [C.POINTER(CSRFormatF64), C.POINTER(DbsrMat)]),
    "bsp_balanceDbsr": (C.c_int, [C.POINTER(DbsrMat)]),
    "bsp_csrRecorderSelf": (C.c_int, [C.POINTER(CSRFormatF64), C.POINTER(CSRFormatF64)]),
    "bsp_csrToBcsr": (C.c_int, [vp, vp, vp, i64, i64, C.c_int, C.c_int, i64p, C.POINTER(vp),
                                C.POINTER(vp), C.POINTER(vp)]),
    "bsp_free": (VOID, [vp]),
    "bsp_amgT_spmv_fp64_dbsr": (C.c_int, [C.POINTER(DbsrMat), vp, vp, i32, i32, f64p]),
    "bsp_amgT_spmv_fp64_dbsr_balance": (C.c_int, [C.POINTER(DbsrMat), vp, vp, i32, i32, f64p]),
    "bsp_amgT_spmv_fp64": (C.c_int, [vp, vp, vp, vp, vp, i32, i32, i32, C.c_int, f64p]),
    "bsp_bcsr_spmv_fp64": (VOID, [vp, vp, vp, vp, vp, i64, i64, i64, C.c_int]),
    "bsp_spmv_all": (VOID, [C.c_char_p, vp, vp, vp, vp, vp, vp, C.c_int, C.c_int, C.c_int,
                            C.c_int, f64, C.c_int, f64p]),
    "bsp_MySpmvWithPrefp64": (VOID, [vp, vp, i64, i64, i64, vp, vp, vp]),
    "bsp_MySpmvWithPrefp16": (VOID, [vp, vp, i32, i32, i32, vp, vp, vp]),
    "bsp_smatSpmmfp16": (VOID, [vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp, vp]),
    "bsp_cusparse_spmv_fp64": (C.c_int, [vp, vp, vp, vp, vp, i64, i64, i64, C.c_int, f64p]),
    "bsp_bspmm_build": (None, [vp, C.c_int, vp, C.POINTER(vp)]),
    "bsp_bspmm_view_get": (None, [vp, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), i64p]),
    "bsp_bspmm_fwd": (None, [vp, vp, i64, vp, vp]),
    "bsp_bspmm_bwd": (None, [vp, vp, vp, i64, vp, vp, vp]),
    "bsp_bspmm_destroy": (None, [vp]),
    "bsp_spmm_plan_create": (None, [C.POINTER(CsrDesc), vp, C.POINTER(vp)]),
    "bsp_spmm": (None, [vp, vp, C.c_int, vp, C.c_int, i64, vp, vp]),
    "bsp_spmm_t": (None, [vp, vp, C.c_int, vp, C.c_int, i64, vp, vp]),
    "bsp_sddmm": (None, [vp, vp, vp, C.c_int, i64, vp, vp]),
    "bsp_spmm_t_sddmm": (None, [vp, vp, C.c_int, vp, vp, C.c_int, i64, vp, vp, vp]),
    "bsp_spmm_plan_destroy": (None, [vp]),
    "bsp_dasp_build": (None, [C.POINTER(CsrDesc), f64, i32, vp, C.POINTER(vp)]),
    "bsp_dasp_view_get": (None, [vp, C.POINTER(DaspView)]),
    "bsp_dasp_spmv": (None, [vp, vp, vp, C.c_int, vp]),
    "bsp_dasp_spmv_ex": (None, [vp, C.c_int, vp, vp, C.c_int, vp]),
    "bsp_dasp_destroy": (None, [vp]),
    "bsp_dia_build": (None, [C.POINTER(CsrDesc), vp, C.POINTER(vp)]),
    "bsp_dia_info": (None, [vp, vp, vp, vp]),
    "bsp_dia_spmv": (None, [vp, vp, vp, vp]),
    "bsp_dia_destroy": (None, [vp]),
    "bsp_autotune": (None, [C.POINTER(CsrDesc), C.POINTER(TuneOptions), vp, C.POINTER(vp)]),
    "bsp_tuned_from_record": (None, [C.POINTER(CsrDesc), C.POINTER(TuneRecord), vp,
                                     C.POINTER(vp)]),
    "bsp_tuned_record": (None, [vp, C.POINTER(TuneRecord)]),
    "bsp_tuned_trials": (None, [vp, C.POINTER(TuneRecord), i32, C.POINTER(i32)]),
    "bsp_tuned_spmv": (None, [vp, vp, vp, vp]),
    "bsp_tuned_destroy": (None, [vp]),
    "bsp_copy_to_host": (None, [vp, vp, C.c_size_t]),
    "bsp_bsr_build": (None, [C.POINTER(CsrDesc), C.c_int, C.c_int, vp, C.POINTER(vp)]),
    "bsp_bsr_view_get": (None, [vp, C.POINTER(BsrView)]),
    "bsp_bsr_spmv": (None, [vp, C.c_int, vp, vp, vp]),
    "bsp_bsr_destroy": (None, [vp]),
    "bsp_dbsr_build": (None, [C.POINTER(CsrDesc), vp, C.POINTER(vp)]),
    "bsp_dbsr_view_get": (None, [vp, C.POINTER(DbsrView)]),
    "bsp_dbsr_spmv": (None, [vp, vp, vp, C.c_int, vp]),
    "bsp_dbsr_spmv_ex": (None, [vp, C.c_int, vp, vp, C.c_int, vp]),
    "bsp_dbsr_destroy": (None, [vp]),
    "bsp_status_string": (C.c_char_p, [C.c_int]),
    "bsp_last_error": (C.c_char_p, []),
    "bsp_version": (C.c_char_p, []),
    "bsp_csr_plan_create": (None, [C.POINTER(CsrDesc), C.c_int, vp, C.POINTER(vp)]),
    "bsp_csr_plan_create_ex": (None, [C.POINTER(CsrDesc), C.c_int, C.c_void_p, vp,
                                       C.POINTER(vp)]),
    "bsp_csr_plan_segments": (None, [vp, i64p, i64p]),
    "bsp_csr_plan_narrow": (None, [vp, C.POINTER(i32)]),
    "bsp_csr_check": (None, [C.POINTER(CsrDesc), vp, C.POINTER(i32)]),
    "bsp_csr_plan_probe": (None, [vp, C.c_int, vp, vp, i64, i64p, vp]),
    "bsp_gather_probe": (None, [vp, i64, C.c_int32, i64, vp, i64, i64p, vp]),
    "bsp_line_probe": (None, [vp, i64, C.c_int32, i64, vp, i64, i64p, vp]),
    "bsp_csr_spmv": (None, [vp, vp, vp, vp]),
    "bsp_csr_plan_info": (None, [vp, i64p, i64p]),
    "bsp_plan_destroy": (None, [vp]),
    "bsp_csr_spmv_noalloc": (None, [C.POINTER(CsrDesc), vp, vp, vp]),
    "bsp_gen_uniform": (None, [i64, u64, vp]),
    "bsp_gen_normal": (None, [i64, u64, vp]),
    "bsp_gen_stencil5": (None, [i64, i64, C.POINTER(vp)]),
    "bsp_gen_stencil3d": (None, [i64, i64, i64, C.c_int, u64, C.POINTER(vp)]),
    "bsp_gen_powerlaw": (None, [i64, i64, f64, f64, i64, u64, C.POINTER(vp)]),
    "bsp_gen_blocksparse": (None, [i64, i64, C.c_int, i64, u64, C.POINTER(vp)]),
    "bsp_gen_graph": (None, [i64, f64, f64, C.c_int, u64, C.POINTER(vp)]),
    "bsp_host_csr_info": (None, [vp, i64p, i64p, i64p]),
    "bsp_host_csr_copy": (None, [vp, vp, vp, vp]),
    "bsp_host_csr_free": (C.c_int, [vp]),
    "bsp_spmv_fp64_warpper": (C.c_int, [vp, vp, i32, i32, i32, vp, vp, vp, C.c_int, f64p]),
    "bsp_spmv_fp64_balance_warpper": (C.c_int, [vp, vp, i32, i32, i32, vp, vp, vp, C.c_int, f64p]),
}

_lib = None


class _Checked:
    """Wraps a bsp_status-returning function: raises BspError on failure."""

    def __init__(self, lib, name, fn):
        self._lib, self._name, self._fn = lib, name, fn

    def __call__(self, *args):
        st = self._fn(*args)
        if st != 0:
            msg = self._lib.bsp_last_error().decode(errors="replace")
            raise BspError(self._name, st, msg)
        return st


class _Lib:
    def __init__(self, cdll):
        self.cdll = cdll
        for name, (res, args) in _SIGS.items():
            fn = getattr(cdll, name)
            fn.argtypes = args
            fn.restype = C.c_int if res is None else (None if res is VOID else res)
            setattr(self, name, _Checked(cdll, name, fn) if res is None else fn)

    def declare(self, name, restype, argtypes):
        """Declares an extra symbol (restype None = checked bsp_status)."""
        fn = getattr(self.cdll, name)
        fn.argtypes = argtypes
        fn.restype = C.c_int if restype is None else restype
        setattr(self, name, _Checked(self.cdll, name, fn) if restype is None else fn)
        return getattr(self, name)


def lib() -> _Lib:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m benchsp_amd.build` "
                              "(or __graft_entry__.build())")
        # torch first, so its HIP runtime is the one libbsp binds to.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        _lib = _Lib(C.CDLL(LIB_PATH))
    return _lib


def exported_symbols() -> list[str]:
    """Every bsp_* name this module declares."""
    return sorted(_SIGS)

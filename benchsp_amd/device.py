"""Device-resident matrices and the CSR SpMV plan, over the C-ABI.

torch supplies device memory and the current HIP stream; every computation
runs in libbsp.so kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import CsrDesc, CsrOptions, i64, lib
from .host import Csr

F64, F32, F16, BF16 = 0, 1, 2, 3
_DT = {torch.float64: F64, torch.float32: F32, torch.float16: F16, torch.bfloat16: BF16}

ADAPTIVE, SCALAR, BALANCED, SPLIT = 0, 1, 2, 3
METHODS = {"adaptive": ADAPTIVE, "scalar": SCALAR, "balanced": BALANCED, "split": SPLIT}


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def ptr(t: torch.Tensor | None) -> int:
    return 0 if t is None else t.data_ptr()


class DeviceCsr:
    """CSR with int32 indices in device memory."""

    def __init__(self, rowptr: torch.Tensor, col: torch.Tensor, val: torch.Tensor, ncols: int):
        assert rowptr.dtype == torch.int32 and col.dtype == torch.int32
        self.rowptr, self.col, self.val = rowptr.contiguous(), col.contiguous(), val.contiguous()
        self.nrows = rowptr.numel() - 1
        self.ncols = int(ncols)
        self.nnz = col.numel()

    @classmethod
    def from_host(cls, A: Csr, device="cuda", dtype=torch.float64) -> "DeviceCsr":
        if A.nnz >= 2**31:
            raise ValueError("nnz must fit int32 offsets")
        rp = torch.from_numpy(A.rowptr.astype(np.int32)).to(device)
        col = torch.from_numpy(A.col).to(device)
        val = torch.from_numpy(A.val).to(device=device, dtype=dtype)
        return cls(rp, col, val, A.ncols)

    def desc(self) -> CsrDesc:
        return CsrDesc(self.nrows, self.ncols, self.nnz, ptr(self.rowptr), ptr(self.col),
                       ptr(self.val), _DT[self.val.dtype])


class CsrSpmv:
    """y = A*x with a per-matrix schedule built on the GPU (bsp_csr_plan_create)."""

    def __init__(self, A: DeviceCsr, method: str | int = "adaptive", seg_cols: int = 0,
                 chunk: int = 0, seg_mode: int = 0, narrow_cols: int = 0):
        self.A = A
        self.method = METHODS[method] if isinstance(method, str) else int(method)
        self._desc = A.desc()
        self._opt = CsrOptions(seg_cols, chunk, seg_mode, narrow_cols)
        h = C.c_void_p()
        lib().bsp_csr_plan_create_ex(C.byref(self._desc), self.method, C.byref(self._opt),
                                     stream_handle(), C.byref(h))
        self._h = h

    def segments(self) -> tuple[int, int]:
        """(column-segment width, number of segments); width 0 = unsegmented."""
        w, k = i64(), i64()
        lib().bsp_csr_plan_segments(self._h, C.byref(w), C.byref(k))
        return w.value, k.value

    def narrow(self) -> bool:
        """True when the plan reads 16-bit column offsets."""
        v = C.c_int32()
        lib().bsp_csr_plan_narrow(self._h, C.byref(v))
        return bool(v.value)

    def info(self) -> tuple[int, int]:
        nb, nl = i64(), i64()
        lib().bsp_csr_plan_info(self._h, C.byref(nb), C.byref(nl))
        return nb.value, nl.value

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        assert x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.A.ncols
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_csr_spmv(self._h, ptr(x), ptr(y), stream_handle())
        return y

    def close(self):
        if getattr(self, "_h", None):
            lib().bsp_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _to_host(ptr_value, n, dtype) -> np.ndarray:
    out = np.empty(max(int(n), 1), dtype)
    if n > 0 and ptr_value:
        lib().bsp_copy_to_host(out.ctypes.data, ptr_value, int(n) * out.itemsize)
    return out[: int(n)]


class Bsr:
    """Block CSR with tm x tk blocks, built on the GPU (bsp_bsr_build)."""

    VALU, MFMA, MFMA_BALANCED, VALU_BALANCED = 1, 2, 3, 4

    def __init__(self, A: DeviceCsr, tm: int = 4, tk: int = 4):
        from ._lib import BsrView
        self.A = A
        d = A.desc()
        h = C.c_void_p()
        lib().bsp_bsr_build(C.byref(d), tm, tk, stream_handle(), C.byref(h))
        self._h = h
        self.view = BsrView()
        lib().bsp_bsr_view_get(self._h, C.byref(self.view))

    def arrays(self) -> dict:
        v = self.view
        out = dict(brows=v.brows, bcols=v.bcols, nblocks=v.nblocks, tm=v.tm, tk=v.tk,
                   rowptr=_to_host(v.rowptr, v.brows + 1, np.int32),
                   colidx=_to_host(v.colidx, v.nblocks, np.int32),
                   val=_to_host(v.val, v.nblocks * v.tm * v.tk, np.float64),
                   avg_nnz=v.avg_nnz, stand=v.stand, warpnum=v.warpnum)
        if v.map:
            out["map"] = _to_host(v.map, v.nblocks + 1, np.uint16)
            out["rowptr_by_warp"] = _to_host(v.rowptr_by_warp, v.brows + 1, np.int32)
            out["rowidx_by_warp"] = _to_host(v.rowidx_by_warp, v.warpnum, np.int32)
        return out

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None, kernel: int = 1):
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_bsr_spmv(self._h, kernel, ptr(x), ptr(y), stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_bsr_destroy(self._h)
                self._h = None
        except Exception:
            pass


class Dbsr:
    """DBSR (row-sorted, position-packed 4x4 blocks), built on the GPU."""

    def __init__(self, A: DeviceCsr):
        from ._lib import DbsrView
        self.A = A
        d = A.desc()
        h = C.c_void_p()
        lib().bsp_dbsr_build(C.byref(d), stream_handle(), C.byref(h))
        self._h = h
        self.view = DbsrView()
        lib().bsp_dbsr_view_get(self._h, C.byref(self.view))

    def arrays(self) -> dict:
        v = self.view
        return dict(brows=v.brows, nblocks=v.nblocks, warpnum=v.warpnum,
                    overflow=bool(v.overflow),
                    recorder=_to_host(v.recorder, v.nrows, np.int32),
                    blcptr=_to_host(v.blcptr, v.brows + 1, np.int32),
                    colidx=_to_host(v.colidx, v.nblocks * 16, np.int32),
                    val=_to_host(v.val, v.nblocks * 16, np.float64),
                    map=_to_host(v.map, v.nblocks, np.uint16),
                    prefix=_to_host(v.prefix, v.warpnum + 1, np.int32))

    VALU, MFMA = 0, 1

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None, sorted_order=False,
                 kernel: int = 0):
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_dbsr_spmv_ex(self._h, int(kernel), ptr(x), ptr(y), int(sorted_order),
                               stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_dbsr_destroy(self._h)
                self._h = None
        except Exception:
            pass


DASP_INTS = ["row_long", "row_block", "row_zero", "short_row_1", "short_row_2", "short_row_3",
             "short_row_4", "common_13", "fill0_nnz_short13", "fill0_nnz_short34",
             "fill0_nnz_short22", "fill0_nnz_short", "threadblock13", "threadblock34",
             "threadblock22", "blocknum", "fill0_nnz_reg", "nnz_irreg", "warp_number",
             "fill0_nnz_long", "nnz_long"]


class Dasp:
    """DASP layout built on the GPU (bsp_dasp_build)."""

    def __init__(self, A: DeviceCsr, threshold: float = 0.75, block_longest: int = 256):
        from ._lib import DaspView
        self.A = A
        d = A.desc()
        h = C.c_void_p()
        lib().bsp_dasp_build(C.byref(d), float(threshold), int(block_longest), stream_handle(),
                             C.byref(h))
        self._h = h
        self.view = DaspView()
        lib().bsp_dasp_view_get(self._h, C.byref(self.view))

    def arrays(self) -> dict:
        v = self.view
        out = {k: getattr(v, k) for k in DASP_INTS}
        out.update(
            order_rid=_to_host(v.order_rid, v.nrows, np.int32),
            short_val=_to_host(v.short_val, v.fill0_nnz_short, np.float64),
            short_cid=_to_host(v.short_cid, v.fill0_nnz_short, np.int32),
            long_rpt=_to_host(v.long_rpt, v.row_long + 1, np.int32),
            long_val=_to_host(v.long_val, v.fill0_nnz_long, np.float64),
            long_cid=_to_host(v.long_cid, v.fill0_nnz_long, np.int32),
            rid_by_warp=_to_host(v.rid_by_warp, v.warp_number, np.int32),
            block_ptr=_to_host(v.block_ptr, v.blocknum + 1, np.int32),
            irreg_rpt=_to_host(v.irreg_rpt, v.row_block + 1, np.int32),
            reg_val=_to_host(v.reg_val, v.fill0_nnz_reg, np.float64),
            reg_cid=_to_host(v.reg_cid, v.fill0_nnz_reg, np.int32),
            irreg_val=_to_host(v.irreg_val, v.nnz_irreg, np.float64),
            irreg_cid=_to_host(v.irreg_cid, v.nnz_irreg, np.int32))
        return out

    VALU, MFMA = 0, 1

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None, permuted=False,
                 kernel: int = 0):
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_dasp_spmv_ex(self._h, int(kernel), ptr(x), ptr(y), int(permuted),
                               stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_dasp_destroy(self._h)
                self._h = None
        except Exception:
            pass


class Dia:
    """Diagonal layout (bsp_dia_build): every distinct diagonal one contiguous
    run of values, a 32-bit presence mask per row; at most 32 diagonals."""

    def __init__(self, A: DeviceCsr):
        self.A = A
        d = A.desc()
        h = C.c_void_p()
        lib().bsp_dia_build(C.byref(d), stream_handle(), C.byref(h))
        self._h = h

    def info(self) -> dict:
        n = C.c_int32()
        offs = (C.c_int32 * 32)()
        b = C.c_int64()
        lib().bsp_dia_info(self._h, C.byref(n), offs, C.byref(b))
        return {"ndiag": n.value, "offsets": list(offs[:n.value]), "bytes": b.value}

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        assert x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.A.ncols
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_dia_spmv(self._h, ptr(x), ptr(y), stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_dia_destroy(self._h)
                self._h = None
        except Exception:
            pass


TUNE_FORMATS = {"csr": 0, "bsr4": 1, "dbsr": 2, "dasp": 3, "bsr8": 4, "dia": 5}
_TUNE_NAMES = {v: k for k, v in TUNE_FORMATS.items()}


def _record_dict(r) -> dict:
    return {"format": _TUNE_NAMES.get(r.format, r.format), "kernel": r.kernel, "chunk": r.chunk,
            "seg_cols": r.seg_cols, "threshold": r.threshold,
            "block_longest": r.block_longest, "time_us": r.time_us}


class Tuned:
    """The fastest SpMV layout of one FP64 matrix (bsp_autotune), or the layout
    a stored record names (Tuned.from_record, no timing). `record` is a plain
    dict (JSON-serialisable); y comes back in the original row order."""

    def __init__(self, A: DeviceCsr, formats=None, reps: int = 0, _record=None):
        from ._lib import TuneOptions, TuneRecord
        self.A = A
        d = A.desc()
        h = C.c_void_p()
        if _record is None:
            mask = 0 if formats is None else sum(1 << TUNE_FORMATS[f] for f in formats)
            lib().bsp_autotune(C.byref(d), C.byref(TuneOptions(mask, int(reps))),
                               stream_handle(), C.byref(h))
        else:
            r = TuneRecord()
            r.format = TUNE_FORMATS[_record["format"]]
            r.kernel = int(_record.get("kernel", 0))
            r.chunk = int(_record.get("chunk", 0))
            r.seg_cols = int(_record.get("seg_cols", 0))
            r.threshold = float(_record.get("threshold", 0.0))
            r.block_longest = int(_record.get("block_longest", 0))
            lib().bsp_tuned_from_record(C.byref(d), C.byref(r), stream_handle(), C.byref(h))
        self._h = h

    @classmethod
    def from_record(cls, A: DeviceCsr, record: dict) -> "Tuned":
        return cls(A, _record=record)

    @property
    def record(self) -> dict:
        from ._lib import TuneRecord
        r = TuneRecord()
        lib().bsp_tuned_record(self._h, C.byref(r))
        return _record_dict(r)

    def trials(self) -> list[dict]:
        from ._lib import TuneRecord
        n = C.c_int32()
        lib().bsp_tuned_trials(self._h, None, 0, C.byref(n))
        recs = (TuneRecord * max(n.value, 1))()
        lib().bsp_tuned_trials(self._h, recs, n.value, C.byref(n))
        return [_record_dict(recs[i]) for i in range(n.value)]

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        assert x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.A.ncols
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float64, device=x.device)
        lib().bsp_tuned_spmv(self._h, ptr(x), ptr(y), stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_tuned_destroy(self._h)
                self._h = None
        except Exception:
            pass


class BlockSpmm:
    """Block-sparse SpMM on FP16/BF16 MFMA over a 16x16 Bsr (bsp_bspmm_*)."""

    def __init__(self, bsr: "Bsr", half: torch.dtype = torch.float16):
        assert bsr.view.tm == 16 and bsr.view.tk == 16
        self.bsr = bsr
        self.half = half
        self.M, self.K = bsr.view.nrows, bsr.view.ncols
        self.nblocks = bsr.view.nblocks
        h = C.c_void_p()
        lib().bsp_bspmm_build(bsr._h, BF16 if half == torch.bfloat16 else F16, stream_handle(),
                              C.byref(h))
        self._h = h

    def arrays(self) -> dict:
        sc, hv, hvt, nb = C.c_void_p(), C.c_void_p(), C.c_void_p(), i64()
        lib().bsp_bspmm_view_get(self._h, C.byref(sc), C.byref(hv), C.byref(hvt), C.byref(nb))
        n = nb.value
        return dict(scale=_to_host(sc.value, n, np.float32), hv=_to_host(hv.value, n * 256, np.uint16),
                    hvt=_to_host(hvt.value, n * 256, np.uint16))

    def fwd(self, B: torch.Tensor) -> torch.Tensor:
        assert B.dtype == self.half and B.is_contiguous() and B.shape[0] == self.K
        Cm = torch.empty(self.M, B.shape[1], dtype=torch.float32, device=B.device)
        lib().bsp_bspmm_fwd(self._h, ptr(B), B.shape[1], ptr(Cm), stream_handle())
        return Cm

    def bwd(self, B: torch.Tensor, dC: torch.Tensor):
        assert B.dtype == self.half and dC.dtype == self.half
        N = B.shape[1]
        dB = torch.empty(self.K, N, dtype=torch.float32, device=B.device)
        dv = torch.empty(self.nblocks, 16, 16, dtype=torch.float32, device=B.device)
        lib().bsp_bspmm_bwd(self._h, ptr(B), ptr(dC), N, ptr(dB), ptr(dv), stream_handle())
        return dB, dv

    def __del__(self):
        try:
            if self._h:
                lib().bsp_bspmm_destroy(self._h)
                self._h = None
        except Exception:
            pass


def block_info(A: DeviceCsr, block_len: int = 4) -> dict:
    """b x b block counts with and without left-packing of each row
    (bsp_block_info)."""
    from ._lib import BlockInfo
    out = BlockInfo()
    d = A.desc()
    lib().bsp_block_info(C.byref(d), block_len, stream_handle(), C.byref(out))
    return {k: getattr(out, k) for k, _ in BlockInfo._fields_}


class BatchOptions(C.Structure):
    """struct bsp_batch_options."""
    _fields_ = [("tune", C.c_int32), ("grid", C.c_int32), ("reserved", i64 * 7)]


PERSISTENT, PLAIN = 0, 1


class Batch:
    """Many device CSRs run together (bsp_batch_*): per step an FP64 SpMV and
    an 8-column FP16 SpMM (FP32 accumulation) for every matrix, over
    concatenated vectors laid out by `row_off` / `col_off`."""

    COLS = 8

    def __init__(self, mats: list[DeviceCsr], tune: int = 1, grid: int = 0):
        self.mats = mats  # keeps the borrowed device arrays alive
        descs = (CsrDesc * len(mats))(*[m.desc() for m in mats])
        opt = BatchOptions(int(tune), int(grid))
        h = C.c_void_p()
        lib().bsp_batch_create(descs, len(mats), C.byref(opt), stream_handle(), C.byref(h))
        self._h = h
        M = len(mats)
        ro, co = np.zeros(M + 1, np.int64), np.zeros(M + 1, np.int64)
        lib().bsp_batch_layout(h, ro.ctypes.data, co.ctypes.data)
        self.row_off, self.col_off = ro, co
        self.nrows, self.ncols = int(ro[-1]), int(co[-1])

    CSR, BSR4, BSR8, DASP, DIA = 0, 1, 2, 3, 4  # formats (bsp_batch_choice)

    def choice(self) -> dict:
        M = len(self.mats)
        f, v, m, fm = (np.zeros(M, np.int32) for _ in range(4))
        lib().bsp_batch_choice(self._h, f.ctypes.data, v.ctypes.data, m.ctypes.data)
        lib().bsp_batch_spmm_format(self._h, fm.ctypes.data)
        return dict(format=f, spmv_window=v, spmm_window=m, spmm_format=fm)

    def info(self) -> dict:
        a, b, g = i64(), i64(), i64()
        lib().bsp_batch_info(self._h, C.byref(a), C.byref(b), C.byref(g))
        return dict(spmv_tiles=a.value, spmm_tiles=b.value, grid=g.value)

    def _check(self, x, y, X, Y):
        if x is not None:
            assert x.dtype == torch.float64 and x.is_contiguous() and x.numel() == self.ncols
            assert y.dtype == torch.float64 and y.is_contiguous() and y.numel() == self.nrows
        if X is not None:
            assert X.dtype == torch.float16 and X.is_contiguous()
            assert tuple(X.shape) == (self.ncols, self.COLS) and X.data_ptr() % 16 == 0
            assert Y.dtype == torch.float32 and Y.is_contiguous()
            assert tuple(Y.shape) == (self.nrows, self.COLS) and Y.data_ptr() % 16 == 0

    def run(self, x=None, y=None, X=None, Y=None, schedule: int = PERSISTENT):
        self._check(x, y, X, Y)
        lib().bsp_batch_run(self._h, schedule, ptr(x), ptr(y), ptr(X), ptr(Y), stream_handle())

    def capture(self, x=None, y=None, X=None, Y=None, schedule: int = PERSISTENT):
        self._check(x, y, X, Y)
        lib().bsp_batch_graph_capture(self._h, schedule, ptr(x), ptr(y), ptr(X), ptr(Y))

    def replay(self):
        lib().bsp_batch_graph_launch(self._h, stream_handle())

    def __del__(self):
        try:
            if self._h:
                lib().bsp_batch_destroy(self._h)
                self._h = None
        except Exception:
            pass


class RocsparseSpmv:
    """rocSPARSE SpMV on a device CSR (bsp_rocsparse_*), the vendor comparator."""

    def __init__(self, A: DeviceCsr, alg: int = 0):
        self.A = A
        self._desc = A.desc()
        h = C.c_void_p()
        lib().bsp_rocsparse_plan_create(C.byref(self._desc), alg, stream_handle(), C.byref(h))
        self._h = h

    def __call__(self, x: torch.Tensor, y: torch.Tensor | None = None) -> torch.Tensor:
        if y is None:
            y = torch.empty(self.A.nrows, dtype=torch.float32 if self.A.val.dtype != torch.float64
                            else torch.float64, device=x.device)
        lib().bsp_rocsparse_spmv(self._h, ptr(x), ptr(y), stream_handle())
        return y

    def __del__(self):
        try:
            if self._h:
                lib().bsp_rocsparse_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


class RocsparseSpmm:
    """rocSPARSE C = op(A) B (B row-major in A's value type, C FP32), the
    vendor comparator for the SpMM kernels (bsp_rocsparse_spmm_*)."""

    def __init__(self, A: DeviceCsr, N: int, transpose: bool = False, alg: int = 0):
        self.A, self.N, self.transpose = A, N, transpose
        self._desc = A.desc()
        h = C.c_void_p()
        lib().bsp_rocsparse_spmm_plan_create(C.byref(self._desc), N, int(transpose), alg,
                                             stream_handle(), C.byref(h))
        self._h = h

    def __call__(self, B: torch.Tensor, Cout: torch.Tensor | None = None) -> torch.Tensor:
        rows = self.A.ncols if self.transpose else self.A.nrows
        assert B.dtype == self.A.val.dtype and B.is_contiguous() and B.shape[1] == self.N
        if Cout is None:
            Cout = torch.empty(rows, self.N, dtype=torch.float32, device=B.device)
        lib().bsp_rocsparse_spmm(self._h, ptr(B), ptr(Cout), stream_handle())
        return Cout

    def __del__(self):
        try:
            if self._h:
                lib().bsp_rocsparse_spmm_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass

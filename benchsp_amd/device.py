"""Device-resident matrices and the CSR SpMV plan, over the C-ABI.

torch supplies device memory and the current HIP stream; every computation
runs in libbsp.so kernels.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import CsrDesc, i64, lib
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

    def __init__(self, A: DeviceCsr, method: str | int = "adaptive"):
        self.A = A
        self.method = METHODS[method] if isinstance(method, str) else int(method)
        self._desc = A.desc()
        h = C.c_void_p()
        lib().bsp_csr_plan_create(C.byref(self._desc), self.method, stream_handle(), C.byref(h))
        self._h = h

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

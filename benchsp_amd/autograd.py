"""Differentiable sparse-dense product Y = A(vals) @ X behind
torch.autograd.Function, computed by libbsp's HIP kernels:

  forward   Y  = A X            bsp_spmm    (FP32 accumulation, FP32 output)
  backward  dX = A^T dY         bsp_spmm_t  (CSC traversal of the pattern, values
                                             read through the CSR permutation)
            dvals = SDDMM(dY, X) bsp_sddmm   (dY[row_j] . X[col_j] for every entry)
  (both gradients at once: bsp_spmm_t_sddmm, one CSC pass, the same bits)

Values and dense operands may be float32 or bfloat16 (bf16 inputs, FP32
accumulation); gradients are returned in the dtype of the input they belong to.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import CsrDesc, lib
from .device import BF16, F32, ptr, stream_handle

_DT = {torch.float32: F32, torch.bfloat16: BF16}


class SparsePattern:
    """A CSR sparsity pattern resident on the GPU with its SpMM plan (work
    lists and the CSC permutation used by the transposed product)."""

    def __init__(self, rowptr: torch.Tensor, col: torch.Tensor, shape: tuple[int, int]):
        assert rowptr.dtype == torch.int32 and col.dtype == torch.int32 and rowptr.is_cuda
        self.rowptr, self.col = rowptr.contiguous(), col.contiguous()
        self.shape = (int(shape[0]), int(shape[1]))
        self.nnz = int(col.numel())
        d = CsrDesc(self.shape[0], self.shape[1], self.nnz, ptr(self.rowptr), ptr(self.col), None,
                    F32)
        h = C.c_void_p()
        lib().bsp_spmm_plan_create(C.byref(d), stream_handle(), C.byref(h))
        self._h = h

    @classmethod
    def from_host(cls, A, device="cuda") -> "SparsePattern":
        rp = torch.from_numpy(A.rowptr.astype(np.int32)).to(device)
        col = torch.from_numpy(A.col).to(device)
        return cls(rp, col, (A.nrows, A.ncols))

    def spmm(self, vals: torch.Tensor, X: torch.Tensor) -> torch.Tensor:
        Y = torch.empty(self.shape[0], X.shape[1], dtype=torch.float32, device=X.device)
        lib().bsp_spmm(self._h, ptr(vals), _DT[vals.dtype], ptr(X), _DT[X.dtype], X.shape[1],
                       ptr(Y), stream_handle())
        return Y

    def spmm_t(self, vals: torch.Tensor, dY: torch.Tensor) -> torch.Tensor:
        dX = torch.empty(self.shape[1], dY.shape[1], dtype=torch.float32, device=dY.device)
        lib().bsp_spmm_t(self._h, ptr(vals), _DT[vals.dtype], ptr(dY), _DT[dY.dtype],
                         dY.shape[1], ptr(dX), stream_handle())
        return dX

    def sddmm(self, dY: torch.Tensor, X: torch.Tensor) -> torch.Tensor:
        assert dY.dtype == X.dtype
        dv = torch.empty(self.nnz, dtype=torch.float32, device=X.device)
        lib().bsp_sddmm(self._h, ptr(dY), ptr(X), _DT[X.dtype], X.shape[1], ptr(dv),
                        stream_handle())
        return dv

    def spmm_t_sddmm(self, vals: torch.Tensor, dY: torch.Tensor, X: torch.Tensor):
        """(A^T dY, SDDMM(dY, X)) in one pass over the CSC: the same bits as
        spmm_t and sddmm."""
        assert dY.dtype == X.dtype
        dX = torch.empty(self.shape[1], dY.shape[1], dtype=torch.float32, device=dY.device)
        dv = torch.empty(self.nnz, dtype=torch.float32, device=X.device)
        lib().bsp_spmm_t_sddmm(self._h, ptr(vals), _DT[vals.dtype], ptr(dY), ptr(X),
                               _DT[X.dtype], X.shape[1], ptr(dX), ptr(dv), stream_handle())
        return dX, dv

    def __del__(self):
        try:
            if self._h:
                lib().bsp_spmm_plan_destroy(self._h)
                self._h = None
        except Exception:
            pass


class SpMMFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals: torch.Tensor, X: torch.Tensor, pattern: SparsePattern):
        vals, X = vals.contiguous(), X.contiguous()
        if vals.dtype not in _DT or X.dtype not in _DT:
            raise TypeError("values and X must be float32 or bfloat16")
        ctx.pattern = pattern
        ctx.save_for_backward(vals, X)
        return pattern.spmm(vals, X)

    @staticmethod
    def backward(ctx, dY: torch.Tensor):
        vals, X = ctx.saved_tensors
        p: SparsePattern = ctx.pattern
        # the dense operands of the backward kernels share X's precision
        dYc = dY.contiguous().to(X.dtype)
        dvals = dX = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            dX, dvals = p.spmm_t_sddmm(vals, dYc, X)
            dX, dvals = dX.to(X.dtype), dvals.to(vals.dtype)
        elif ctx.needs_input_grad[1]:
            dX = p.spmm_t(vals, dYc).to(X.dtype)
        elif ctx.needs_input_grad[0]:
            dvals = p.sddmm(dYc, X).to(vals.dtype)
        return dvals, dX, None


def spmm(vals: torch.Tensor, X: torch.Tensor, pattern: SparsePattern) -> torch.Tensor:
    """Y = A(vals) @ X with A's pattern given by `pattern` (differentiable in
    vals and X)."""
    return SpMMFunction.apply(vals, X, pattern)

"""Shared parity checks.

FP64 SpMV parity is stated componentwise: |y - y_ref|_i <= tol * (|A||x|)_i,
which is the bound any summation order meets (it is ill-posed to divide by
|y_ref|_i when a stencil row cancels to ~0). A norm-wise check
||y - y_ref||_inf <= tol * ||y_ref||_inf is applied as well.
"""
from __future__ import annotations

import numpy as np

FP64_RTOL = 1e-12


def abs_spmv(oracle, A, x):
    return oracle.spmv_f64(A.rowptr, A.col, np.abs(A.val), np.abs(x))


def assert_spmv_close(y, y_ref, absAx, tol=FP64_RTOL, what=""):
    y = np.asarray(y, np.float64)
    err = np.abs(y - y_ref)
    bound = tol * absAx + 1e-300
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, (
        f"{what}: {bad.size} rows beyond {tol:g}*(|A||x|); first row {bad[0]}: "
        f"y={y[bad[0]]!r} ref={y_ref[bad[0]]!r} bound={bound[bad[0]]:.3e}")
    ninf = np.max(np.abs(y_ref)) if y_ref.size else 0.0
    assert np.max(err, initial=0.0) <= tol * ninf or ninf == 0.0, what


def rel_fro(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (d if d > 0 else 1.0))

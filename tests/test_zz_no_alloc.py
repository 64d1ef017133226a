"""The allocation-free CSR SpMV (bsp_csr_spmv_noalloc) through the C-ABI:
bit-identical to bsp_csr_spmv on the unsegmented 2048-window adaptive plan
and within 1e-12 componentwise of the oracle, on two calls in a row (the
second with a smaller matrix, so stale schedule entries of the first would
show), run alone or after the other tests."""
import ctypes as C

import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu


def _ragged(rng):
    # empty rows, rows around the 1024-entry long-row cut, two long rows in
    # one window, one very long row
    lens = np.concatenate([[0, 3, 5000, 0, 1025, 1024, 1025, 1025, 3000, 2],
                           rng.integers(0, 40, 3000), [1100, 1100, 1100], np.zeros(50, int),
                           [9000]])
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(12000, L, replace=False)) for L in lens])
    return Csr(len(lens), 12000, rp, col.astype(np.int32), rng.uniform(-1, 1, int(rp[-1])))


def _plan_y(L, desc, x, y):
    # the existing entry point: an unsegmented adaptive plan, 2048-unit windows
    from benchsp_amd._lib import CsrOptions, lib
    opt = CsrOptions(-1, 2048, 0)
    h = C.c_void_p()
    lib().bsp_csr_plan_create_ex(C.byref(desc), 0, C.byref(opt), None, C.byref(h))
    try:
        lib().bsp_csr_spmv(h, x, y, None)
    finally:
        lib().bsp_plan_destroy(h)


@pytest.mark.parametrize("order", ["large_then_small", "small_then_large"])
def test_noalloc_matches_plan_and_oracle(oracle, bsp_host, order):
    import torch
    from benchsp_amd._lib import CsrDesc, lib
    from benchsp_amd.device import DeviceCsr, ptr
    rng = np.random.default_rng(3)
    mats = [("powerlaw-1M", bsp_host.powerlaw_1m()), ("stencil27-64", bsp_host.stencil27_64()),
            ("ragged", _ragged(rng)), ("poisson5-256", bsp_host.poisson5_256())]
    if order == "small_then_large":
        mats = mats[::-1]
    for k, (name, A) in enumerate(mats):
        print(f"[noalloc] call {k}: {name} n={A.nrows} nnz={A.nnz}", flush=True)
        d = DeviceCsr.from_host(A)
        desc = CsrDesc(A.nrows, A.ncols, A.nnz, ptr(d.rowptr), ptr(d.col), ptr(d.val), 0)
        x = bsp_host.uniform(A.ncols, 40 + k)
        xd = torch.from_numpy(x).cuda()
        y = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        yp = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        lib().bsp_csr_spmv_noalloc(C.byref(desc), ptr(xd), ptr(y), None)
        _plan_y(lib(), desc, ptr(xd), ptr(yp))
        torch.cuda.synchronize()
        yh = y.cpu().numpy()
        assert np.array_equal(yh, yp.cpu().numpy()), name
        assert_spmv_close(yh, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                          what=f"{name} noalloc")


def test_noalloc_empty_and_rowless(bsp_host):
    import torch
    from benchsp_amd._lib import CsrDesc, lib
    from benchsp_amd.device import DeviceCsr, ptr
    A = Csr(37, 29, np.zeros(38, np.int64), np.zeros(0, np.int32), np.zeros(0))
    d = DeviceCsr.from_host(A)
    desc = CsrDesc(37, 29, 0, ptr(d.rowptr), ptr(d.col), ptr(d.val), 0)
    x = torch.ones(29, dtype=torch.float64, device="cuda")
    y = torch.full((37,), float("nan"), dtype=torch.float64, device="cuda")
    lib().bsp_csr_spmv_noalloc(C.byref(desc), ptr(x), ptr(y), None)
    torch.cuda.synchronize()
    assert not y.cpu().numpy().any()
    desc = CsrDesc(0, 5, 0, ptr(d.rowptr), None, None, 0)
    lib().bsp_csr_spmv_noalloc(C.byref(desc), ptr(x), None, None)

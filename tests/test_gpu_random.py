"""Seeded random matrices of varied structure (uniform, lognormal row
lengths, banded, clustered blocks, a few very long rows, many empty rows,
wide and tall shapes) through every SpMV method and format, each against the
oracle's CPU loop (src/ref.cpp:26-37)."""
import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu


def _matrix(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = seed % 6
    n = int(rng.integers(1, 60_000))
    ncols = int(rng.integers(1, 80_000)) if kind in (4, 5) else n
    if kind == 0:    # uniform lengths
        lens = rng.integers(0, 30, n)
    elif kind == 1:  # lognormal lengths with a cap
        lens = np.minimum(np.rint(np.exp(rng.normal(2.0, 1.3, n))), 3000).astype(np.int64)
    elif kind == 2:  # banded
        lens = np.full(n, int(rng.integers(1, 40)))
    elif kind == 3:  # mostly empty rows, a few very long ones
        lens = np.where(rng.random(n) < 0.9, 0, rng.integers(1, 8, n))
        lens[rng.integers(0, n, 3)] = rng.integers(2000, 9000, 3)
    elif kind == 4:  # wide
        lens = rng.integers(0, 60, n)
    else:            # tall/clustered: runs of equal lengths
        lens = np.repeat(rng.integers(0, 50, n // 16 + 1), 16)[:n]
    lens = np.minimum(lens, ncols)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    cols = []
    for r, L in enumerate(lens):
        if L == 0:
            continue
        if kind == 2:  # a band around the diagonal
            lo = max(0, min(r - L // 2, ncols - L))
            cols.append(np.arange(lo, lo + L))
        elif kind == 5 and L < ncols:  # a block of consecutive columns
            lo = int(rng.integers(0, ncols - L + 1))
            cols.append(np.arange(lo, lo + L))
        else:
            cols.append(np.sort(rng.choice(ncols, L, replace=False)))
    col = (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)
    return Csr(n, ncols, rp, col, rng.uniform(-1, 1, int(rp[-1])))


@pytest.mark.parametrize("seed", range(12))
def test_random_matrix_every_method(oracle, bsp_host, seed):
    import torch
    from benchsp_amd.device import Bsr, CsrSpmv, Dasp, Dbsr, DeviceCsr
    A = _matrix(seed)
    x = bsp_host.uniform(A.ncols, seed=seed)
    y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    bound = abs_spmv(oracle, A, x)
    dA = DeviceCsr.from_host(A)
    xd = torch.from_numpy(x).cuda()
    ops = {m: CsrSpmv(dA, m) for m in ("adaptive", "scalar", "balanced", "split")}
    if A.ncols > 8:
        ops["adaptive-segments"] = CsrSpmv(dA, "adaptive", seg_cols=max(1, A.ncols // 5))
        ops["adaptive-sorted"] = CsrSpmv(dA, "adaptive", seg_cols=-(-A.ncols // 2), seg_mode=2)
    ops["dasp"] = Dasp(dA)
    ops["dbsr"] = Dbsr(dA)
    B4 = Bsr(dA, 4, 4)
    ops["bsr4-mfma"] = lambda v: B4(v, kernel=Bsr.MFMA)
    ops["bsr4-valu"] = lambda v: B4(v, kernel=Bsr.VALU)
    B3 = Bsr(dA, 3, 3)
    ops["bsr3-valu"] = lambda v: B3(v, kernel=Bsr.VALU)
    for name, op in ops.items():
        print(f"[random] seed={seed} n={A.nrows} ncols={A.ncols} nnz={A.nnz} {name}", flush=True)
        y = op(xd)
        torch.cuda.synchronize()
        assert_spmv_close(y.cpu().numpy(), y_ref, bound, what=f"seed {seed} {name}")

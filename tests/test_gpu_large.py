"""Inputs past 2^25 rows. One dispatch holds fewer than 2^32 work-items, so
every launch that gives a wave to each row (block row, row block) is issued
in slices of at most 2^31 work-items; at this size those launches take two
slices each: the column-segment scatter of the adaptive plan, DBSR placement,
the 1x1 BSR SpMV (a block row per row), DASP's row-block fill and the CSC key
pass of the transposed SpMM. Results against the oracle, and the transposed
SpMM against its closed form on sampled columns; the diagonal layout and the
fused backward (inverse permutation, CSC-ordered SDDMM put back in CSR
order) at the same size."""
import numpy as np
import pytest

from benchsp_amd.host import Csr
from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu

N = (1 << 25) + 4099  # just past one slice of a wave-per-row launch
OFFS = (-1, 0, 1, N // 3, 2 * N // 3)  # five distinct columns per row, two far off the diagonal


@pytest.fixture(scope="module")
def big(oracle):
    import torch
    from benchsp_amd.device import DeviceCsr
    r = np.arange(N, dtype=np.int64)
    cols = np.stack([(r + o) % N for o in OFFS], axis=1)
    cols.sort(axis=1)
    A = Csr(N, N, np.arange(0, 5 * N + 1, 5, dtype=np.int64), cols.reshape(-1).astype(np.int32),
            np.random.default_rng(11).uniform(-1, 1, 5 * N))
    del cols
    x = np.random.default_rng(12).uniform(-1, 1, N)
    print(f"[large] n={N} nnz={A.nnz}", flush=True)
    dA = DeviceCsr.from_host(A)
    xd = torch.from_numpy(x).cuda()
    y_ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    return A, dA, xd, y_ref, abs_spmv(oracle, A, x)


def _check(y, big, what):
    import torch
    torch.cuda.synchronize()
    _, _, _, y_ref, bound = big
    assert_spmv_close(y.cpu().numpy(), y_ref, bound, what=what)


def test_adaptive_segmented(big):
    from benchsp_amd.device import CsrSpmv
    P = CsrSpmv(big[1])
    assert P.segments()[1] > 1  # 40% of the entries lie far from the diagonal
    _check(P(big[2]), big, "adaptive, column segments")
    P.close()


def test_sorted_segments(big):
    # eight segments stored as column-sorted blocks: over 30,000 blocks,
    # a radix sort of every entry, the row-order slots
    from benchsp_amd.device import CsrSpmv
    ncols = big[1].ncols
    P = CsrSpmv(big[1], "adaptive", seg_cols=-(-ncols // 8), seg_mode=2)
    assert P.segments()[1] == 8
    _check(P(big[2]), big, "adaptive, sorted-block segments")
    P.close()


def test_dbsr(big):
    from benchsp_amd.device import Dbsr
    D = Dbsr(big[1])
    _check(D(big[2]), big, "dbsr")
    del D


def test_bsr_1x1(big):
    from benchsp_amd.device import Bsr
    B = Bsr(big[1], 1, 1)
    _check(B(big[2], kernel=Bsr.VALU), big, "bsr 1x1")
    del B


def test_dasp(big):
    from benchsp_amd.device import Dasp
    D = Dasp(big[1])
    _check(D(big[2]), big, "dasp")
    del D


def test_dia_bit_identical(big):
    # the wrapped offsets give nine diagonals: the diagonal layout holds the
    # matrix in 9 x 33.6 M slots and returns the CPU loop's bits
    import torch
    from benchsp_amd.device import Dia
    D = Dia(big[1])
    assert D.info()["ndiag"] == 9
    y = D(big[2])
    torch.cuda.synchronize()
    assert np.array_equal(y.cpu().numpy(), big[3])
    del D, y
    torch.cuda.empty_cache()


def test_spmm_transposed_sampled(big):
    # dX = A^T dY, F = 64: column c of A appears in rows (c - o) mod N, one
    # entry per offset o; checked on 256 sampled columns
    import torch
    from benchsp_amd.autograd import SparsePattern
    A, dA = big[0], big[1]
    S = SparsePattern(dA.rowptr, dA.col, (N, N))
    vals = dA.val.float()
    dY = torch.randn(N, 64, device="cuda", generator=torch.Generator("cuda").manual_seed(5))
    dX = S.spmm_t(vals, dY)
    # the fused backward at this size: the same dX bits, and SDDMM entries on
    # a sample against their dot products
    X = torch.randn(N, 64, device="cuda", generator=torch.Generator("cuda").manual_seed(6))
    fdX, fdv = S.spmm_t_sddmm(vals, dY, X)
    assert torch.equal(fdX, dX)
    js = torch.from_numpy(np.random.default_rng(14).choice(A.nnz, 4096, replace=False)).cuda()
    rj = torch.div(js, 5, rounding_mode="floor")
    cj = dA.col.long()[js]
    ref = (dY[rj].double() * X[cj].double()).sum(1)
    scale = (dY[rj].double().abs() * X[cj].double().abs()).sum(1)
    assert torch.all((fdv[js].double() - ref).abs() <= 1e-5 * scale + 1e-30)
    del X, fdX, fdv
    torch.cuda.synchronize()
    cs = np.random.default_rng(13).choice(N, 256, replace=False)
    got = dX[torch.from_numpy(cs).cuda()].cpu().numpy().astype(np.float64)
    rows = np.stack([(cs - o) % N for o in OFFS], axis=1)  # 256 x 5
    dy = dY[torch.from_numpy(rows.reshape(-1)).cuda()].cpu().numpy().astype(np.float64)
    dy = dy.reshape(256, 5, 64)
    v = np.empty((256, 5))
    for i, c in enumerate(cs):
        for k, r in enumerate(rows[i]):
            seg = A.col[5 * r:5 * r + 5]
            v[i, k] = np.float32(A.val[5 * r + int(np.nonzero(seg == c)[0][0])])
    want = np.einsum("ik,ikf->if", v, dy)
    scale = np.einsum("ik,ikf->if", np.abs(v), np.abs(dy))
    assert np.all(np.abs(got - want) <= 1e-5 * scale + 1e-30)
    del S, dY, dX
    torch.cuda.empty_cache()

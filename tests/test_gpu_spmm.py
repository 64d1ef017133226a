"""GPU parity of the differentiable CSR SpMM (forward, transpose-free A^T dY,
SDDMM) against the FP64 oracle, finite differences and a dense FP64
gradient, in FP32 and in BF16 with FP32 accumulation."""
import numpy as np
import pytest

from tests._parity import rel_fro

pytestmark = pytest.mark.gpu

FP32_TOL = 2e-6   # relative Frobenius error vs FP64, FP32 inputs
BF16_TOL = 1e-2   # relative Frobenius error vs FP64, BF16 inputs (stated tolerance)


def _bf16_round(a):
    import torch
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float64).numpy()


def _setup(A, F, dtype, seed=0):
    import torch
    from benchsp_amd.autograd import SparsePattern
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(A.ncols, F)).astype(np.float32)
    dY = rng.normal(size=(A.nrows, F)).astype(np.float32)
    P = SparsePattern.from_host(A)
    tv = torch.from_numpy(A.val.astype(np.float32)).cuda().to(dtype)
    tX = torch.from_numpy(X).cuda().to(dtype)
    tdY = torch.from_numpy(dY).cuda().to(dtype)
    return P, tv, tX, tdY, X, dY


@pytest.mark.parametrize("dtype_name", ["float32", "bfloat16"])
@pytest.mark.parametrize("F", [64, 128, 256])
@pytest.mark.parametrize("symmetric", [True, False])
def test_spmm_fwd_bwd_small(oracle, bsp_host, dtype_name, F, symmetric):
    import torch
    dtype = getattr(torch, dtype_name)
    A = bsp_host.graph(3000, 30.0, 2.1, symmetric, seed=7)
    P, tv, tX, tdY, X, dY = _setup(A, F, dtype)
    Y = P.spmm(tv, tX).cpu().numpy()
    dX = P.spmm_t(tv, tdY).cpu().numpy()
    dv = P.sddmm(tdY, tX).cpu().numpy()
    # the fused backward gives the same bits as the two separate products
    fdX, fdv = P.spmm_t_sddmm(tv, tdY, tX)
    assert np.array_equal(fdX.cpu().numpy(), dX)
    assert np.array_equal(fdv.cpu().numpy(), dv)
    # dense FP64 references
    D = A.to_dense()
    if dtype == torch.bfloat16:
        tol, Dq, Xq, dYq = BF16_TOL, D, X, dY
    else:
        tol, Dq, Xq, dYq = FP32_TOL, A.to_dense(), X.astype(np.float64), dY.astype(np.float64)
    assert rel_fro(Y, Dq @ Xq) <= tol
    assert rel_fro(dX, Dq.T @ dYq) <= tol
    rows = np.repeat(np.arange(A.nrows), A.row_lengths())
    assert rel_fro(dv, (dYq @ Xq.T)[rows, A.col]) <= tol
    if not symmetric:
        # a backward that used A instead of A^T would fail here
        assert rel_fro(D @ dYq, Dq.T @ dYq) > 0.1


def test_autograd_function_matches_fp64_and_fd(oracle, bsp_host):
    import torch
    from benchsp_amd.autograd import spmm
    A = bsp_host.graph(2000, 25.0, 2.1, False, seed=3)
    P, tv, tX, tdY, X, dY = _setup(A, 64, torch.float32, seed=5)
    tv.requires_grad_(True)
    tX.requires_grad_(True)
    Y = spmm(tv, tX, P)
    (Y * tdY).sum().backward()
    gv = tv.grad.cpu().numpy()
    gX = tX.grad.cpu().numpy()
    v64 = A.val.astype(np.float32).astype(np.float64)
    X64 = X.astype(np.float64)
    dY64 = dY.astype(np.float64)
    assert rel_fro(gX, oracle.csr_spmm_t_f64(A.rowptr, A.col, v64, dY64, A.ncols)) <= FP32_TOL
    assert rel_fro(gv, oracle.sddmm_f64(A.rowptr, A.col, dY64, X64)) <= FP32_TOL
    # finite differences of L = <dY, A X> in FP64, eps = 1e-3, on sampled entries
    rng = np.random.default_rng(0)
    eps = 1e-3
    rows = np.repeat(np.arange(A.nrows), A.row_lengths())
    for j in rng.choice(A.nnz, 64, replace=False):
        r = rows[j]
        s, e = A.rowptr[r], A.rowptr[r + 1]
        def row_loss(v):  # L restricted to row r, the only row val_j enters
            return float(dY64[r] @ (v[s:e] @ X64[A.col[s:e]]))
        vp, vm = v64.copy(), v64.copy()
        vp[j] += eps
        vm[j] -= eps
        fd = (row_loss(vp) - row_loss(vm)) / (2 * eps)
        assert abs(fd - gv[j]) <= 1e-4 * (1 + abs(fd))


# BF16 inputs, FP32 accumulation, gradients returned in BF16: one rounding
# of the output to BF16 (at most 2^-9 relative, stated with a factor-2
# margin) plus the FP32 accumulation of 64 products (64 * 2^-24 of the sum
# of their magnitudes, stated as 1e-5).
BF16_FD_REL = 2.0**-8
BF16_FD_ABS = 1e-5


def test_autograd_bf16_matches_fd(oracle, bsp_host):
    import torch
    from benchsp_amd.autograd import spmm
    A = bsp_host.graph(2000, 25.0, 2.1, False, seed=3)
    P, tv, tX, tdY, _, _ = _setup(A, 64, torch.bfloat16, seed=6)
    tv.requires_grad_(True)
    tX.requires_grad_(True)
    Y = spmm(tv, tX, P)
    (Y * tdY.float()).sum().backward()
    assert tv.grad.dtype == torch.bfloat16 and tX.grad.dtype == torch.bfloat16
    gv = tv.grad.float().cpu().numpy().astype(np.float64)
    gX = tX.grad.float().cpu().numpy().astype(np.float64)
    # the operands the kernels saw, exactly, in FP64
    v64 = tv.detach().float().cpu().numpy().astype(np.float64)
    X64 = tX.detach().float().cpu().numpy().astype(np.float64)
    dY64 = tdY.float().cpu().numpy().astype(np.float64)
    assert rel_fro(Y.detach().cpu().numpy(), oracle.csr_spmm_f64(A.rowptr, A.col, v64, X64)) <= 1e-5
    eps = 1e-3
    rng = np.random.default_rng(2)
    rows = np.repeat(np.arange(A.nrows), A.row_lengths())
    # dvals: L = <dY, A X> restricted to row r, the only row val_j enters
    for j in rng.choice(A.nnz, 64, replace=False):
        r = rows[j]
        s, e = A.rowptr[r], A.rowptr[r + 1]
        def row_loss(v):
            return float(dY64[r] @ (v[s:e] @ X64[A.col[s:e]]))
        vp, vm = v64.copy(), v64.copy()
        vp[j] += eps
        vm[j] -= eps
        fd = (row_loss(vp) - row_loss(vm)) / (2 * eps)
        mag = float(np.abs(dY64[r]) @ np.abs(X64[A.col[j]]))
        assert abs(fd - gv[j]) <= BF16_FD_REL * abs(fd) + BF16_FD_ABS * mag, (j, fd, gv[j])
    # dX: L restricted to the entries of column c, the only ones X[c, f] enters
    for c, f in zip(rng.choice(A.ncols, 32, replace=False), rng.integers(0, 64, 32)):
        js = np.nonzero(A.col == c)[0]
        def col_loss(xcf):
            return float(np.sum(v64[js] * dY64[rows[js], f] * xcf))
        fd = (col_loss(X64[c, f] + eps) - col_loss(X64[c, f] - eps)) / (2 * eps)
        mag = float(np.sum(np.abs(v64[js] * dY64[rows[js], f])))
        assert abs(fd - gX[c, f]) <= BF16_FD_REL * abs(fd) + BF16_FD_ABS * mag, (c, f, fd)


@pytest.mark.parametrize("dtype_name", ["float32", "bfloat16"])
def test_spmm_gnn_200k(oracle, bsp_host, dtype_name):
    import torch
    dtype = getattr(torch, dtype_name)
    for symmetric in (True, False):
        A = bsp_host.zipf_graph_200k(symmetric=symmetric)
        print(f"[spmm] n={A.nrows} nnz={A.nnz} sym={symmetric} {dtype_name}", flush=True)
        P, tv, tX, tdY, X, dY = _setup(A, 64, dtype, seed=11)
        Y = P.spmm(tv, tX).cpu().numpy()
        dX = P.spmm_t(tv, tdY).cpu().numpy()
        dv = P.sddmm(tdY, tX).cpu().numpy()
        if dtype == torch.bfloat16:
            tol = BF16_TOL
        else:
            tol = FP32_TOL
        v = A.val.astype(np.float32).astype(np.float64)
        assert rel_fro(Y, oracle.csr_spmm_f64(A.rowptr, A.col, v, X)) <= tol
        assert rel_fro(dX, oracle.csr_spmm_t_f64(A.rowptr, A.col, v, dY, A.ncols)) <= tol
        # dvals on a sample of 4096 entries, FP64 dot products
        rng = np.random.default_rng(1)
        js = rng.choice(A.nnz, 4096, replace=False)
        rows = np.searchsorted(A.rowptr, js, side="right") - 1
        ref = np.einsum("ij,ij->i", dY[rows].astype(np.float64), X[A.col[js]].astype(np.float64))
        assert rel_fro(dv[js], ref) <= tol
        # deterministic: a second run is bit-identical
        assert np.array_equal(P.spmm(tv, tX).cpu().numpy(), Y)
        fdX, fdv = P.spmm_t_sddmm(tv, tdY, tX)
        assert np.array_equal(fdX.cpu().numpy(), dX)
        assert np.array_equal(fdv.cpu().numpy(), dv)


def test_spmm_rejects_unaligned_operands(bsp_host):
    # rows of four features are read with 16-byte loads: a dense operand or
    # output that is not 16-byte aligned is an argument error
    import torch
    from benchsp_amd._lib import BspError
    A = bsp_host.graph(500, 10.0, 2.1, True, seed=3)
    P, tv, tX, tdY, _, _ = _setup(A, 64, torch.float32)
    Xu = torch.zeros(A.ncols * 64 + 1, device="cuda")[1:].view(A.ncols, 64)
    Xu.copy_(tX)
    assert Xu.data_ptr() % 16 != 0
    with pytest.raises(BspError):
        P.spmm(tv, Xu)
    with pytest.raises(BspError):
        P.sddmm(tdY, Xu)
    dYu = torch.zeros(A.nrows * 64 + 1, device="cuda")[1:].view(A.nrows, 64)
    with pytest.raises(BspError):
        P.spmm_t(tv, dYu)

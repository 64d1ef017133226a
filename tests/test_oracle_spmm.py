"""Pins the FP64 SpMM / SDDMM / block-SpMM oracles against numpy dense
products (a library routine), and the SDDMM against finite differences of
the loss L = <dY, A X> (it must be that gradient)."""
import numpy as np

from benchsp_amd.host import Csr


def _rand(rng, n, m, deg):
    lens = rng.integers(0, deg, n)
    rp = np.zeros(n + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(m, L, replace=False)) for L in lens]).astype(np.int32)
    return Csr(n, m, rp, col, rng.normal(size=int(rp[-1])))


def test_spmm_transpose_sddmm_vs_dense(oracle):
    rng = np.random.default_rng(0)
    A = _rand(rng, 70, 55, 12)  # rectangular and not symmetric
    D = A.to_dense()
    X = rng.normal(size=(55, 8))
    dY = rng.normal(size=(70, 8))
    np.testing.assert_allclose(oracle.csr_spmm_f64(A.rowptr, A.col, A.val, X), D @ X,
                               rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(oracle.csr_spmm_t_f64(A.rowptr, A.col, A.val, dY, 55), D.T @ dY,
                               rtol=1e-13, atol=1e-13)
    rows = np.repeat(np.arange(70), A.row_lengths())
    np.testing.assert_allclose(oracle.sddmm_f64(A.rowptr, A.col, dY, X),
                               (dY @ X.T)[rows, A.col], rtol=1e-13, atol=1e-13)


def test_sddmm_is_the_gradient(oracle):
    rng = np.random.default_rng(1)
    A = _rand(rng, 40, 33, 9)
    X = rng.normal(size=(33, 6))
    dY = rng.normal(size=(40, 6))
    g = oracle.sddmm_f64(A.rowptr, A.col, dY, X)

    def loss(v):
        return float(np.sum(dY * oracle.csr_spmm_f64(A.rowptr, A.col, v, X)))

    eps = 1e-3
    for j in rng.choice(A.nnz, 25, replace=False):
        vp, vm = A.val.copy(), A.val.copy()
        vp[j] += eps
        vm[j] -= eps
        assert abs((loss(vp) - loss(vm)) / (2 * eps) - g[j]) <= 1e-9 * (1 + abs(g[j]))
    # and A^T dY is dL/dX
    gX = oracle.csr_spmm_t_f64(A.rowptr, A.col, A.val, dY, 33)
    for (r, c) in [(0, 0), (5, 3), (32, 5), (17, 1)]:
        Xp, Xm = X.copy(), X.copy()
        Xp[r, c] += eps
        Xm[r, c] -= eps
        fd = (np.sum(dY * oracle.csr_spmm_f64(A.rowptr, A.col, A.val, Xp)) -
              np.sum(dY * oracle.csr_spmm_f64(A.rowptr, A.col, A.val, Xm))) / (2 * eps)
        assert abs(fd - gX[r, c]) <= 1e-9 * (1 + abs(gX[r, c]))


def test_bsr_spmm_vs_dense(oracle, bsp_host):
    A = bsp_host.blocksparse(6, 7, 16, 11, seed=2)
    b = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 16, 16)
    B = np.random.default_rng(3).normal(size=(A.ncols, 24))
    np.testing.assert_allclose(oracle.bsr_spmm_f64(b["rowptr"], b["colidx"], b["val"], 16, B),
                               A.to_dense() @ B, rtol=1e-12, atol=1e-12)

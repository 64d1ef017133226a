"""The diagonal layout (bsp_dia_*): y bit-identical to the oracle's
row-sequential FP64 sum on stencils and on banded matrices with holes,
rectangular shapes, empty rows and non-finite x; matrices it cannot hold
are refused with BSP_ERR_UNSUPPORTED."""
import numpy as np
import pytest

from benchsp_amd.host import Csr

pytestmark = pytest.mark.gpu


def _banded(rng, nrows, ncols, offsets, keep):
    """Entries on the given diagonals, each kept with probability `keep`
    (holes inside diagonals), values U[-1, 1] with some exact zeros."""
    rows, cols = [], []
    for r in range(nrows):
        for d in sorted(offsets):
            c = r + d
            if 0 <= c < ncols and rng.random() < keep:
                rows.append(r)
                cols.append(c)
    rows, cols = np.array(rows, np.int64), np.array(cols, np.int64)
    rp = np.zeros(nrows + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp = np.cumsum(rp)
    val = rng.uniform(-1, 1, len(cols))
    val[rng.random(len(cols)) < 0.05] = 0.0
    return Csr(nrows, ncols, rp, cols.astype(np.int32), val)


def _check(oracle, A, x, what):
    import torch
    from benchsp_amd.device import DeviceCsr, Dia
    D = Dia(DeviceCsr.from_host(A))
    y = D(torch.from_numpy(x).cuda()).cpu().numpy()
    ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    assert np.array_equal(y, ref, equal_nan=True), what
    return D


def test_dia_stencils_bit_exact(oracle, bsp_host):
    for name, A, nd in (("poisson5-256", bsp_host.poisson5_256(), 5),
                        ("stencil27-64", bsp_host.stencil27_64(), 27),
                        ("stencil7-40x30x20", bsp_host.stencil3d(40, 30, 20, 7, seed=3), 7)):
        print(f"[dia] {name}", flush=True)
        x = bsp_host.uniform(A.ncols, 5)
        D = _check(oracle, A, x, name)
        info = D.info()
        assert info["ndiag"] == nd, name
        assert info["offsets"] == sorted(set((A.col - np.repeat(np.arange(A.nrows),
                                                                np.diff(A.rowptr))).tolist()))
        assert info["bytes"] == 8 * nd * A.nrows + 4 * A.nrows


def test_dia_stencil27_offsets(bsp_host):
    import torch
    from benchsp_amd.device import DeviceCsr, Dia
    A = bsp_host.stencil27_64()
    got = Dia(DeviceCsr.from_host(A)).info()["offsets"]
    want = sorted(dz * 4096 + dy * 64 + dx for dz in (-1, 0, 1) for dy in (-1, 0, 1)
                  for dx in (-1, 0, 1))
    assert got == want
    torch.cuda.synchronize()


@pytest.mark.parametrize("nrows,ncols,noffs,keep", [(1000, 1000, 9, 0.8), (777, 1200, 32, 0.6),
                                                     (1500, 300, 17, 1.0), (64, 64, 1, 0.5),
                                                     (3, 5000, 4, 1.0)])
def test_dia_banded_with_holes(oracle, nrows, ncols, noffs, keep):
    rng = np.random.default_rng(nrows + noffs)
    offs = rng.choice(np.arange(-nrows + 1, ncols), noffs, replace=False)
    A = _banded(rng, nrows, ncols, offs, keep)
    x = rng.uniform(-1, 1, ncols)
    _check(oracle, A, x, f"{nrows}x{ncols} {noffs} diagonals")


def test_dia_non_finite_x_only_where_referenced(oracle, bsp_host):
    # x = inf / nan at columns some rows reach through a diagonal they do not
    # hold: those rows stay finite (absent slots are skipped, not 0 * inf)
    rng = np.random.default_rng(11)
    A = _banded(rng, 500, 500, [-3, -1, 0, 1, 3], 0.7)
    x = rng.uniform(-1, 1, 500)
    x[[10, 200, 201]] = np.inf
    x[[50, 377]] = np.nan
    _check(oracle, A, x, "non-finite x")


def test_dia_y_not_16_byte_aligned(oracle, bsp_host):
    # an even row count takes two rows per lane (16-byte y stores) only when
    # y is 16-byte aligned; an 8-byte aligned y takes the row-per-lane kernel
    import torch
    from benchsp_amd.device import DeviceCsr, Dia
    A = bsp_host.poisson5_256()
    x = bsp_host.uniform(A.ncols, 6)
    D = Dia(DeviceCsr.from_host(A))
    ref = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    buf = torch.full((A.nrows + 1,), 7.0, dtype=torch.float64, device="cuda")
    y = D(torch.from_numpy(x).cuda(), buf[1:])
    assert np.array_equal(y.cpu().numpy(), ref)
    assert buf[0].item() == 7.0


def test_dia_empty_and_degenerate(oracle):
    rng = np.random.default_rng(2)
    for name, A in (("no entries", Csr(37, 29, np.zeros(38, np.int64), np.zeros(0, np.int32),
                                       np.zeros(0))),
                    ("empty rows", _banded(rng, 40, 40, [0], 0.2)),
                    ("one row", _banded(rng, 1, 100, [0, 5, 99], 1.0)),
                    ("one col", _banded(rng, 100, 1, [0, -7, -99], 1.0))):
        _check(oracle, A, rng.uniform(-1, 1, A.ncols), name)


def test_dia_refuses_what_it_cannot_hold(bsp_host):
    from benchsp_amd._lib import BspError
    from benchsp_amd.device import DeviceCsr, Dia
    rng = np.random.default_rng(4)
    many = _banded(rng, 400, 400, list(range(-20, 20)), 0.9)  # 40 diagonals
    unsorted = Csr(3, 4, np.array([0, 2, 3, 3]), np.array([2, 1, 0], np.int32),
                   np.array([1.0, 2.0, 3.0]))
    long_rows = bsp_host.powerlaw(2000, 40.0, 1.0, 300, seed=1)
    for name, A, status in (("40 diagonals", many, 5), ("unsorted", unsorted, 4),
                            ("long rows", long_rows, 5)):
        with pytest.raises(BspError) as e:
            Dia(DeviceCsr.from_host(A))
        assert e.value.status == status, name  # BSP_ERR_UNSUPPORTED / BSP_ERR_UNSORTED

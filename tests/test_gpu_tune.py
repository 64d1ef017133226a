"""The per-matrix SpMV autotuner (bsp_autotune): whatever layout it picks,
y equals the oracle's within the FP64 bound; a stored record rebuilds the
same layout and gives the same bits; format masks restrict the candidates."""
import json

import numpy as np
import pytest

from tests._parity import abs_spmv, assert_spmv_close
from tests.test_gpu_formats import _degenerate, _rand_csr

pytestmark = pytest.mark.gpu


def _mats(bsp_host):
    rng = np.random.default_rng(9)
    yield "random", _rand_csr(rng, rng.integers(0, 40, 301), 263)
    yield "stencil27-64", bsp_host.stencil27_64()
    yield "powerlaw-100k", bsp_host.powerlaw(100_000, 20.0, 1.2, 2000, seed=4)
    yield "blocks-4x4", bsp_host.blocksparse(2000, 2000, 4, 20_000, seed=6)  # BSR4 candidates
    yield "blocks-8x8", bsp_host.blocksparse(1000, 1000, 8, 8_000, seed=10)  # BSR4 and BSR8
    yield from _degenerate(rng)


def test_autotune_parity_and_replay(oracle, bsp_host):
    import torch
    from benchsp_amd.device import DeviceCsr, Tuned
    seen = set()
    for name, A in _mats(bsp_host):
        dA = DeviceCsr.from_host(A)
        x = bsp_host.uniform(A.ncols, 17)
        xd = torch.from_numpy(x).cuda()
        T = Tuned(dA, reps=3)
        rec, trials = T.record, T.trials()
        print(f"[tune] {name}: {rec['format']} ({len(trials)} candidates)", flush=True)
        assert rec in trials and rec["time_us"] == min(t["time_us"] for t in trials)
        formats = {t["format"] for t in trials}
        assert {"csr", "dbsr", "dasp"} <= formats
        assert ("bsr4" in formats) == (name in ("blocks-4x4", "blocks-8x8"))
        assert ("bsr8" in formats) == (name == "blocks-8x8")
        if name in ("stencil27-64", "powerlaw-100k"):  # 27 diagonals / rows too long
            assert ("dia" in formats) == (name == "stencil27-64")
        seen.add(rec["format"])
        y = T(xd)
        torch.cuda.synchronize()
        assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                          abs_spmv(oracle, A, x), what=f"{name} tuned {rec['format']}")
        # the record survives a JSON round trip and rebuilds the same layout
        R = Tuned.from_record(dA, json.loads(json.dumps(rec)))
        assert R.record == dict(rec, time_us=0.0) and R.trials() == []
        y2 = R(xd)
        torch.cuda.synchronize()
        assert torch.equal(y, y2), name


@pytest.mark.parametrize("fmt", ["csr", "dbsr", "dasp", "bsr4", "bsr8"])
def test_autotune_format_mask(oracle, bsp_host, fmt):
    import torch
    from benchsp_amd.device import DeviceCsr, Tuned
    A = bsp_host.blocksparse(250, 250, 8, 2000, seed=7)  # every format qualifies
    dA = DeviceCsr.from_host(A)
    T = Tuned(dA, formats=[fmt], reps=2)
    assert {t["format"] for t in T.trials()} == {fmt} and T.record["format"] == fmt
    if fmt == "dasp":
        assert {(t["threshold"], t["block_longest"]) for t in T.trials()} == {
            (t, b) for t in (0.5, 0.75, 1.0) for b in (64, 256)}
    # kernel variants: VALU and MFMA for DBSR and DASP (at 0.75 / 256), the
    # four 4x4 BSR kernels, VALU and MFMA for 8x8
    kernels = {t["kernel"] for t in T.trials()}
    assert kernels == {"dbsr": {0, 1}, "dasp": {0, 1}, "bsr4": {1, 2, 3, 4}, "bsr8": {1, 2},
                       "csr": {0}}[fmt]
    x = bsp_host.uniform(A.ncols, 3)
    y = T(torch.from_numpy(x).cuda()).cpu().numpy()
    assert_spmv_close(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x), abs_spmv(oracle, A, x),
                      what=f"mask {fmt}")


def test_autotune_sorted_block_candidate(oracle, bsp_host):
    # a matrix that segments into 4 XCD-pinned pieces also gets the sorted
    # blocks (CSR kernel 1) as a candidate; its record rebuilds the plan
    import torch
    from benchsp_amd.device import DeviceCsr, Tuned
    A = bsp_host.powerlaw_1m()
    dA = DeviceCsr.from_host(A)
    T = Tuned(dA, formats=["csr"], reps=2)
    sorted_recs = [t for t in T.trials() if t["kernel"] == 1]
    assert {t["kernel"] for t in T.trials()} == {0, 1} and len(sorted_recs) == 1
    x = bsp_host.uniform(A.ncols, 5)
    xd = torch.from_numpy(x).cuda()
    R = Tuned.from_record(dA, sorted_recs[0])
    y, y2 = R(xd), R(xd)
    torch.cuda.synchronize()
    assert torch.equal(y, y2)
    assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, x),
                      abs_spmv(oracle, A, x), what="tuned sorted blocks")


def test_autotune_sparse_matrix_skips_bsr4(bsp_host):
    # about one entry per 4x4 block: BSR4 is screened out before any build,
    # so a BSR4-only mask leaves no candidate
    from benchsp_amd._lib import BspError
    from benchsp_amd.device import DeviceCsr, Tuned
    A = bsp_host.powerlaw(20_000, 8.0, 1.0, 500, seed=8)
    with pytest.raises(BspError) as e:
        Tuned(DeviceCsr.from_host(A), formats=["bsr4"], reps=2)
    assert e.value.status == 5  # BSP_ERR_UNSUPPORTED


def test_autotune_dia_mask(oracle, bsp_host):
    # the diagonal layout alone on a 7-point stencil: y is the oracle's bits
    import torch
    from benchsp_amd.device import DeviceCsr, Tuned
    A = bsp_host.stencil3d(30, 20, 10, 7, seed=2)
    T = Tuned(DeviceCsr.from_host(A), formats=["dia"], reps=2)
    assert [t["format"] for t in T.trials()] == ["dia"] and T.record["format"] == "dia"
    x = bsp_host.uniform(A.ncols, 8)
    y = T(torch.from_numpy(x).cuda()).cpu().numpy()
    assert np.array_equal(y, oracle.spmv_f64(A.rowptr, A.col, A.val, x))

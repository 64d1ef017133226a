"""The batched workload at its full size: 512 matrices with n log-uniform in
[1k, 64k] (one third 5-point, one third 7-point stencils, one third
lognormal-random, 60.2 M entries), each with an FP64 SpMV and an 8-column
FP16 SpMM, fixed and autotuned, eagerly and as graph replays; every matrix
against the oracle."""
import numpy as np
import pytest

from tests.test_gpu_batch import COLS, _check, _device, _vectors

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full_mix(bsp_host):
    mats = bsp_host.batch_mix(512)
    assert len(mats) == 512
    n = np.array([A.nrows for A in mats])
    assert n.min() >= 900 and n.max() <= 66_000
    return mats


@pytest.mark.parametrize("tune", [0, 1, 4])
def test_batch_full_workload(oracle, bsp_host, full_mix, tune):
    import torch
    from benchsp_amd.device import Batch
    mats = full_mix
    xs, Xs = _vectors(bsp_host, mats, 101)
    dm, x, X = _device(mats, xs, Xs)
    print(f"[batch-full] tune={tune} build", flush=True)
    B = Batch(dm, tune=tune)
    ch = B.choice()
    if tune == 0:
        assert np.all(ch["format"] == 0)
    if tune == 4:  # every stencil in the diagonal layout, the random third in CSR
        stencil = np.arange(len(mats)) % 3 < 2
        assert np.all((ch["format"] == Batch.DIA) == stencil)
        assert np.all((ch["spmm_format"] == Batch.DIA) == stencil)
    y = torch.full((B.nrows,), float("nan"), dtype=torch.float64, device="cuda")
    Y = torch.full((B.nrows, COLS), float("nan"), dtype=torch.float32, device="cuda")
    print(f"[batch-full] tune={tune} eager", flush=True)
    B.run(x, y, X, Y)
    torch.cuda.synchronize()
    _check(oracle, mats, xs, Xs, y.cpu().numpy(), Y.cpu().numpy(), B)
    # graph replays with new inputs between them, each checked
    B.capture(x, y, X, Y)
    for rep in range(2):
        print(f"[batch-full] tune={tune} replay {rep}", flush=True)
        xs, Xs = _vectors(bsp_host, mats, 200 + 13 * rep)
        x.copy_(torch.from_numpy(np.concatenate(xs)))
        X.copy_(torch.from_numpy(np.concatenate(Xs).reshape(-1, COLS)))
        y.fill_(float("nan"))
        Y.fill_(float("nan"))
        torch.cuda.synchronize()
        B.replay()
        torch.cuda.synchronize()
        _check(oracle, mats, xs, Xs, y.cpu().numpy(), Y.cpu().numpy(), B)
        _ = (x * 2).sum().item()  # eager work between replays

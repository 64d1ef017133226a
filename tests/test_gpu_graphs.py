"""Every SpMV plan and format handle runs inside a captured HIP graph (torch's
CUDA-graph capture on a side stream): no allocation, synchronisation or
memset happens at call time. Each graph is replayed several times with a new
x copied in between, with eager work and synchronisation between replays,
and every replay is checked against the oracle (src/ref.cpp:26-37)."""
import numpy as np
import pytest

from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu


def _ops(bsp_host):
    from benchsp_amd.device import Bsr, CsrSpmv, Dasp, Dbsr, DeviceCsr
    mats = {"27pt-64": bsp_host.stencil27_64(),
            "powerlaw-300k": bsp_host.powerlaw(300_000, 20.0, 1.2, 5000, seed=6)}
    for name, A in mats.items():
        dA = DeviceCsr.from_host(A)
        for method in ("adaptive", "scalar", "balanced", "split"):
            yield f"{name}/{method}", A, dA, CsrSpmv(dA, method)
        yield f"{name}/dasp", A, dA, Dasp(dA)
        yield f"{name}/dbsr", A, dA, Dbsr(dA)
        B = Bsr(dA)
        yield f"{name}/bsr4-mfma", A, dA, (lambda b: (lambda x, y: b(x, y, kernel=Bsr.MFMA)))(B)
    # forced column segments, both modes (XCD-pinned one launch; one launch per segment)
    A = mats["powerlaw-300k"]
    dA = DeviceCsr.from_host(A)
    for mode in (0, 1):
        yield f"powerlaw-300k/segments-mode{mode}", A, dA, CsrSpmv(dA, "adaptive", seg_cols=40_000,
                                                                  seg_mode=mode)


def test_spmv_plans_replay_in_graphs(oracle, bsp_host):
    import torch
    for name, A, dA, op in _ops(bsp_host):
        print(f"[graph] {name}", flush=True)
        x = torch.zeros(A.ncols, dtype=torch.float64, device="cuda")
        y = torch.full((A.nrows,), float("nan"), dtype=torch.float64, device="cuda")
        op(x, y)  # first call outside the graph
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            op(x, y)
        for rep in range(3):
            xh = bsp_host.uniform(A.ncols, seed=100 + rep)
            x.copy_(torch.from_numpy(xh))
            y.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert_spmv_close(y.cpu().numpy(), oracle.spmv_f64(A.rowptr, A.col, A.val, xh),
                              abs_spmv(oracle, A, xh), what=f"{name} replay {rep}")
            _ = (x * 2).sum().item()  # eager work between replays
        del g

"""Every SpMV plan and format handle runs inside a captured HIP graph (torch's
CUDA-graph capture on a side stream): no allocation, synchronisation or
memset happens at call time. Each graph is replayed several times with a new
x copied in between, with eager work and synchronisation between replays,
and every replay is checked against the oracle (src/ref.cpp:26-37)."""
import numpy as np
import pytest

from tests._parity import abs_spmv, assert_spmv_close

pytestmark = pytest.mark.gpu


def _noalloc(dA, x, y):
    import ctypes as C
    from benchsp_amd._lib import lib
    from benchsp_amd.device import ptr, stream_handle
    d = dA.desc()
    lib().bsp_csr_spmv_noalloc(C.byref(d), ptr(x), ptr(y), stream_handle())


def _ops(bsp_host):
    from benchsp_amd.device import Bsr, CsrSpmv, Dasp, Dbsr, DeviceCsr, Dia
    mats = {"27pt-64": bsp_host.stencil27_64(),
            "powerlaw-300k": bsp_host.powerlaw(300_000, 20.0, 1.2, 5000, seed=6)}
    for name, A in mats.items():
        dA = DeviceCsr.from_host(A)
        for method in ("adaptive", "scalar", "balanced", "split"):
            yield f"{name}/{method}", A, dA, CsrSpmv(dA, method)
        yield f"{name}/dasp", A, dA, Dasp(dA)
        yield f"{name}/dasp-mfma", A, dA, (lambda h: (lambda x, y: h(x, y, kernel=Dasp.MFMA)))(
            Dasp(dA))
        yield f"{name}/dbsr", A, dA, Dbsr(dA)
        yield f"{name}/dbsr-mfma", A, dA, (lambda h: (lambda x, y: h(x, y, kernel=Dbsr.MFMA)))(
            Dbsr(dA))
        B = Bsr(dA)
        for kname, k in (("mfma", Bsr.MFMA), ("mfma-balanced", Bsr.MFMA_BALANCED),
                         ("valu-balanced", Bsr.VALU_BALANCED)):
            yield f"{name}/bsr4-{kname}", A, dA, (lambda b, k: (lambda x, y: b(x, y, kernel=k)))(B, k)
        yield f"{name}/noalloc", A, dA, (lambda d: (lambda x, y: _noalloc(d, x, y)))(dA)
        if name == "27pt-64":
            yield f"{name}/dia", A, dA, Dia(dA)
    # forced column segments, every mode (XCD-pinned one launch; one launch per
    # segment; XCD-pinned sorted blocks)
    A = mats["powerlaw-300k"]
    dA = DeviceCsr.from_host(A)
    for mode in (0, 1, 2):
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


def test_spmm_and_block_spmm_replay_in_graphs(bsp_host):
    # deterministic kernels: every replay equals the eager call on the same
    # operands bit for bit (after one eager call has sized the workspaces)
    import torch
    from benchsp_amd.autograd import SparsePattern
    from benchsp_amd.device import BlockSpmm, Bsr, DeviceCsr
    rng = np.random.default_rng(4)
    A = bsp_host.graph(20_000, 30.0, 2.1, False, seed=9)
    P = SparsePattern.from_host(A)
    for dt in (torch.float32, torch.bfloat16):
        print(f"[graph] spmm {dt}", flush=True)
        vals = torch.from_numpy(A.val.astype(np.float32)).cuda().to(dt)
        X = torch.zeros(A.ncols, 64, device="cuda", dtype=dt)
        dY = torch.zeros(A.nrows, 64, device="cuda", dtype=dt)
        outs = {}

        def step():
            outs["Y"] = P.spmm(vals, X)
            outs["dX"] = P.spmm_t(vals, dY)
            outs["dv"] = P.sddmm(dY, X)
            outs["fdX"], outs["fdv"] = P.spmm_t_sddmm(vals, dY, X)
        step()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            step()
        static = dict(outs)
        for rep in range(3):
            X.copy_(torch.from_numpy(rng.normal(size=(A.ncols, 64)).astype(np.float32)).to(dt))
            dY.copy_(torch.from_numpy(rng.normal(size=(A.nrows, 64)).astype(np.float32)).to(dt))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = {k: v.clone() for k, v in static.items()}
            step()  # eager, same operands
            torch.cuda.synchronize()
            for k in got:
                assert torch.equal(got[k], outs[k]), (str(dt), k, rep)
        del g
    B16 = bsp_host.blocksparse(64, 48, 16, 700, seed=2)
    bsr = Bsr(DeviceCsr.from_host(B16), 16, 16)
    for half in (torch.float16, torch.bfloat16):
        print(f"[graph] block spmm {half}", flush=True)
        S = BlockSpmm(bsr, half)
        Bm = torch.zeros(B16.ncols, 128, device="cuda", dtype=half)
        dC = torch.zeros(B16.nrows, 128, device="cuda", dtype=half)
        outs = {}

        def bstep():
            outs["C"] = S.fwd(Bm)
            outs["dB"], outs["dv"] = S.bwd(Bm, dC)
        bstep()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            bstep()
        static = dict(outs)
        for rep in range(3):
            Bm.copy_(torch.from_numpy(rng.normal(size=(B16.ncols, 128)).astype(np.float32)).to(half))
            dC.copy_(torch.from_numpy(rng.normal(size=(B16.nrows, 128)).astype(np.float32)).to(half))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            got = {k: v.clone() for k, v in static.items()}
            bstep()
            torch.cuda.synchronize()
            for k in got:
                assert torch.equal(got[k], outs[k]), (str(half), k, rep)
        del g

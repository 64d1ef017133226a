"""Block-sparse SpMM on FP16/BF16 MFMA (the SMaT counterpart, src/smat.cu):
builder (scales, one rounding, transposed copy), forward and backward
against the FP64 oracle."""
import numpy as np
import pytest

from tests._parity import rel_fro

pytestmark = pytest.mark.gpu

STATED_TOL = 1e-2  # relative Frobenius error vs the FP64 oracle on the unquantized A
ACCUM_TOL = 2e-5   # vs FP64 on the quantized operands: FP32 accumulation only


def _bf16_bits_rne(x64):
    """IEEE RNE of FP64 values to bfloat16 bit patterns (normal range)."""
    b = np.ascontiguousarray(x64, np.float64).view(np.uint64)
    sign = (b >> np.uint64(63)).astype(np.uint32)
    exp = ((b >> np.uint64(52)) & np.uint64(0x7FF)).astype(np.int64)
    man = (b & np.uint64((1 << 52) - 1)) | np.uint64(1 << 52)  # 53-bit significand
    keep = man >> np.uint64(45)                                   # top 8 bits
    rem = man & np.uint64((1 << 45) - 1)
    half = np.uint64(1 << 44)
    up = (rem > half) | ((rem == half) & ((keep & np.uint64(1)) == np.uint64(1)))
    keep = keep + up.astype(np.uint64)
    carry = keep >> np.uint64(8)
    keep = np.where(carry > 0, keep >> np.uint64(1), keep)
    exp = exp + carry.astype(np.int64)
    e8 = exp - 1023 + 127
    bits = (sign << 15) | (e8.astype(np.uint32) << 7) | (keep.astype(np.uint32) & 0x7F)
    zero = (b & np.uint64(0x7FFFFFFFFFFFFFFF)) == 0
    return np.where(zero, sign << 15, bits).astype(np.uint16)


def _half_bits(x64, half):
    if half == "float16":
        return np.asarray(x64, np.float64).astype(np.float16).view(np.uint16)
    return _bf16_bits_rne(x64)


def _bits_to_f64(bits, half):
    if half == "float16":
        return bits.view(np.float16).astype(np.float64)
    return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def _build(A, half):
    import torch
    from benchsp_amd.device import Bsr, BlockSpmm, DeviceCsr
    bsr = Bsr(DeviceCsr.from_host(A), 16, 16)
    return bsr, BlockSpmm(bsr, getattr(torch, half))


@pytest.mark.parametrize("half", ["float16", "bfloat16"])
def test_builder_scales_rounding_transpose(oracle, bsp_host, half):
    A = bsp_host.blocksparse(40, 36, 16, 300, seed=5)
    A.val[::97] *= 1e4   # blocks of very different magnitude
    A.val[5::89] *= 1e-6
    bsr, S = _build(A, half)
    g = S.arrays()
    o = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 16, 16)
    blocks = o["val"].reshape(-1, 256)
    m = np.abs(blocks).max(axis=1)
    sc = g["scale"].astype(np.float64)
    assert np.all(np.log2(sc) == np.round(np.log2(sc)))   # powers of two
    assert np.all((m / sc >= 0.5) & (m / sc < 1.0))
    expect = _half_bits(blocks / sc[:, None], half).ravel()
    assert np.array_equal(g["hv"], expect)                 # one rounding from FP64
    hv = g["hv"].reshape(-1, 16, 16)
    assert np.array_equal(g["hvt"].reshape(-1, 16, 16), hv.transpose(0, 2, 1))


def _quantized_dense(A, g, o, half):
    """Dense FP64 matrix of the quantized operand scale_b * hv_b."""
    q = _bits_to_f64(g["hv"], half).reshape(-1, 16, 16) * g["scale"].astype(np.float64)[:, None, None]
    D = np.zeros((o["brows"] * 16, o["bcols"] * 16))
    for br in range(o["brows"]):
        for k in range(o["rowptr"][br], o["rowptr"][br + 1]):
            bc = o["colidx"][k]
            D[br * 16:(br + 1) * 16, bc * 16:(bc + 1) * 16] = q[k]
    return D[:A.nrows, :A.ncols]


@pytest.mark.parametrize("half", ["float16", "bfloat16"])
@pytest.mark.parametrize("N", [32, 64, 128, 256])
@pytest.mark.parametrize("shape", ["blocks", "ragged", "wide", "empty"])
def test_fwd_bwd_small(oracle, bsp_host, half, N, shape):
    import torch
    rng = np.random.default_rng(N)
    if shape == "blocks":
        A = bsp_host.blocksparse(24, 20, 16, 90, seed=1)
    elif shape == "empty":  # no blocks at all: C, dB zero, dvals empty
        A = bsp_host.blocksparse(3, 2, 16, 0, seed=1)
    elif shape == "wide":
        # two full block rows of 300 blocks (more than 64 per wave, so block
        # lists are read in several 64-block batches), block scales spread
        # over nine powers of two
        A = bsp_host.blocksparse(2, 300, 16, 600, seed=3)
        rows = np.repeat(np.arange(A.nrows), np.diff(A.rowptr))
        A.val *= np.exp2((A.col // 16 + 5 * (rows // 16)) % 9 - 4.0)
    else:  # sizes not multiples of 16, scattered entries
        A = bsp_host.powerlaw(100, 8.0, 1.0, 40, seed=2, ncols=90)
    bsr, S = _build(A, half)
    g = S.arrays()
    o = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 16, 16)
    td = getattr(torch, half)
    B = torch.from_numpy(rng.normal(size=(A.ncols, N)).astype(np.float32)).to(td).cuda()
    dC = torch.from_numpy(rng.normal(size=(A.nrows, N)).astype(np.float32)).to(td).cuda()
    C = S.fwd(B).cpu().numpy()
    dB, dv = S.bwd(B, dC)
    dB, dv = dB.cpu().numpy(), dv.cpu().numpy()
    B64 = B.float().cpu().numpy().astype(np.float64)
    dC64 = dC.float().cpu().numpy().astype(np.float64)
    Dq = _quantized_dense(A, g, o, half)
    assert rel_fro(C, Dq @ B64) <= ACCUM_TOL
    assert rel_fro(dB, Dq.T @ dC64) <= ACCUM_TOL
    # dvals: gradient of the unscaled block values, block order of the BSR
    ref = np.zeros_like(dv)
    dCp = np.zeros((o["brows"] * 16, N))
    dCp[:A.nrows] = dC64
    Bp = np.zeros((o["bcols"] * 16, N))
    Bp[:A.ncols] = B64
    for br in range(o["brows"]):
        for k in range(o["rowptr"][br], o["rowptr"][br + 1]):
            bc = o["colidx"][k]
            ref[k] = dCp[br * 16:(br + 1) * 16] @ Bp[bc * 16:(bc + 1) * 16].T
    assert rel_fro(dv, ref) <= ACCUM_TOL
    # error against the FP64 oracle on the original (unquantized) values
    Cref = oracle.bsr_spmm_f64(o["rowptr"], o["colidx"], o["val"], 16,
                               np.vstack([B64, np.zeros((o["bcols"] * 16 - A.ncols, N))]))
    assert rel_fro(C, Cref[:A.nrows]) <= STATED_TOL


def test_operands_not_16_byte_aligned_are_refused(bsp_host):
    # 16-byte operand loads (B, dC) and dvals stores: a view one element into
    # a buffer is an argument error, before any launch
    import ctypes as C
    import torch
    from benchsp_amd._lib import BspError, lib
    from benchsp_amd.device import ptr, stream_handle
    A = bsp_host.blocksparse(4, 4, 16, 6, seed=1)
    bsr, S = _build(A, "float16")
    N = 32
    Bb = torch.zeros(A.ncols * N + 8, dtype=torch.float16, device="cuda")
    dCb = torch.zeros(A.nrows * N + 8, dtype=torch.float16, device="cuda")
    out = torch.zeros(max(A.nrows, A.ncols) * N, dtype=torch.float32, device="cuda")
    dvb = torch.zeros(6 * 256 + 4, dtype=torch.float32, device="cuda")
    ok_B, bad_B = ptr(Bb), ptr(Bb[1:])
    with pytest.raises(BspError) as e:
        lib().bsp_bspmm_fwd(S._h, bad_B, N, ptr(out), stream_handle())
    assert e.value.status == 1
    for B_, dC_, dv_ in ((bad_B, ptr(dCb), ptr(dvb)), (ok_B, ptr(dCb[1:]), ptr(dvb)),
                         (ok_B, ptr(dCb), ptr(dvb[1:]))):
        with pytest.raises(BspError) as e:
            lib().bsp_bspmm_bwd(S._h, B_, dC_, N, ptr(out), dv_, stream_handle())
        assert e.value.status == 1
    lib().bsp_bspmm_fwd(S._h, ok_B, N, ptr(out), stream_handle())
    lib().bsp_bspmm_bwd(S._h, ok_B, ptr(dCb), N, ptr(out), ptr(dvb), stream_handle())
    torch.cuda.synchronize()


@pytest.mark.parametrize("half", ["float16", "bfloat16"])
def test_blocksparse_16k_fwd_bwd(oracle, bsp_host, half):
    import torch
    A = bsp_host.blocksparse_16k()
    print(f"[bspmm] 16384^2 blocks={A.nnz // 256} {half}", flush=True)
    bsr, S = _build(A, half)
    o = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 16, 16)
    td = getattr(torch, half)
    B64 = np.asarray(np.random.default_rng(1).normal(size=(A.ncols, 128)), np.float32)
    B = torch.from_numpy(B64).to(td).cuda()
    dC = torch.from_numpy(np.random.default_rng(2).normal(size=(A.nrows, 128)).astype(np.float32)).to(td).cuda()
    C = S.fwd(B).cpu().numpy()
    dB, dv = S.bwd(B, dC)
    Bq = B.float().cpu().numpy().astype(np.float64)
    dCq = dC.float().cpu().numpy().astype(np.float64)
    Cref = oracle.bsr_spmm_f64(o["rowptr"], o["colidx"], o["val"], 16, Bq)
    err = rel_fro(C, Cref)
    print(f"[bspmm] C rel err vs FP64 = {err:.3e}", flush=True)
    assert err <= STATED_TOL
    # dB = A^T dC: through the FP64 oracle's transposed CSR product
    v = A.val
    dBref = oracle.csr_spmm_t_f64(A.rowptr, A.col, v, dCq, A.ncols)
    assert rel_fro(dB.cpu().numpy(), dBref) <= STATED_TOL
    # dvals on a sample of blocks
    dv = dv.cpu().numpy()
    rng = np.random.default_rng(3)
    brow = np.repeat(np.arange(o["brows"]), np.diff(o["rowptr"]))
    ks = rng.choice(o["nblocks"], 200, replace=False)
    ref = np.stack([dCq[brow[k] * 16:(brow[k] + 1) * 16] @ Bq[o["colidx"][k] * 16:(o["colidx"][k] + 1) * 16].T
                    for k in ks])
    assert rel_fro(dv[ks], ref) <= ACCUM_TOL

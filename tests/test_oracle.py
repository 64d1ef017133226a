"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker:
golden fixtures from the reference's compiled dbsr.h, the reference itself
where it builds (oracle/_ref), closed forms, dense brute force, numpy's
IEEE half conversion, and structural invariants read off the reference
kernels' addressing."""
import glob
import os

import numpy as np
import pytest

from benchsp_amd.host import Csr

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _rand_csr(rng, lens, ncols):
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) if L else
                          np.zeros(0, np.int64) for L in lens]).astype(np.int32)
    return Csr(len(lens), ncols, rp, col, rng.uniform(-1, 1, int(rp[-1])))


# ---------------------------------------------------------------- SpMV -----

def test_spmv_closed_form_poisson(oracle, bsp_host):
    A = bsp_host.stencil5(31, 17)
    y = oracle.spmv_f64(A.rowptr, A.col, A.val, np.ones(A.ncols))
    # row sums of the 5-point operator: 0 inside, 1 on an edge, 2 at a corner
    i, j = np.divmod(np.arange(A.nrows), 17)
    expect = (i == 0).astype(float) + (i == 30) + (j == 0) + (j == 16)
    assert np.array_equal(y, expect)


def test_spmv_matches_dense(oracle):
    rng = np.random.default_rng(1)
    A = _rand_csr(rng, rng.integers(0, 40, 300), 257)
    x = rng.uniform(-1, 1, 257)
    y = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    yd = A.to_dense() @ x
    bound = 1e-14 * (np.abs(A.to_dense()) @ np.abs(x)) + 1e-300
    assert np.all(np.abs(y - yd) <= bound)


def test_spmv_summation_order_is_left_to_right(oracle):
    # one row whose value depends on the order: (1e16 + 1) - 1e16 = 0 left
    # to right, 1 right to left
    A = Csr(1, 3, np.array([0, 3]), np.array([0, 1, 2], np.int32), np.array([1e16, 1.0, -1e16]))
    assert oracle.spmv_f64(A.rowptr, A.col, A.val, np.ones(3))[0] == 0.0


def test_spmv_threads_bit_identical(oracle, bsp_host):
    A = bsp_host.powerlaw(20000, 15.0, 1.2, 400, seed=3)
    x = bsp_host.uniform(A.ncols, 5)
    y1 = oracle.spmv_f64(A.rowptr, A.col, A.val, x)
    y8 = oracle.spmv_f64(A.rowptr, A.col, A.val, x, threads=8)
    assert np.array_equal(y1, y8)


def test_f16_conversion_matches_numpy(oracle):
    rng = np.random.default_rng(0)
    vals = np.concatenate([
        rng.normal(0, 1, 2000), rng.normal(0, 1e-5, 2000), rng.uniform(-7e4, 7e4, 2000),
        [0.0, -0.0, 65504.0, 65519.99, 65520.0, 1e-8, 2.0**-24, 2.0**-25, 3 * 2.0**-26,
         1.0 + 2.0**-11, 1.0 + 3 * 2.0**-11, np.inf, -np.inf]]).astype(np.float32)
    ours = np.array([oracle.f32_to_f16_bits(v) for v in vals], np.uint16)
    assert np.array_equal(ours, vals.astype(np.float16).view(np.uint16))
    halfs = np.arange(0, 1 << 16, dtype=np.uint16)
    halfs = halfs[(halfs & 0x7c00) != 0x7c00]  # finite values
    back = np.array([oracle.f16_bits_to_f32(h) for h in halfs[::7]], np.float32)
    assert np.array_equal(back, halfs[::7].view(np.float16).astype(np.float32))


def test_spmv_f16_matches_numpy_emulation(oracle):
    rng = np.random.default_rng(4)
    A = _rand_csr(rng, rng.integers(0, 12, 60), 50)
    x = rng.normal(0, 1, 50).astype(np.float32)
    v32 = A.val.astype(np.float32)
    y = oracle.spmv_f16(A.rowptr, A.col, v32, x)
    vh = v32.astype(np.float16).astype(np.float32)
    xh = x.astype(np.float16).astype(np.float32)
    expect = np.zeros(A.nrows, np.float32)
    for i in range(A.nrows):
        s = np.float32(0)
        for j in range(A.rowptr[i], A.rowptr[i + 1]):
            s = np.float32(s + np.float32(vh[j] * xh[A.col[j]]))
        expect[i] = np.float32(np.float16(s))
    assert np.array_equal(y, expect)


def _f32_row_sums(A, v32, x32, mode):
    """numpy emulation of a row loop over FP32 inputs. mode "ref" is
    ref.cpp:14-24 (product rounded to FP32, then added, left to right);
    "fma" contracts each product into the add; "f64" accumulates in FP64 and
    rounds once. The last two are the plausible mistakes an oracle could make."""
    y = np.zeros(A.nrows, np.float32)
    for i in range(A.nrows):
        s32, s64 = np.float32(0), 0.0
        for j in range(A.rowptr[i], A.rowptr[i + 1]):
            p = float(v32[j]) * float(x32[A.col[j]])  # exact: 24+24 bits < 53
            if mode == "ref":
                s32 = np.float32(s32 + np.float32(p))
            elif mode == "fma":
                s32 = np.float32(float(s32) + p)  # one rounding of s + v*x
            else:
                s64 += p
        y[i] = s64 if mode == "f64" else s32
    return y


def test_spmv_f32_matches_numpy_emulation(oracle):
    rng = np.random.default_rng(5)
    A = _rand_csr(rng, rng.integers(0, 30, 200), 90)
    x = rng.normal(0, 1, 90).astype(np.float32)
    v32 = A.val.astype(np.float32)
    y = oracle.spmv_f32(A.rowptr, A.col, v32, x)
    assert y.dtype == np.float32
    assert np.array_equal(y, _f32_row_sums(A, v32, x, "ref"))


def test_spmv_f32_rounds_each_product_and_each_add(oracle):
    # row 0: -(1+2^-11)*1 + (1+2^-12)^2. The product is 1 + 2^-11 + 2^-24,
    # a tie that FP32 rounds to 1 + 2^-11, so the ref loop gives 0; an
    # FMA-contracted or FP64-accumulated loop gives 2^-24.
    # row 1: 2^24 + 1 - 2^24 is 0 in FP32 left to right, 1 in FP64.
    a = np.float32(1 + 2.0**-12)
    A = Csr(2, 4, np.array([0, 2, 5]), np.array([0, 1, 2, 3, 2], np.int32),
            np.array([-(1 + 2.0**-11), float(a), 2.0**24, 1.0, -(2.0**24)]))
    x = np.array([1.0, a, 1.0, 1.0], np.float32)
    v32 = A.val.astype(np.float32)
    y = oracle.spmv_f32(A.rowptr, A.col, v32, x)
    assert np.array_equal(y, np.array([0.0, 0.0], np.float32))
    assert np.array_equal(_f32_row_sums(A, v32, x, "ref"), y)
    # the crafted rows tell the two mistaken loops apart from the ref loop
    assert _f32_row_sums(A, v32, x, "fma")[0] == np.float32(2.0**-24)
    assert np.array_equal(_f32_row_sums(A, v32, x, "f64"), np.array([2.0**-24, 1.0], np.float32))


# ---------------------------------------------------------------- split ----

def test_split_maps(oracle):
    rng = np.random.default_rng(2)
    lens = rng.integers(0, 300, 500)
    lens[:5] = [128, 129, 0, 127, 1000]
    A = _rand_csr(rng, lens, 2000)
    sm, lm = oracle.split(A.rowptr, 128)
    assert np.array_equal(sm, np.nonzero(lens <= 128)[0])
    assert np.array_equal(lm, np.nonzero(lens > 128)[0])


# ---------------------------------------------------------------- BCSR -----

@pytest.mark.parametrize("tm,tk", [(4, 4), (3, 3), (2, 5), (16, 16), (1, 1)])
def test_bcsr_reconstructs_dense(oracle, tm, tk):
    rng = np.random.default_rng(tm * 31 + tk)
    A = _rand_csr(rng, rng.integers(0, 9, 37), 41)  # sizes not multiples of the tile
    b = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, tm, tk)
    assert b["brows"] == -(-37 // tm) and b["bcols"] == -(-41 // tk)
    D = np.zeros((b["brows"] * tm, b["bcols"] * tk))
    blocks = b["val"].reshape(-1, tm, tk)
    for br in range(b["brows"]):
        cols = b["colidx"][b["rowptr"][br]:b["rowptr"][br + 1]]
        assert np.all(np.diff(cols) > 0), "block columns ascending"
        for k in range(b["rowptr"][br], b["rowptr"][br + 1]):
            bc = b["colidx"][k]
            assert np.any(blocks[k] != 0) or True
            D[br * tm:(br + 1) * tm, bc * tk:(bc + 1) * tk] = blocks[k]
    assert np.array_equal(D[:37, :41], A.to_dense())
    assert not D[37:, :].any() and not D[:, 41:].any()
    # a block is stored iff it holds an entry
    occ = {(r // tm, c // tk) for r in range(37) for c in A.col[A.rowptr[r]:A.rowptr[r + 1]]}
    stored = {(br, int(bc)) for br in range(b["brows"])
              for bc in b["colidx"][b["rowptr"][br]:b["rowptr"][br + 1]]}
    assert occ == stored


def test_bsr4_map_and_balance(oracle, bsp_host):
    A = bsp_host.stencil3d(9, 10, 11, 27, seed=2)
    b = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 4, 4)
    # map bit r*4+c set exactly where the CSR has an entry (amgT.cu:781-834)
    mp = np.zeros(b["nblocks"], np.uint16)
    for r in range(A.nrows):
        for j in range(A.rowptr[r], A.rowptr[r + 1]):
            c = A.col[j]
            lo, hi = b["rowptr"][r // 4], b["rowptr"][r // 4 + 1]
            k = lo + np.searchsorted(b["colidx"][lo:hi], c // 4)
            mp[k] |= np.uint16(1 << ((r % 4) * 4 + c % 4))
    assert np.array_equal(b["map"][:-1], mp) and b["map"][-1] == 0
    lens = np.diff(b["rowptr"])
    w = -(-lens // 64)
    assert np.array_equal(b["rowptr_by_warp"], np.concatenate([[0], np.cumsum(w)]))
    assert np.array_equal(b["rowidx_by_warp"], np.repeat(np.arange(b["brows"]), w))
    assert b["avg_nnz"] == A.nnz / b["nblocks"]
    assert abs(b["stand"] - np.sqrt(np.mean((lens - lens.mean()) ** 2))) < 1e-12


def test_bsr4_stencil27_block_count(oracle, bsp_host):
    A = bsp_host.stencil27_64()
    b = oracle.bcsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, 4, 4)
    # closed form: a block row is 4 consecutive x cells of one (z, y) line.
    # For each neighbour line (dz, dy) its columns span x cells 4b-1..4b+4,
    # i.e. 3 aligned 4-blocks for the 14 inner b and 2 for b = 0 and 15:
    # 46 blocks per line pair. Summed over z, the number of valid dz is
    # 62*3 + 2*2 = 190, the same for y: 46 * 190 * 190.
    assert b["nblocks"] == 46 * 190 * 190 == 1_660_600


# ---------------------------------------------------------------- DBSR -----

def _check_dbsr(d, g):
    assert d["warpnum"] == int(g["warpnum"])
    for k, gk in [("recorder", "recorder"), ("blcptr", "blcptr"), ("colidx", "colidx"),
                  ("val", "blcval"), ("map", "map"), ("prefix", "prefix")]:
        assert np.array_equal(d[k], g[gk]), k


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "dbsr_*.npz"))),
                         ids=lambda p: os.path.basename(p)[5:-4])
def test_dbsr_golden(oracle, path):
    g = np.load(path)
    d = oracle.dbsr(int(g["nrows"]), int(g["ncols"]), g["rowptr"].astype(np.int64), g["col"],
                    g["val"])
    assert not d["overflow"]
    _check_dbsr(d, g)


def test_dbsr_matches_compiled_reference(oracle):
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(11)
    checked = 0
    for trial in range(40):
        n = int(rng.integers(1, 120))
        lens = rng.integers(0, int(rng.choice([6, 30, 300])), n)
        A = _rand_csr(rng, lens, 700)
        d = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
        if d["overflow"]:
            continue  # undefined in the reference
        r = oracle.ref_dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
        _check_dbsr(d, {"warpnum": r["warpnum"], "recorder": r["recorder"],
                        "blcptr": r["blcptr"], "colidx": r["colidx"], "blcval": r["val"],
                        "map": r["map"], "prefix": r["prefix"]})
        checked += 1
    assert checked >= 20


def test_dbsr_unsorted_matches_compiled_reference(oracle):
    # csrTodbsr + balanceDbsr on the rows as given (no recorder sort)
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(12)
    checked = 0
    for trial in range(40):
        n = int(rng.integers(1, 120))
        lens = rng.integers(0, int(rng.choice([6, 30, 300])), n)
        A = _rand_csr(rng, lens, 700)
        d = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, sort_rows=False)
        assert np.array_equal(d["recorder"], np.arange(A.nrows))
        if d["overflow"]:
            continue  # undefined in the reference
        r = oracle.ref_dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val, sort_rows=False)
        _check_dbsr(d, {"warpnum": r["warpnum"], "recorder": r["recorder"],
                        "blcptr": r["blcptr"], "colidx": r["colidx"], "blcval": r["val"],
                        "map": r["map"], "prefix": r["prefix"]})
        checked += 1
    assert checked >= 20


def test_recorder_self_matches_compiled_reference(oracle):
    # csrRecorderSelf: the CSR permuted by the oracle's recorder order
    if not oracle.ref_available():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = np.random.default_rng(13)
    for trial in range(10):
        lens = rng.integers(0, 40, int(rng.integers(1, 200)))
        A = _rand_csr(rng, lens, 300)
        rp, col, val = oracle.ref_recorder_self(A.nrows, A.ncols, A.rowptr, A.col, A.val)
        rec = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)["recorder"]
        assert np.array_equal(np.diff(rp), lens[rec])
        for i, o in enumerate(rec):
            s, e = A.rowptr[o], A.rowptr[o + 1]
            assert np.array_equal(col[rp[i]:rp[i + 1]], A.col[s:e])
            assert np.array_equal(val[rp[i]:rp[i + 1]], A.val[s:e])


def test_dbsr_balance_overflow_is_defined(oracle):
    # block-row widths [1,1,4,17,18]: pass 1 packs all 41 blocks in one warp,
    # pass 2 starts from now=41 and opens a second warp at the last block row
    rng = np.random.default_rng(0)
    A = _rand_csr(rng, [0, 1, 2, 3, 4, 5, 9, 16, 17, 33, 70, 0, 2, 64, 65, 8, 1], 90)
    d = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
    assert d["overflow"] and d["warpnum"] == 2
    assert np.array_equal(d["prefix"], [0, 4, 5])


def test_dbsr_reconstructs_csr(oracle, bsp_host):
    A = bsp_host.powerlaw(5000, 9.0, 1.0, 200, seed=8)
    d = oracle.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
    rec = d["recorder"]
    lens = np.diff(A.rowptr)
    assert np.array_equal(np.sort(rec), np.arange(A.nrows))
    key = lens[rec] * (A.nrows + 1) + rec
    assert np.all(np.diff(key) > 0), "sorted by (len, index)"
    vals = d["val"].reshape(-1, 16)
    cols = d["colidx"].reshape(-1, 16)
    for i in range(A.nrows):
        o = rec[i]
        L = lens[o]
        z = np.arange(L)
        blk = d["blcptr"][i // 4] + z // 4
        slot = (i % 4) * 4 + z % 4
        assert np.array_equal(vals[blk, slot], A.val[A.rowptr[o]:A.rowptr[o + 1]])
        assert np.array_equal(cols[blk, slot], A.col[A.rowptr[o]:A.rowptr[o + 1]])
    assert int(np.count_nonzero(d["map"].astype(np.int64) != 0)) > 0
    bits = np.unpackbits(d["map"].view(np.uint8)).sum()
    assert bits == A.nnz


# ---------------------------------------------------------------- DASP -----

def _dasp_rows(D, A):
    """Re-reads every row's entries out of the DASP arrays the way the
    reference kernel addresses them (src/dasp.cu:405-812) and returns
    {row: (cols, vals)} with padding dropped."""
    out = {}
    o = D["order_rid"]
    rl, rb = D["row_long"], D["row_block"]
    lens = np.diff(A.rowptr)
    # long rows: contiguous from long_rpt[i]*64 (dasp.cu:421-444)
    for i in range(rl):
        r = o[i]
        s = D["long_rpt"][i] * 64
        out[r] = (D["long_cid"][s:s + lens[r]], D["long_val"][s:s + lens[r]])
    # row blocks: slab i holds 32 entries, row g of the block at g*4+t
    # (fragA index tID_in_group + groupID*4, +32 per step: dasp.cu:496-503)
    for b in range(D["blocknum"]):
        s, e = D["block_ptr"][b], D["block_ptr"][b + 1]
        nslab = (e - s) // 32
        for g in range(8):
            k = b * 8 + g
            if k >= rb:
                continue
            idx = s + np.arange(nslab)[:, None] * 32 + g * 4 + np.arange(4)[None, :]
            cols, vals = D["reg_cid"][idx.ravel()], D["reg_val"][idx.ravel()]
            r = o[rl + k]
            keep = min(lens[r], nslab * 4)
            i0, i1 = D["irreg_rpt"][k], D["irreg_rpt"][k + 1]
            out[r] = (np.concatenate([cols[:keep], D["irreg_cid"][i0:i1]]),
                      np.concatenate([vals[:keep], D["irreg_val"][i0:i1]]))
            assert not np.any(vals[keep:]), "padding is zero"
    base = rl + rb
    s1, c13, s3, s4, s2 = (D["short_row_1"], D["common_13"], D["short_row_3"],
                           D["short_row_4"], D["short_row_2"])
    sv, sc = D["short_val"], D["short_cid"]
    for i in range(s1):
        out[o[base + i]] = (sc[i:i + 1], sv[i:i + 1])
    for t in range(c13 // 8):
        for j in range(8):
            p = s1 + t * 32 + j * 4
            out[o[base + s1 + t * 16 + j]] = (sc[p:p + 1], sv[p:p + 1])
            out[o[base + s1 + t * 16 + 8 + j]] = (sc[p + 1:p + 4], sv[p + 1:p + 4])
    p34 = s1 + D["fill0_nnz_short13"]
    ob = base + s1 + 2 * c13
    for i in range(s3):
        out[o[ob + i]] = (sc[p34 + 4 * i:p34 + 4 * i + 3], sv[p34 + 4 * i:p34 + 4 * i + 3])
    for i in range(s4):
        q = p34 + 4 * (s3 + i)
        out[o[ob + s3 + i]] = (sc[q:q + 4], sv[q:q + 4])
    p22 = p34 + D["fill0_nnz_short34"]
    for j in range(s2):
        t, jj = divmod(j, 16)
        q = p22 + t * 32 + (jj % 8) * 4 + (jj // 8) * 2
        out[o[ob + s3 + s4 + j]] = (sc[q:q + 2], sv[q:q + 2])
    return out


@pytest.mark.parametrize("case", ["powerlaw", "stencil", "mixed"])
def test_dasp_layout_holds_every_row(oracle, bsp_host, case):
    rng = np.random.default_rng(5)
    if case == "powerlaw":
        A = bsp_host.powerlaw(30000, 8.0, 1.3, 900, seed=6)
    elif case == "stencil":
        A = bsp_host.stencil3d(11, 9, 13, 27, seed=1)
    else:
        lens = np.concatenate([np.full(300, 1), np.full(200, 3), np.full(77, 2), np.full(50, 4),
                               np.full(9, 0), rng.integers(5, 300, 400), [256, 255, 1000]])
        rng.shuffle(lens)
        A = _rand_csr(rng, lens, 3000)
    D = oracle.dasp(A.rowptr, A.col, A.val)
    n = A.nrows
    assert np.array_equal(np.sort(D["order_rid"]), np.arange(n))
    lens = np.diff(A.rowptr)
    assert D["row_zero"] == np.sum(lens == 0)
    assert D["row_long"] == np.sum(lens >= 256)
    assert D["row_block"] == np.sum((lens >= 5) & (lens < 256))
    # row-block rows sorted by length, longest first, ties by index
    rbr = D["order_rid"][D["row_long"]:D["row_long"] + D["row_block"]]
    key = -lens[rbr].astype(np.int64) * (n + 1) + rbr
    assert np.all(np.diff(key) > 0)
    rows = _dasp_rows(D, A)
    assert len(rows) == n - D["row_zero"]
    for r, (c, v) in rows.items():
        assert np.array_equal(c, A.col[A.rowptr[r]:A.rowptr[r + 1]]), r
        assert np.array_equal(v, A.val[A.rowptr[r]:A.rowptr[r + 1]]), r


def test_dasp_common13_threshold(oracle):
    rng = np.random.default_rng(3)
    for n13, expect in [(127, 0), (128, 128), (135, 128), (200, 200)]:
        A = _rand_csr(rng, [1] * n13 + [3] * (n13 + 5), 500)
        D = oracle.dasp(A.rowptr, A.col, A.val)
        assert D["common_13"] == expect, n13
        assert D["short_row_1"] == n13 - expect and D["short_row_3"] == n13 + 5 - expect


@pytest.mark.parametrize("L,regular,irreg", [(7, 64, 0), (6, 32, 2), (8, 64, 0), (5, 32, 1)])
def test_dasp_slab_threshold(oracle, L, regular, irreg):
    # 8 rows of length L: slab 2 holds 8*(L-4) entries, kept iff >= 0.75*32
    rng = np.random.default_rng(L)
    A = _rand_csr(rng, [L] * 8, 100)
    D = oracle.dasp(A.rowptr, A.col, A.val)
    assert D["blocknum"] == 4  # ceil(8/8) rounded up to a multiple of 4
    assert D["block_ptr"][1] == regular
    assert np.all(np.diff(D["irreg_rpt"]) == irreg)


# ----------------------------------------------------------- blockinfo ----

def test_blockinfo_closed_forms(oracle, bsp_host):
    # 27-pt 64^3, b = 4: blocks as in test_bsr4_stencil27_block_count; packed:
    # a group is 4 x-cells of one (z, y) line whose longest row has 3a
    # entries, a = (valid dz)(valid dy) in {9, 6, 4} on 62^2, 4*62 and 4
    # lines: 16 groups per line times ceil(3a/4) in {7, 5, 3}
    A = bsp_host.stencil27_64()
    blocks, packed = oracle.blockinfo(A.nrows, A.ncols, A.rowptr, A.col, 4)
    assert blocks == 46 * 190 * 190
    assert packed == 16 * (62 * 62 * 7 + 4 * 62 * 5 + 4 * 3) == 450_560


@pytest.mark.parametrize("b", [1, 3, 4, 16])
def test_blockinfo_brute_force(oracle, b):
    rng = np.random.default_rng(b)
    A = _rand_csr(rng, rng.integers(0, 30, 77), 91)
    blocks, packed = oracle.blockinfo(A.nrows, A.ncols, A.rowptr, A.col, b)
    rows = np.repeat(np.arange(A.nrows), A.row_lengths())
    pos = np.arange(A.nnz) - A.rowptr[rows]
    assert blocks == len({(r // b, c // b) for r, c in zip(rows, A.col)})
    assert packed == len({(r // b, k // b) for r, k in zip(rows, pos)})
    L = A.row_lengths()
    assert packed == sum(int(max(-(-L[g:g + b] // b))) if L[g:g + b].size else 0
                         for g in range(0, A.nrows, b))

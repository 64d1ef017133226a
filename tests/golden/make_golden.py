"""Writes tests/golden/dbsr_*.npz: inputs and outputs of the reference's own
DBSR builder (csrTodbsrWithRecorder, include/dbsr.h:208-253, compiled from the
reference tree into oracle/_ref/libdbsr_ref.so by oracle/Makefile).

Run in a checkout where the reference tree exists:
    make -C oracle && python tests/golden/make_golden.py
Each fixture holds the CSR input (int32 rowptr/col, float64 val) and every
structural array the reference returns, so tests pin the oracle's
restatement without the reference present.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def cases():
    from benchsp_amd import host
    from benchsp_amd.host import Csr
    rng = np.random.default_rng(2024)

    def from_lens(lens, ncols):
        rp = np.zeros(len(lens) + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        col = np.concatenate([np.sort(rng.choice(ncols, L, replace=False)) if L else
                              np.zeros(0, np.int64) for L in lens]).astype(np.int32)
        return Csr(len(lens), ncols, rp, col, rng.uniform(-1, 1, int(rp[-1])))

    yield "poisson5_8x8", host.stencil5(8, 8)
    yield "stencil27_4", host.stencil3d(4, 4, 4, 27, seed=7)
    yield "mixed_lengths", from_lens([0, 1, 2, 3, 4, 5, 9, 16, 17, 33, 70, 0, 2, 64, 65, 8, 1], 90)
    # block rows wider than 64 blocks: balanceDbsr's multi-warp branch
    yield "long_rows", from_lens([3, 600, 5, 257, 0, 0, 1, 300, 12, 2, 260, 7, 9], 1000)
    yield "powerlaw_600", host.powerlaw(600, 12.0, 1.2, 300, seed=9)
    # ties in length across non-adjacent rows (sort key (len, idx))
    yield "ties", from_lens([5, 3, 5, 3, 5, 3, 0, 5, 3, 1, 1, 5], 40)


def main():
    from oracle import oracle as O
    if not O.ref_available():
        sys.exit("oracle/_ref/libdbsr_ref.so missing: run `make -C oracle` with the reference tree")
    for name, A in cases():
        if O.dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)["overflow"]:
            # the reference's balanceDbsr writes past its warp list here
            print(name, "skipped: reference behaviour undefined (warp list overflow)")
            continue
        r = O.ref_dbsr(A.nrows, A.ncols, A.rowptr, A.col, A.val)
        np.savez_compressed(
            os.path.join(HERE, f"dbsr_{name}.npz"),
            nrows=A.nrows, ncols=A.ncols, rowptr=A.rowptr.astype(np.int32), col=A.col,
            val=A.val, brows=r["brows"], nblocks=r["nblocks"], warpnum=r["warpnum"],
            recorder=r["recorder"], blcptr=r["blcptr"], colidx=r["colidx"], blcval=r["val"],
            map=r["map"], prefix=r["prefix"])
        print(name, A.nrows, A.nnz, "blocks", r["nblocks"], "warps", r["warpnum"])


if __name__ == "__main__":
    main()

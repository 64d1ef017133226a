"""MatrixMarket reader and COO -> CSR assembly (bsp_mtx_read, bsp_coo_to_csr),
checked against scipy's reader on the same files."""
import numpy as np
import pytest
import scipy.io
import scipy.sparse

from benchsp_amd import host


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


def _same_as_scipy(path, A):
    ref = scipy.sparse.csr_matrix(scipy.io.mmread(path))
    ref.sum_duplicates()
    ours = scipy.sparse.csr_matrix((A.val, A.col, A.rowptr), shape=(A.nrows, A.ncols))
    ours.sum_duplicates()
    assert ours.shape == ref.shape
    assert (abs(ours - ref)).max() == 0 if ref.nnz else ours.nnz == 0


def _sorted_rows(A):
    for r in range(A.nrows):
        c = A.col[A.rowptr[r]:A.rowptr[r + 1]]
        assert np.all(np.diff(c) >= 0)


@pytest.mark.parametrize("field,sym", [("real", "general"), ("integer", "general"),
                                       ("pattern", "general"), ("real", "symmetric"),
                                       ("real", "skew-symmetric"), ("pattern", "symmetric")])
def test_read_matches_scipy(tmp_path, field, sym):
    rng = np.random.default_rng(3)
    n = 40
    entries = set()
    while len(entries) < 150:
        i, j = int(rng.integers(1, n + 1)), int(rng.integers(1, n + 1))
        if sym != "general" and i < j:
            i, j = j, i
        if sym == "skew-symmetric" and i == j:
            continue
        entries.add((i, j))
    lines = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% a comment", "%",
             f"{n} {n} {len(entries)}"]
    for i, j in sorted(entries, key=lambda e: rng.random()):
        if field == "pattern":
            lines.append(f"{i} {j}")
        elif field == "integer":
            lines.append(f"{i} {j} {int(rng.integers(-9, 10))}")
        else:
            lines.append(f"{i} {j} {rng.normal():.17g}")
    path = _write(tmp_path, "m.mtx", "\n".join(lines) + "\n")
    A = host.read_mtx(path)
    _sorted_rows(A)
    expect = len(entries) if sym == "general" else 2 * len(entries) - sum(i == j for i, j in entries)
    assert A.nnz == expect
    _same_as_scipy(path, A)


def test_read_rectangular_and_duplicates(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real general\n3 5 5\n"
            "3 5 1.5\n1 2 -2\n3 5 2.5\n1 1 4e-3\n2 4 7\n")
    A = host.read_mtx(_write(tmp_path, "d.mtx", text))
    assert (A.nrows, A.ncols, A.nnz) == (3, 5, 5)
    assert list(A.rowptr) == [0, 2, 3, 5]
    assert list(A.col) == [0, 1, 3, 4, 4]
    assert list(A.val) == [4e-3, -2.0, 7.0, 1.5, 2.5]  # duplicates kept, file order


def test_coo_to_csr_order_and_errors():
    A = host.coo_to_csr(3, 4, [2, 0, 2, 0, 1], [1, 3, 0, 3, 2], [1.0, 2.0, 3.0, 4.0, 5.0])
    assert list(A.rowptr) == [0, 2, 3, 5]
    assert list(A.col) == [3, 3, 2, 0, 1]
    assert list(A.val) == [2.0, 4.0, 5.0, 3.0, 1.0]
    P = host.coo_to_csr(2, 2, [1, 0], [0, 1])
    assert list(P.val) == [1.0, 1.0]
    from benchsp_amd._lib import BspError
    with pytest.raises(BspError):
        host.coo_to_csr(2, 2, [2], [0])


@pytest.mark.parametrize("text", [
    "%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n",
    "%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n",
    "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n",
    # an entry count the file cannot hold is refused before any array is sized
    "%%MatrixMarket matrix coordinate real symmetric\n2000000000 2000000000 1000000000\n"
    "1 1 1.0\n",
    # symmetric storage mirrors (i, j) to (j, i): a non-square header is
    # refused (the mirror of (1, 3) would be row 3 of a 2-row matrix)
    "%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 3 1.0\n",
    "%%MatrixMarket matrix coordinate pattern skew-symmetric\n3 2 1\n3 1\n",
    "not a matrix market file\n"])
def test_read_rejects(tmp_path, text):
    from benchsp_amd._lib import BspError
    with pytest.raises(BspError):
        host.read_mtx(_write(tmp_path, "bad.mtx", text))

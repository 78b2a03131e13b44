"""Matrix Market reader / writer and COO->CSR (readCOOMatrix, convertCOO2CSR).

Behaviour follows the Matrix Market format with the reference's banner checks
and return codes (utils/basicMatrix.cpp:84-132); its index-base, padding and
square-ify quirks are deliberately not reproduced."""
import numpy as np
import pytest


def _w(tmp_path, text, name="m.mtx"):
    p = tmp_path / name
    p.write_text(text)
    return p


def test_general_real_one_based_and_duplicates(P, tmp_path):
    p = _w(tmp_path, "%%MatrixMarket matrix coordinate real general\n% comment\n3 4 5\n"
                     "1 1 1.5\n3 4 -2\n1 1 0.25\n2 2 3e-1\n1 3 7\n")
    A = P.read_mtx(p)
    assert (A.numRows, A.numCols, A.nnz) == (3, 4, 4)
    assert A.rowOffsets.tolist() == [0, 2, 3, 4]
    assert A.cols.tolist() == [0, 2, 1, 3]
    assert A.vals.tolist() == [1.75, 7.0, 0.3, -2.0]


def test_symmetric_pattern_integer_skew_complex(P, tmp_path):
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate real symmetric\n3 3 3\n1 1 2\n2 1 -1\n3 2 4\n"))
    D = A.to_scipy().toarray()
    assert np.array_equal(D, [[2, -1, 0], [-1, 0, 4], [0, 4, 0]])
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n", "p.mtx"))
    assert A.vals.tolist() == [1.0, 1.0]
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate integer general\n2 2 1\n2 2 -7\n", "i.mtx"))
    assert A.vals.tolist() == [-7.0]
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 3\n", "s.mtx"))
    assert np.array_equal(A.to_scipy().toarray(), [[0, -3], [3, 0]])
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 2 5 9\n", "c.mtx"))
    assert A.vals.tolist() == [5.0]


def test_empty_rows_and_zero_nnz(P, tmp_path):
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate real general\n5 5 2\n2 1 1\n4 5 2\n"))
    assert A.rowOffsets.tolist() == [0, 0, 1, 1, 2, 2]
    A = P.read_mtx(_w(tmp_path, "%%MatrixMarket matrix coordinate real general\n4 4 0\n", "z.mtx"))
    assert A.nnz == 0 and A.rowOffsets.tolist() == [0] * 5


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarketX matrix coordinate real general\n1 1 1\n1 1 1\n", -1),
    ("%%MatrixMarket vector coordinate real general\n1 1 1\n1 1 1\n", -2),
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -3),
    ("%%MatrixMarket matrix coordinate real general\nfoo\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n3 3 4\n1 1 1\n2 2 2\n", -6),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n4 1 1\n", -6),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 0 1\n", -6),
    ("%%MatrixMarket matrix coordinate real general\n3 3 1\n1 x 1\n", -6),
])
def test_reader_errors(P, tmp_path, text, code):
    with pytest.raises(P.PavaError) as e:
        P.read_mtx(_w(tmp_path, text))
    assert e.value.code == code


def test_missing_file(P, tmp_path):
    with pytest.raises(P.PavaError) as e:
        P.read_mtx(tmp_path / "nope.mtx")
    assert e.value.code == -7


def test_write_read_round_trip_bit_exact(P, tmp_path):
    A = P.rmat(10, 8, seed=2)
    p = tmp_path / "r.mtx"
    P.write_mtx(p, A)
    B = P.read_mtx(p)
    assert np.array_equal(A.rowOffsets, B.rowOffsets)
    assert np.array_equal(A.cols, B.cols)
    assert np.array_equal(A.vals, B.vals)


def test_bench_file_cpu_filters(P, tmp_path):
    # too small: skipped before any device work, as benchFile does (main.cpp:82-99)
    p = _w(tmp_path, "%%MatrixMarket matrix coordinate real general\n3 3 1\n1 1 1\n")
    assert P.lib().pava_bench_file(str(p).encode(), 1) == -1

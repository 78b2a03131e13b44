"""Fast input path: the parallel Matrix Market parser must give the same CSR
as a straightforward parse, and the binary cache must round-trip exactly and be
preferred when newer than the text file."""
import os
import time

import numpy as np
import pytest


def _write_mm(path, m, n, rows, cols, vals, sym="general", comments_every=0):
    with open(path, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real {sym}\n% generated\n{m} {n} {len(rows)}\n")
        for k, (r, c, v) in enumerate(zip(rows, cols, vals)):
            if comments_every and k % comments_every == 0:
                f.write("% a comment line inside the body\n")
            f.write(f"{int(r) + 1} {int(c) + 1} {float(v)!r}\n")


def test_parallel_parse_matches_scipy_large(P, tmp_path):
    """A file above the parallel threshold (4 MiB), with duplicates and
    comment lines between entries, against scipy's COO->CSR summation."""
    import scipy.sparse as sp
    rng = np.random.default_rng(1)
    m, n, L = 20000, 15000, 260000
    rows = rng.integers(0, m, L)
    cols = rng.integers(0, n, L)
    vals = rng.normal(size=L)
    p = tmp_path / "big.mtx"
    _write_mm(p, m, n, rows, cols, vals, comments_every=997)
    assert os.path.getsize(p) > (1 << 22)
    A = P.read_mtx(p)
    S = sp.coo_matrix((vals, (rows, cols)), shape=(m, n)).tocsr()
    S.sum_duplicates()
    S.sort_indices()
    assert np.array_equal(A.rowOffsets, S.indptr)
    assert np.array_equal(A.cols, S.indices)
    assert np.allclose(A.vals, S.data, rtol=1e-15, atol=1e-15)


def test_parallel_parse_symmetric_and_errors(P, tmp_path):
    rng = np.random.default_rng(2)
    m = 30000
    L = 240000
    r = rng.integers(0, m, L)
    c = rng.integers(0, m, L)
    lo, hi = np.minimum(r, c), np.maximum(r, c)
    v = rng.normal(size=L)
    p = tmp_path / "sym.mtx"
    _write_mm(p, m, m, hi, lo, v, sym="symmetric")
    A = P.read_mtx(p)
    D = A.to_scipy()
    assert abs(D - D.T).max() == 0
    assert A.nnz <= 2 * L
    # an out-of-range entry deep in the body (in a later parallel piece)
    q = tmp_path / "bad.mtx"
    rows = list(hi)
    rows[L - 10] = m + 5
    _write_mm(q, m, m, rows, lo, v)
    with pytest.raises(P.PavaError) as e:
        P.read_mtx(q)
    assert e.value.code == -6 and f"entry {L - 9} " in str(e.value)


def test_binary_cache_round_trip_and_preference(P, tmp_path):
    A = P.rmat(11, 8, seed=7)
    mtx = tmp_path / "g.mtx"
    P.write_mtx(mtx, A)
    cache = tmp_path / "g.mtx.csrbin"
    P.write_csr_binary(cache, A)
    B = P.read_csr_binary(cache)
    for f in ("rowOffsets", "cols", "vals"):
        assert np.array_equal(getattr(A, f), getattr(B, f))
    # the cache is used when newer than the .mtx: make it distinguishable
    Z = P.CSR(A.numRows, A.numCols, A.rowOffsets, A.cols, np.zeros(A.nnz))
    P.write_csr_binary(cache, Z)
    now = time.time()
    os.utime(mtx, (now - 100, now - 100))
    assert not P.read_mtx(mtx).vals.any()
    # an older cache is ignored
    os.utime(cache, (now - 1000, now - 1000))
    assert np.array_equal(P.read_mtx(mtx).vals, A.vals)
    with pytest.raises(P.PavaError):
        P.read_csr_binary(mtx)  # not a cache file

"""FP64 serial parity of the CVR-64 order (CPU emulation).

Rows that the CVR-64 layout splits (steals, chunk boundaries) are summed as
per-segment FMA chains; the reference sums every row as one FMA chain
(utils/basicMatrix.cpp:538-556). The parity pass recomputes, in serial order,
every split row whose running error bound cannot show agreement within
parity_tol. These tests pin the bound (it must hold wherever rounding is
systematic and the error approaches it) and the pass (every row within
parity_tol of the serial sum, each row equal to either its CVR-order or its
serial value bit for bit)."""
import numpy as np
import pytest

from matrices import csr_from_dense_pattern, edge_cases, random_csr, split_rows

TOL = 1e-12
U = 2.0 ** -53


@pytest.fixture(scope="module")
def mats(P):
    return edge_cases(P)


def _three(P, O, A, x, S, tol=TOL):
    lay = O.cvr64_build(A, S, 0)
    ycvr, bound, _ = O.cvr64_spmv(A, lay, x, parity_tol=0.0, details=True)
    ypar, _, nrc = O.cvr64_spmv(A, lay, x, parity_tol=tol, details=True)
    yser = O.spmv_fma(A, x)
    return lay, ycvr, bound, ypar, nrc, yser


def _rel(a, b):
    d = np.abs(a - b)
    den = np.abs(b)
    return np.where(den > 0, d / np.where(den > 0, den, 1.0), np.where(d > 0, np.inf, 0.0))


@pytest.mark.parametrize("S", [8, 64])
def test_parity_pass_on_edge_cases(P, O, mats, S):
    for name, A in mats.items():
        x = P.fill_uniform(A.numCols, 7)
        lay, ycvr, bound, ypar, nrc, yser = _three(P, O, A, x, S)
        # every row is its CVR-order value or its serial value, bit for bit
        assert np.all((ypar == ycvr) | (ypar == yser)), name
        assert np.all(_rel(ypar, yser) <= TOL), name
        # the bound holds for the plain CVR order; unsplit rows have none
        assert np.all(np.abs(ycvr - yser) <= bound), name
        split = split_rows(A, lay, S)
        unsplit = np.ones(A.numRows, bool)
        unsplit[list(split)] = False
        assert np.all(bound[unsplit] == 0) and np.all(ycvr[unsplit] == yser[unsplit]), name


def test_bound_holds_under_systematic_rounding(P, O):
    # identical terms (v = 1, x = 0.1): rounding errors of an FMA chain all
    # lean the same way, the case where the running bound is nearly reached
    rows = [list(range(L)) for L in (4000, 2500, 1200, 700, 333, 64, 65, 3)]
    A = csr_from_dense_pattern(P, len(rows), 4000, rows)
    A.vals[:] = 1.0
    x = np.full(4000, 0.1)
    for S in (8, 16, 64):
        lay, ycvr, bound, ypar, nrc, yser = _three(P, O, A, x, S)
        d = np.abs(ycvr - yser)
        assert np.all(d <= bound)
        # the bound is nearly reached here (about a third of it on the worst row)
        assert np.max(d[bound > 0] / bound[bound > 0]) > 0.2
        assert np.all(_rel(ypar, yser) <= TOL)


def test_bound_holds_on_random_shapes(P, O):
    rng = np.random.default_rng(5)
    worst = 0.0
    for _ in range(12):
        A = random_csr(P, rng, 6000)
        x = rng.uniform(-1, 1, A.numCols)
        S = int(rng.choice([8, 16, 64]))
        lay, ycvr, bound, ypar, nrc, yser = _three(P, O, A, x, S)
        d = np.abs(ycvr - yser)
        assert np.all(d <= bound)
        m = bound > 0
        if m.any():
            worst = max(worst, float(np.max(d[m] / bound[m])))
        assert np.all((ypar == ycvr) | (ypar == yser))
        assert np.all(_rel(ypar, yser) <= TOL)
    assert worst > 0


def cancelling(P, rng, m=400, n=20000):
    """Rows of 40..3000 random terms whose last value is set so that the exact
    row sum is tiny: strongly cancelling (kappa 1e6 .. 1e12) split rows."""
    lens = rng.integers(40, 3000, m)
    rows = [np.sort(rng.choice(n, L, replace=False)) for L in lens]
    A = csr_from_dense_pattern(P, m, n, [list(r) for r in rows], seed=3)
    x = rng.uniform(-1, 1, n)
    rp = A.rowOffsets
    for r in range(m):
        b, e = int(rp[r]), int(rp[r + 1])
        last = e - 1
        partial = float(np.dot(A.vals[b:last], x[A.cols[b:last]]))
        xl = x[A.cols[last]]
        if abs(xl) > 1e-3:
            A.vals[last] = -partial / xl * (1 + rng.uniform(-1e-9, 1e-9))
    return A, x


def test_cancelling_split_rows_are_recomputed(P, O):
    rng = np.random.default_rng(1)
    A, x = cancelling(P, rng)
    for S in (8, 64):
        lay, ycvr, bound, ypar, nrc, yser = _three(P, O, A, x, S)
        bad = _rel(ycvr, yser) > TOL
        assert bad.sum() > 10, "the plain CVR order must miss 1e-12 on some rows here"
        assert np.all(ypar[bad] == yser[bad])
        assert np.all(_rel(ypar, yser) <= TOL)
        assert nrc >= bad.sum()


def test_parity_off_is_the_cvr_order(P, O, mats):
    for name, A in mats.items():
        x = P.fill_uniform(A.numCols, 3)
        lay = O.cvr64_build(A, 16, 0)
        y0, _, n0 = O.cvr64_spmv(A, lay, x, parity_tol=0.0, details=True)
        assert n0 == 0
        assert np.array_equal(y0, O.cvr64_spmv(A, lay, x, parity_tol=0.0)), name


def test_long_rows_use_the_probabilistic_bound(P, O):
    # rows longer than 4,096 elements: the rigorous bound alone would
    # recompute them; the probabilistic one (lambda = 10) accepts those whose
    # result is not cancelling (here all terms are positive, kappa = 1), and
    # the result stays within 1e-12
    rng = np.random.default_rng(8)
    rows = [list(np.sort(rng.choice(60000, L, replace=False))) for L in (5000, 9000, 20000, 30000)]
    A = csr_from_dense_pattern(P, len(rows), 60000, rows, seed=2)
    A.vals[:] = rng.uniform(0.0, 1.0, A.nnz)
    x = rng.uniform(0.5, 1.0, 60000)  # no cancellation: kappa = 1
    lay, ycvr, bound, ypar, nrc, yser = _three(P, O, A, x, 64)
    lens = np.diff(A.rowOffsets)
    # rigorous bound above 1e-12 |y| on the longest rows, yet not recomputed
    assert np.any(bound > TOL * np.abs(ycvr))
    assert nrc < len(rows)
    assert np.all(_rel(ypar, yser) <= TOL)
    assert lens.max() > 4096

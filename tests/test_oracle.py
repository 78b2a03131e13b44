"""Pins the CPU oracle against the reference's known answer, closed forms,
exact arithmetic and an independent library (scipy.sparse fixtures)."""
import json
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


def _csr(P, m, n, rp, ci, va):
    return P.CSR(m, n, np.array(rp), np.array(ci), np.array(va, dtype=np.float64))


def test_esb_5x5_known_answer(P, O):
    g = json.loads((GOLD / "esb_5x5.json").read_text())
    A = _csr(P, g["numRows"], g["numCols"], g["rowOffsets"], g["cols"], g["vals"])
    x = np.array(g["x"], dtype=np.float64)
    for f in (O.spmv_fma, O.spmv_muladd, O.spmv_dd):
        assert f(A, x).tolist() == g["y"]
    # the product driver's own verification SpMV agrees too
    assert P.reference_spmv(A, x).tolist() == g["y"]


def test_laplacian_closed_forms(P, O):
    nx, ny = 37, 23
    A = P.laplacian2d(nx, ny)
    assert A.nnz == 5 * nx * ny - 2 * nx - 2 * ny
    y = O.spmv_fma(A, np.ones(A.numCols))
    ix = np.arange(nx * ny) % nx
    iy = np.arange(nx * ny) // nx
    deg = (ix > 0).astype(int) + (ix < nx - 1) + (iy > 0) + (iy < ny - 1)
    assert np.array_equal(y, 4.0 - deg)
    # affine x: interior rows vanish exactly
    x = 3.0 + 2.0 * ix - 5.0 * iy
    y = O.spmv_fma(A, x.astype(np.float64))
    interior = (deg == 4)
    assert np.all(y[interior] == 0.0)
    assert not np.all(y[~interior] == 0.0)


def test_dense_brute_force_integers(P, O):
    rng = np.random.default_rng(1)
    for trial in range(20):
        m, n = rng.integers(1, 30, 2)
        D = rng.integers(-9, 10, (m, n)) * (rng.random((m, n)) < 0.4)
        rp = np.concatenate([[0], np.cumsum((D != 0).sum(1))])
        ci = np.nonzero(D)[1]
        va = D[D != 0].astype(np.float64)
        A = _csr(P, int(m), int(n), rp, ci, va)
        x = rng.integers(-5, 6, n).astype(np.float64)
        want = D.astype(np.float64) @ x
        assert np.array_equal(O.spmv_fma(A, x), want)
        assert np.array_equal(O.spmv_muladd(A, x), want)


def test_fma_vs_muladd_distinguished(P, O):
    # row: (-1)*1 + (1+2^-30)^2 ; exact = 2^-29 + 2^-60 (representable)
    v = 1.0 + 2.0 ** -30
    A = _csr(P, 1, 2, [0, 2], [0, 1], [-1.0, v])
    x = np.array([1.0, v])
    assert O.spmv_fma(A, x)[0] == 2.0 ** -29 + 2.0 ** -60
    assert O.spmv_muladd(A, x)[0] == 2.0 ** -29


def _exact_rows(A, x):
    out = []
    for r in range(A.numRows):
        s = Fraction(0)
        for k in range(A.rowOffsets[r], A.rowOffsets[r + 1]):
            s += Fraction(float(A.vals[k])) * Fraction(float(x[A.cols[k]]))
        out.append(s)
    return out


def test_dd_is_near_exact_and_fma_within_higham_bound(P, O):
    rng = np.random.default_rng(7)
    m, n = 40, 300
    lens = rng.integers(1, 200, m)
    rp = np.concatenate([[0], np.cumsum(lens)])
    ci = np.concatenate([np.sort(rng.choice(n, L, replace=False)) for L in lens])
    va = rng.uniform(-1, 1, rp[-1])
    A = _csr(P, m, n, rp, ci, va)
    x = rng.uniform(-1, 1, n)
    exact = _exact_rows(A, x)
    ydd, yf = O.spmv_dd(A, x), O.spmv_fma(A, x)
    sabs = O.row_abs(A, x)
    u = 2.0 ** -53
    for r in range(m):
        ex = float(exact[r])
        # double-double: within 2 ulp of the correctly rounded value
        assert abs(ydd[r] - ex) <= 2 * np.spacing(abs(ex)) + 4 * (lens[r] * u) ** 2 * sabs[r]
        # sequential fma: Higham's gamma_n bound
        n_r = int(lens[r])
        assert abs(Fraction(float(yf[r])) - exact[r]) <= Fraction(n_r * u / (1 - n_r * u)) * Fraction(float(sabs[r]))


def test_row_abs_matches_numpy(P, O):
    d = np.load(GOLD / "scipy_spmv.npz")
    m, n = d["m2_shape"]
    A = _csr(P, int(m), int(n), d["m2_rp"], d["m2_ci"], d["m2_va"])
    x = d["m2_x"]
    S = A.to_scipy()
    want = abs(S) @ np.abs(x)
    assert np.allclose(O.row_abs(A, x), want, rtol=1e-14, atol=0)


def test_against_scipy_fixtures(P, O):
    d = np.load(GOLD / "scipy_spmv.npz")
    for k in range(int(d["count"])):
        m, n = d[f"m{k}_shape"]
        A = _csr(P, int(m), int(n), d[f"m{k}_rp"], d[f"m{k}_ci"], d[f"m{k}_va"])
        x = d[f"m{k}_x"]
        y = O.spmv_fma(A, x)
        sabs = O.row_abs(A, x)
        lens = np.diff(A.rowOffsets)
        tol = lens * 2.0 ** -52 * sabs + 1e-300
        assert np.all(np.abs(y - d[f"m{k}_y"]) <= tol)


def test_check_results_one_percent_rule(O):
    y_ref = np.array([100.0, 100.0, -50.0, 0.0, 1.0])
    y = np.array([100.9, 101.1, -50.6, 0.0, 1.0])
    assert O.check_results(y, y_ref) == 2  # rows 1 and 2 exceed 1%
    assert O.check_results(y_ref, y_ref) == 0


def test_check_results_product_matches_oracle(P, O):
    rng = np.random.default_rng(3)
    y_ref = rng.normal(size=1000)
    y = y_ref * (1 + rng.normal(scale=0.01, size=1000))
    assert P.check_results(y, y_ref) == O.check_results(y, y_ref)


def test_omp_equals_serial(P, O):
    A = P.rmat(12, 8, seed=5)
    x = P.fill_uniform(A.numCols, 1)
    assert np.array_equal(O.spmv_omp(A, x, 4), O.spmv_fma(A, x))


def test_driver_reference_spmv_equals_oracle_on_strided_and_f32_x(P, O):
    """The driver's reference SpMV (pava's referenceSpMV) equals the oracle bit
    for bit, also when x arrives as a strided column or in FP32."""
    A = P.synthetic(3000, P.RL_POISSON, 12.0, col_struct=P.COL_MIXED, half_band=100, band_frac=0.7, seed=9)
    X = np.random.default_rng(1).uniform(-1, 1, (A.numCols, 4))
    for j in range(4):
        col = X[:, j]  # strided view
        assert np.array_equal(P.reference_spmv(A, col), O.spmv_fma(A, np.ascontiguousarray(col)))
    x32 = X[:, 0].astype(np.float32)
    assert np.array_equal(P.reference_spmv(A, x32), O.spmv_fma(A, x32.astype(np.float64)))

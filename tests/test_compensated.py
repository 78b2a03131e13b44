"""Pins the oracle's FP32 / BF16 CVR-64 accumulation (the GPU path equals it
bit for bit) against exact arithmetic.

Each lane segment is accumulated with Dot2 (Ogita, Rump and Oishi, "Accurate
sum and dot product", SIAM J. Sci. Comput. 26(6), 2005): TwoProduct through
an FP32 fma and TwoSum in FP32, their error terms summed alongside; the
segment's value is s + c in FP64, segments of a split row are added in FP64
and the row is rounded once to FP32. Dot2's bound gives, with u = 2^-24 and n
the row length,

    |y - exact| <= u |exact| + (gamma_n(u)^2 + 64 * 2^-53) * sum |a x| * (1 + u)

where exact is the rational row sum of the stored values times the FP32 x.
Values here have 8 significant bits, so FP32 and BF16 store them exactly. The
rows cancel strongly (kappa ~ 1e3 to 1e5): a plain FP32 sum, or Dot2 without
the TwoProduct term, breaks the bound, as checked below.
"""
from fractions import Fraction

import numpy as np
import pytest

U32 = 2.0 ** -24


def _cancelling_matrix(P, seed):
    rng = np.random.default_rng(seed)
    lens = np.concatenate([rng.integers(2, 60, 1500) * 2, rng.integers(100, 400, 60) * 2, [1500, 3000]])
    rng.shuffle(lens)
    m, n = lens.size, 4096
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols, vals = [], []
    for L in lens:
        cols.append(np.sort(rng.choice(n, L, replace=False)))
        k = rng.integers(64, 256, L // 2)             # 8 significant bits
        v = np.empty(L)
        v[0::2], v[1::2] = k / 128.0, -k / 128.0      # pairs cancel exactly on x = 1
        vals.append(v)
    A = P.CSR(m, n, rp, np.concatenate(cols).astype(np.int32), np.concatenate(vals))
    x = (1.0 + rng.uniform(-1e-3, 1e-3, n)).astype(np.float32)
    return A, x


def _exact(A, x):
    xf = [Fraction(float(v)) for v in x]
    out = []
    for r in range(A.numRows):
        s = Fraction(0)
        for k in range(A.rowOffsets[r], A.rowOffsets[r + 1]):
            s += Fraction(float(A.vals[k])) * xf[A.cols[k]]
        out.append(s)
    return out


def _bound(A, x, exact):
    lens = np.diff(A.rowOffsets).astype(np.float64)
    g = lens * U32 / (1 - lens * U32)
    sabs = np.array([np.sum(np.abs(A.vals[A.rowOffsets[r]:A.rowOffsets[r + 1]]) *
                            np.abs(x[A.cols[A.rowOffsets[r]:A.rowOffsets[r + 1]]].astype(np.float64)))
                     for r in range(A.numRows)])
    ex = np.array([float(e) for e in exact])
    return U32 * np.abs(ex) + (g * g + 64 * 2.0 ** -53) * sabs * (1 + U32), ex, sabs


def _err(y, exact):
    return np.array([abs(Fraction(float(v)) - e) for v, e in zip(y, exact)], dtype=object).astype(np.float64)


@pytest.mark.parametrize("prec,S", [(1, 16), (2, 16), (1, 64)])
def test_dot2_accumulation_within_its_bound(P, O, prec, S):
    A, x = _cancelling_matrix(P, 7)
    exact = _exact(A, x)
    bound, ex, sabs = _bound(A, x, exact)
    kappa = sabs / np.abs(ex)
    assert np.median(kappa) > 1e3
    lay = O.cvr64_build(A, S, prec)
    assert lay["split_n"][: lay["nchunks"]].sum() > 0  # split rows and rows across chunks take part
    y = O.cvr64_spmv(A, lay, x)
    assert y.dtype == np.float32
    err = _err(y, exact)
    assert np.all(err <= bound), (np.max(err / bound), int(np.argmax(err / bound)))


def _plain_fp32(A, x):
    y = np.zeros(A.numRows, np.float32)
    for r in range(A.numRows):
        s = np.float32(0)
        for k in range(A.rowOffsets[r], A.rowOffsets[r + 1]):
            s = np.float32(s + np.float32(A.vals[k]) * x[A.cols[k]])
        y[r] = s
    return y


def _dot2_without_two_product(A, x):
    y = np.zeros(A.numRows, np.float32)
    for r in range(A.numRows):
        s, c = np.float32(0), np.float32(0)
        for k in range(A.rowOffsets[r], A.rowOffsets[r + 1]):
            p = np.float32(np.float32(A.vals[k]) * x[A.cols[k]])
            t = np.float32(s + p)
            z = np.float32(t - s)
            c = np.float32(c + np.float32(np.float32(s - np.float32(t - z)) + np.float32(p - z)))
            s = t
        y[r] = np.float32(np.float64(s) + np.float64(c))
    return y


def test_bound_rejects_uncompensated_variants(P):
    """The bound has teeth on these rows: an FP32 recursive sum, and Dot2 with
    the product errors dropped, both leave it on many rows."""
    A, x = _cancelling_matrix(P, 7)
    exact = _exact(A, x)
    bound, _, _ = _bound(A, x, exact)
    plain = _err(_plain_fp32(A, x), exact) > bound
    no_tp = _err(_dot2_without_two_product(A, x), exact) > bound
    assert plain.mean() > 0.5
    assert no_tp.mean() > 0.2

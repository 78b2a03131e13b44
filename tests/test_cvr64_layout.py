"""CVR-64 layout emulation (CPU): invariants and summation-order parity with
the serial oracle. The GPU conversion is compared against this emulation
array-for-array in test_gpu_cvr.py."""
import numpy as np
import pytest

from cvr_check import reconstruct
from matrices import edge_cases, split_rows

STEPS = [8, 16, 64, 256]


@pytest.fixture(scope="module")
def mats(P):
    return edge_cases(P)


@pytest.mark.parametrize("S", STEPS)
def test_layout_rebuilds_csr(P, O, mats, S):
    for name, A in mats.items():
        lay = O.cvr64_build(A, S, 0)
        errs = reconstruct(A, lay, S)
        assert not errs, (name, errs[:5])
        nc, L, _ = O.cvr64_sizes(A.numRows, A.nnz, S)
        assert L >= A.nnz and L % 512 == 0 or A.nnz == 0
        assert lay["info"][0] <= 160  # split slots per chunk


@pytest.mark.parametrize("S", STEPS)
def test_emulated_spmv_vs_oracle(P, O, mats, S):
    for name, A in mats.items():
        x = P.fill_uniform(A.numCols, 42)
        lay = O.cvr64_build(A, S, 0)
        y = O.cvr64_spmv(A, lay, x)
        yref = O.spmv_fma(A, x)
        sabs = O.row_abs(A, x)
        split = split_rows(A, lay, S)
        lens = np.diff(A.rowOffsets)
        for r in range(A.numRows):
            if r not in split:
                assert y[r] == yref[r], (name, r)
        # split rows: blocked summation, within gamma_(n+64) * sum|a x| of the exact sum
        bound = (lens + 64) * 2.0 ** -53 * sabs * 2
        assert np.all(np.abs(y - yref) <= bound), name


def test_every_lane_finishes_on_last_step(P, O, mats):
    for name, A in mats.items():
        S = 16
        lay = O.cvr64_build(A, S, 0)
        nc, L, _ = O.cvr64_sizes(A.numRows, A.nnz, S)
        for c in range(nc):
            last = (min((c + 1) * 64 * S, L) // 64 - 1) * 64
            assert np.all(lay["cols"][last:last + 64] < 0), (name, c)


def test_segment_count_equals_pieces_plus_steals(P, O):
    # pure row stream with no stealing: one segment per non-empty row piece
    A = P.CSR(64 * 16, 64 * 16, np.arange(64 * 16 + 1), np.arange(64 * 16), np.ones(64 * 16))
    lay = O.cvr64_build(A, 16, 0)
    assert lay["nchunks"] == 1 and lay["seg_count"][0] == 1024 and lay["split_n"][0] == 0
    # a single row of 64*S elements in one chunk: 1 owner + 63 thieves, all split
    S = 8
    A = P.CSR(1, 512, np.array([0, 512]), np.arange(512), np.ones(512))
    lay = O.cvr64_build(A, S, 0)
    assert lay["seg_count"][0] == 64 and lay["split_n"][0] == 1 and lay["info"][0] == 64
    assert lay["split_tab"][0] == 0  # the row itself (code = row id 0)


def test_low_precision_layout_matches_values(P, O):
    A = P.synthetic(2000, P.RL_POISSON, 15.0, col_struct=P.COL_BANDED, half_band=50, seed=2)
    for prec in (1, 2):
        lay = O.cvr64_build(A, 16, prec)
        assert not reconstruct(A, lay, 16)
        l64 = O.cvr64_build(A, 16, 0)
        v = l64["vals"].astype(np.float32)
        if prec == 1:
            assert np.array_equal(lay["vals"], v)
        else:
            got = (lay["vals"].astype(np.uint32) << 16).view(np.float32)
            assert np.all(np.abs(got - v) <= np.abs(v) * 2.0 ** -8)


def test_fp32_emulation_error_vs_fp64(P, O):
    A = P.stencil27(12, 0.01, 0)
    x = P.fill_uniform(A.numCols, 0)
    lay = O.cvr64_build(A, 64, 1)
    y32 = O.cvr64_spmv(A, lay, x.astype(np.float32)).astype(np.float64)
    y64 = O.spmv_fma(A, x)
    sabs = O.row_abs(A, x)
    # inputs rounded to FP32 (2^-24 relative each): error <= ~2^-23 sum|a x|
    assert np.all(np.abs(y32 - y64) <= 2.0 ** -22 * sabs + 1e-30)

"""FP64 serial parity on the GPU (needs an MI355X).

With the default parity_tol (1e-12), every FP64 row the CVR-64 order splits
and whose running error bound cannot show agreement with the serial FMA sum
(utils/basicMatrix.cpp:538-556) is recomputed as one serial chain, inside the
SpMV (rows within a chunk) or by the parity fix-up kernel (rows spanning
chunks). The GPU result must equal the oracle's emulation of the same pass bit
for bit, recompute the same number of rows, and be within 1e-12 of the serial
sum on every row; parity_tol = 0 must give the plain CVR order.
"""
import sys

import numpy as np
import pytest

from matrices import csr_from_dense_pattern, edge_cases, random_csr
from test_parity import cancelling

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    d = np.abs(a - b)
    den = np.abs(b)
    return np.where(den > 0, d / np.where(den > 0, den, 1.0), np.where(d > 0, np.inf, 0.0))


def _check(P, O, A, x, name, **kw):
    with P.CvrHip(A, **kw) as h:
        inf = h.info()
        y = h.spmv_host(x)
        n_gpu = h.last_recompute_count()
    S = inf["steps_per_chunk"]
    ye, _, n_emu = O.cvr64_spmv(A, O.cvr64_build(A, S, 0), x, parity_tol=inf["parity_tol"], details=True)
    assert np.array_equal(y, ye), (name, kw)
    assert n_gpu == n_emu, (name, kw, n_gpu, n_emu)
    if inf["parity_tol"] > 0:
        assert np.all(_rel(y, O.spmv_fma(A, x)) <= inf["parity_tol"]), (name, kw)
    return n_gpu, inf


@pytest.mark.parametrize("S", [8, 64, 0])
def test_parity_edge_cases(P, O, S):
    for name, A in edge_cases(P).items():
        x = P.fill_uniform(A.numCols, 7)
        _say(f"parity edge {name} S={S}")
        _check(P, O, A, x, name, steps_per_chunk=S)


@pytest.mark.parametrize("opts", [dict(steps_per_chunk=8), dict(steps_per_chunk=64), dict(steps_per_chunk=64, x_hot_kb=64),
                                  dict(steps_per_chunk=16, col_index_bits=16), dict(steps_per_chunk=64, x_hot_kb=64,
                                                                                    x_lds_kb=0)])
def test_parity_cancelling_rows(P, O, opts):
    A, x = cancelling(P, np.random.default_rng(1))
    _say(f"parity cancelling {opts}")
    n, inf = _check(P, O, A, x, "cancelling", **opts)
    assert n > 10  # rows the plain CVR order gets wrong by more than 1e-12


def test_parity_off_is_the_cvr_order(P, O):
    A, x = cancelling(P, np.random.default_rng(2))
    with P.CvrHip(A, steps_per_chunk=16, parity_tol=0.0) as h:
        assert h.info()["parity_tol"] == 0.0
        y = h.spmv_host(x)
        assert h.last_recompute_count() == 0
    lay = O.cvr64_build(A, 16, 0)
    assert np.array_equal(y, O.cvr64_spmv(A, lay, x, parity_tol=0.0))
    assert np.any(_rel(y, O.spmv_fma(A, x)) > TOL)


def test_parity_long_rows_across_chunks(P, O):
    # rows of 5,000-30,000 elements span several chunks: closed by the parity
    # fix-up kernel, with the probabilistic bound beyond 4,096 elements; one
    # cancelling long row must be recomputed there
    rng = np.random.default_rng(8)
    lens = (5000, 9000, 20000, 30000, 7000)
    rows = [list(np.sort(rng.choice(60000, L, replace=False))) for L in lens]
    A = csr_from_dense_pattern(P, len(rows), 60000, rows, seed=2)
    A.vals[:] = rng.uniform(0.0, 1.0, A.nnz)
    x = rng.uniform(0.5, 1.0, 60000)
    b, e = int(A.rowOffsets[4]), int(A.rowOffsets[5])
    A.vals[b:e] *= np.where(np.arange(e - b) % 2 == 0, 1.0, -1.0)  # row 4 cancels
    partial = float(np.dot(A.vals[b:e - 1], x[A.cols[b:e - 1]]))
    A.vals[e - 1] = -partial / x[A.cols[e - 1]] * (1 + 1e-10)
    for S in (16, 64):
        _say(f"parity long rows S={S}")
        n, inf = _check(P, O, A, x, "long", steps_per_chunk=S)
        assert n >= 1 and inf["n_fixup_rows"] >= len(lens) - 1


def test_parity_random_shapes(P, O):
    rng = np.random.default_rng(21)
    for i in range(16):
        A = random_csr(P, rng, 20000)
        x = rng.uniform(-1, 1, A.numCols)
        S = int(rng.choice([8, 24, 64, 256]))
        _say(f"parity random {i} m={A.numRows} nnz={A.nnz} S={S}")
        _check(P, O, A, x, f"random{i}", steps_per_chunk=S, x_hot_kb=int(rng.choice([-1, 0, 32])))


def test_parity_device_create_and_graph(P, O):
    import torch
    A, x = cancelling(P, np.random.default_rng(3))
    rp = torch.from_numpy(A.rowOffsets.astype(np.int32)).cuda()
    ci = torch.from_numpy(A.cols.astype(np.int32)).cuda()
    va = torch.from_numpy(A.vals).cuda()
    torch.cuda.synchronize()
    h = P.CvrHip.from_device(A.numRows, A.numCols, A.nnz, rp.data_ptr(), ci.data_ptr(), va.data_ptr(),
                             steps_per_chunk=16)
    del rp, ci, va
    try:
        S = h.info()["steps_per_chunk"]
        want = O.cvr64_spmv(A, O.cvr64_build(A, S, 0), x)
        dx = torch.from_numpy(x).cuda()
        dy = torch.empty(A.numRows, dtype=torch.float64, device="cuda")
        st = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            h.spmv(dx.data_ptr(), dy.data_ptr(), st.cuda_stream)
        for rep in range(3):
            dy.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(dy.cpu().numpy(), want), rep
    finally:
        h.close()


@pytest.mark.parametrize("hot", [False, True])
def test_speculative_chains_both_outcomes_repeated_and_graphed(P, O, hot):
    """Rows of at least 4,096 elements spanning chunks have their serial sum
    started at the beginning of the SpMV (plain kernel, no LDS table), before
    their bound is known. Rows of positive terms pass the bound (the chain is
    abandoned and the CVR-order sum stays); cancelling rows fail it (the
    chain's sum is used, whichever of the chain and the row's last chunk
    finishes first). Every call, with new x each time and under graph replay,
    equals the oracle's emulation bit for bit and recomputes the same rows:
    the per-row state words are left cleared for the next call. With hot-x
    staged (hot), the CSR copy the chains read carries the hot renumbering,
    and their hot columns are read from the hot-x copy."""
    import torch
    rng = np.random.default_rng(31)
    n = 60000
    lens = [5000, 9000, 20000, 4096, 4500, 12000, 30, 7, 6000, 4100]
    rows = [list(np.sort(rng.choice(n, L, replace=False))) for L in lens]
    A = csr_from_dense_pattern(P, len(rows), n, rows, seed=6)
    rp = A.rowOffsets
    # rows 0-2 and 8: positive terms (pass); the others: random signs, with
    # the last value set so that the row cancels (fail)
    pos = {0, 1, 2, 8}
    x0 = rng.uniform(0.5, 1.0, n)
    for r in range(len(rows)):
        b, e = int(rp[r]), int(rp[r + 1])
        if r in pos:
            A.vals[b:e] = rng.uniform(0.1, 1.0, e - b)
        else:
            A.vals[b:e] = rng.uniform(-1.0, 1.0, e - b)
            if e - b > 1:
                part = float(np.dot(A.vals[b:e - 1], x0[A.cols[b:e - 1]]))
                A.vals[e - 1] = -part / x0[A.cols[e - 1]] * (1 + 1e-9)
    S = 8
    opts = dict(x_hot_kb=64, x_lds_kb=0) if hot else dict(x_hot_kb=0)
    with P.CvrHip(A, steps_per_chunk=S, **opts) as h:
        inf = h.info()
        assert inf["n_lds_cols"] == 0 and inf["parity_tol"] > 0
        assert (inf["n_hot_cols"] > 0) == hot
        assert inf["n_spec_rows"] == sum(1 for L in lens if L >= 4096), inf
        lay = O.cvr64_build(A, S, 0)
        dx = torch.empty(n, dtype=torch.float64, device="cuda")
        dy = torch.empty(A.numRows, dtype=torch.float64, device="cuda")
        xs = [x0] + [x0 * rng.uniform(0.9, 1.1, n) for _ in range(3)]

        def want(x):
            ye, _, ne = O.cvr64_spmv(A, lay, x, parity_tol=inf["parity_tol"], details=True)
            return ye, ne

        for it, x in enumerate(xs + xs):
            dx.copy_(torch.from_numpy(x))
            h.spmv(dx.data_ptr(), dy.data_ptr())
            torch.cuda.synchronize()
            ye, ne = want(x)
            _say(f"speculative chains, call {it}: {h.last_recompute_count()} rows recomputed")
            assert np.array_equal(dy.cpu().numpy(), ye), it
            assert h.last_recompute_count() == ne, it
            assert np.all(_rel(dy.cpu().numpy(), O.spmv_fma(A, x)) <= TOL)
        s = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            h.spmv(dx.data_ptr(), dy.data_ptr(), s.cuda_stream)
        for it, x in enumerate(xs):
            dx.copy_(torch.from_numpy(x))
            dy.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert np.array_equal(dy.cpu().numpy(), want(x)[0]), ("replay", it)

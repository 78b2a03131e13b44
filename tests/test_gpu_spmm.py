"""GPU multi-RHS SpMM and the batched persistent launch (needs an MI355X).

Reference: k independent FP64 oracle SpMVs on the same (FP32-valued) inputs.
The GPU accumulates exact FP32 products in FP64 and rounds once, so each
output is within one FP32 rounding (2^-24 relative) plus the FP64 summation
bound 2 (n+64) 2^-53 sum|a x| of the oracle."""
import sys

import numpy as np
import pytest

from matrices import edge_cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _f32_matrix(A):
    A.vals = A.vals.astype(np.float32).astype(np.float64)
    return A


def _oracle_mm(O, A, X):
    return np.stack([O.spmv_fma(A, X[:, j].astype(np.float64)) for j in range(X.shape[1])], axis=1)


def check_queue_counts(b):
    """Every work item of the last batched launch was handed out exactly once:
    pops per queue range = range size + one empty pop per workgroup."""
    counts, nq = b.queue_counts()
    inf = b.info()
    total, grid = inf["items"], inf["grid"]
    assert nq == min(grid, 8)
    for q in range(8):
        want = total * (q + 1) // nq - total * q // nq + grid if q < nq else 0
        assert counts[q] == want, (q, counts, total, grid)


def _bound(O, A, X, yref):
    lens = np.diff(A.rowOffsets)[:, None]
    sabs = np.stack([O.row_abs(A, np.abs(X[:, j]).astype(np.float64)) for j in range(X.shape[1])], axis=1)
    return 2.0 ** -24 * np.abs(yref) + 2 * (lens + 64) * 2.0 ** -53 * sabs + 1e-38


@pytest.mark.parametrize("k", [1, 5, 8, 16, 32])
def test_spmm_single_matrix(P, O, k):
    import torch
    rng = np.random.default_rng(k)
    for name, A in edge_cases(P).items():
        if A.nnz == 0:
            continue
        _say(f"spmm {name} k={k}")
        A = _f32_matrix(A)
        X = rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)
        yref = _oracle_mm(O, A, X)
        with P.CvrHip(A, steps_per_chunk=16, precision=P.PREC_F32) as h:
            dX = torch.from_numpy(X).cuda()
            dY = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            h.spmm(k, dX.data_ptr(), dY.data_ptr())
            torch.cuda.synchronize()
            Y = dY.cpu().numpy().astype(np.float64)
        assert not np.isnan(Y).any(), name
        assert np.all(np.abs(Y - yref) <= _bound(O, A, X, yref)), name


def test_spmm_fp64_handle(P, O):
    import torch
    A = _f32_matrix(P.rmat(13, 8, seed=4))
    X = np.random.default_rng(0).uniform(-1, 1, (A.numCols, 8)).astype(np.float32)
    with P.CvrHip(A, steps_per_chunk=64) as h:
        dX = torch.from_numpy(X).cuda()
        dY = torch.empty((A.numRows, 8), dtype=torch.float32, device="cuda")
        h.spmm(8, dX.data_ptr(), dY.data_ptr())
        torch.cuda.synchronize()
        Y = dY.cpu().numpy().astype(np.float64)
    yref = _oracle_mm(O, A, X)
    assert np.all(np.abs(Y - yref) <= _bound(O, A, X, yref))


@pytest.mark.parametrize("k", [8, 32])
def test_batched_persistent_and_graph_replay(P, O, k):
    import torch
    rng = np.random.default_rng(100 + k)
    mats = []
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e4))))
        A = P.synthetic(n, P.RL_POISSON, float(rng.uniform(5, 60)), col_struct=P.COL_MIXED, half_band=1000,
                        band_frac=0.7, seed=1000 + i)
        mats.append(_f32_matrix(A))
    # 16- and 32-bit column layouts mixed in one batch
    hs = [P.CvrHip(A, steps_per_chunk=32, precision=P.PREC_F32, col_index_bits=16 if i % 2 else 32, x_hot_kb=0)
          for i, A in enumerate(mats)]
    b = P.Batch(hs, k)
    Xs = [torch.from_numpy(rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)).cuda() for A in mats]
    Ys = [torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda") for A in mats]
    b.set_operands([x.data_ptr() for x in Xs], [y.data_ptr() for y in Ys])
    _say("batch", b.info())
    s = torch.cuda.Stream()
    b.spmm(s.cuda_stream)
    s.synchronize()
    check_queue_counts(b)
    first = [y.cpu().numpy().astype(np.float64) for y in Ys]
    for A, X, Y in zip(mats, Xs, first):
        yref = _oracle_mm(O, A, X.cpu().numpy())
        assert np.all(np.abs(Y - yref) <= _bound(O, A, X.cpu().numpy(), yref))
    # graph: capture once, replay several times with eager work in between
    b.capture(s.cuda_stream)
    for rep in range(4):
        for y in Ys:
            y.fill_(float("nan"))
        torch.cuda.synchronize()
        b.replay(s.cuda_stream)
        s.synchronize()
        check_queue_counts(b)
        for Y0, y in zip(first, Ys):
            assert np.array_equal(Y0, y.cpu().numpy().astype(np.float64)), rep
    b.close()
    for h in hs:
        h.close()


@pytest.mark.parametrize("bits", [16, 32])
@pytest.mark.parametrize("k", [3, 8, 12, 16, 32])
def test_spmm_bit_exact_vs_cvr_order(P, O, k, bits):
    """FP32 products are exact in FP64, so every output is the FP32 rounding of
    the FP64 CVR-order sum (the SpMV emulation at the same chunk size, column by
    column), including split rows and rows spanning chunks."""
    import torch
    rng = np.random.default_rng(50 + k)
    mats = dict(edge_cases(P))
    mats["rmat12"] = P.rmat(12, 16, seed=5)
    for name, A in mats.items():
        if A.nnz == 0:
            continue
        A = _f32_matrix(A)
        S = 16
        X = rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)
        with P.CvrHip(A, steps_per_chunk=S, precision=P.PREC_F32, col_index_bits=bits, x_hot_kb=0) as h:
            assert h.info()["col_index_bits"] == bits
            dX = torch.from_numpy(X).cuda()
            dY = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            h.spmm(k, dX.data_ptr(), dY.data_ptr())
            torch.cuda.synchronize()
            Y = dY.cpu().numpy()
        lay = O.cvr64_build(A, S, 0)
        want = np.stack([O.cvr64_spmv(A, lay, X[:, j].astype(np.float64), parity_tol=0.0) for j in range(k)],
                        1).astype(np.float32)
        assert np.array_equal(Y, want), name


@pytest.mark.parametrize("k", [8, 32])
def test_spmm_with_hot_columns_bit_identical(P, O, k):
    """Handles with hot-x renumbering map hot column numbers back through the
    hot-column list (looked up for a whole group before the X-row loads): the
    single-matrix and batched SpMM give the same Y, bit for bit, as handles
    without hot columns."""
    import torch
    rng = np.random.default_rng(70 + k)
    A = _f32_matrix(P.rmat(14, 16, seed=9))
    X = torch.from_numpy(rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)).cuda()
    out = []
    for hot_kb in (64, 0):
        with P.CvrHip(A, steps_per_chunk=32, precision=P.PREC_F32, x_hot_kb=hot_kb) as h:
            if hot_kb:
                assert len(h.hot_cols()) > 0
            Y = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            h.spmm(k, X.data_ptr(), Y.data_ptr())
            torch.cuda.synchronize()
            b = P.Batch([h], k)
            Yb = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            b.set_operands([X.data_ptr()], [Yb.data_ptr()])
            b.spmm(0)
            torch.cuda.synchronize()
            b.close()
            out.append((Y.cpu().numpy(), Yb.cpu().numpy()))
    (y_hot, yb_hot), (y_plain, yb_plain) = out
    assert np.array_equal(y_hot, y_plain)
    assert np.array_equal(yb_hot, yb_plain)
    assert np.array_equal(y_hot, yb_hot)


def test_seeded_sweep_spmm_and_batch(P, O):
    """16 seeded random matrices (ragged, empty and long rows; banded or
    scattered columns) with random chunk sizes and column widths, one random
    k: each single-matrix SpMM output is the FP32 rounding of the FP64
    CVR-order sum, and one batch of all 16 gives the same Y bit for bit."""
    import torch
    rng = np.random.default_rng(77)
    k = int(rng.integers(1, 33))
    mats, hs, Xs, Ys = [], [], [], []
    for i in range(16):
        M = int(rng.integers(1, 8000))
        N = int(rng.choice([M, int(rng.integers(1, 150000))]))
        lens = np.minimum(rng.choice([0, 1, 5, 20, 300], M, p=[0.1, 0.2, 0.4, 0.29, 0.01]), N)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        banded = i % 2 == 0
        cols = []
        for r, l in enumerate(lens):
            if l == 0:
                continue
            if banded:
                lo = max(0, min(N - l, r * N // M - 100))
                cols.append(lo + np.sort(rng.choice(min(N - lo, max(l, 300)), l, replace=False)))
            else:
                cols.append(np.sort(rng.choice(N, l, replace=False)))
        ci = (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)
        if ci.size == 0:
            continue
        A = _f32_matrix(P.CSR(M, N, rp, ci, rng.uniform(-1, 1, int(rp[-1]))))
        S = int(rng.integers(1, 17)) * 8
        bits = int(rng.choice([0, 32]))
        h = P.CvrHip(A, steps_per_chunk=S, precision=P.PREC_F32, col_index_bits=bits)
        X = rng.uniform(-1, 1, (N, k)).astype(np.float32)
        dX = torch.from_numpy(X).cuda()
        dY = torch.full((M, k), float("nan"), dtype=torch.float32, device="cuda")
        h.spmm(k, dX.data_ptr(), dY.data_ptr())
        torch.cuda.synchronize()
        Y = dY.cpu().numpy()
        lay = O.cvr64_build(A, S, 0)
        want = np.stack([O.cvr64_spmv(A, lay, X[:, j].astype(np.float64), parity_tol=0.0) for j in range(k)],
                        1).astype(np.float32)
        _say(f"spmm sweep {i}: M={M} N={N} nnz={A.nnz} S={S} bits={h.info()['col_index_bits']} k={k}")
        assert np.array_equal(Y, want), i
        mats.append(A)
        hs.append(h)
        Xs.append(dX)
        Ys.append((dY, Y))
    b = P.Batch(hs, k)
    outs = [torch.full_like(dY, float("nan")) for dY, _ in Ys]
    b.set_operands([x.data_ptr() for x in Xs], [o.data_ptr() for o in outs])
    b.spmm(0)
    torch.cuda.synchronize()
    check_queue_counts(b)
    for o, (_, Y) in zip(outs, Ys):
        assert np.array_equal(o.cpu().numpy(), Y)
    b.close()
    for h in hs:
        h.close()


@pytest.mark.parametrize("k", [3, 8])
def test_batch_x_window_bit_identical(P, O, k):
    """The LDS window of X (k <= 8; sized for 1 or 2 workgroups per CU) changes
    where X rows are read, not the sums: Y equals the batch without it bit for
    bit, every group of chunks is handed out once, and graph replays
    reproduce it. The matrices include ones wider than the window, narrower
    than it, non-square, with 16- and 32-bit columns and with matrix ends that
    leave groups of fewer than 4 chunks."""
    import torch
    rng = np.random.default_rng(500 + k)
    mats = []
    for i, (m, n, hb) in enumerate([(3000, 3000, 200), (20000, 20000, 1000), (50000, 50000, 3000), (700, 700, 700),
                                     (12000, 30000, 1500), (30000, 9000, 500), (5000, 5000, 5000)]):
        A = P.synthetic(m, P.RL_POISSON, float(rng.uniform(5, 40)), col_struct=P.COL_MIXED, half_band=hb,
                        band_frac=0.7, seed=700 + i)
        if n != m:  # a non-square matrix of the same row structure
            A = P.CSR(m, n, A.rowOffsets, (A.cols.astype(np.int64) * n // m).astype(np.int32), A.vals)
        mats.append(_f32_matrix(A))
    hs = [P.CvrHip(A, steps_per_chunk=int(rng.choice([8, 24, 64])), precision=P.PREC_F32,
                   col_index_bits=int(rng.choice([0, 32])), x_hot_kb=0) for A in mats]
    Xs = [torch.from_numpy(rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)).cuda() for A in mats]

    def run(win):
        b = P.Batch(hs, k)
        if win:
            b.set_x_window(win)
        Ys = [torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda") for A in mats]
        b.set_operands([x.data_ptr() for x in Xs], [y.data_ptr() for y in Ys])
        s = torch.cuda.Stream()
        b.spmm(s.cuda_stream)
        s.synchronize()
        check_queue_counts(b)
        first = [y.cpu().numpy() for y in Ys]
        b.capture(s.cuda_stream)
        for rep in range(2):
            for y in Ys:
                y.fill_(float("nan"))
            torch.cuda.synchronize()
            b.replay(s.cuda_stream)
            s.synchronize()
            check_queue_counts(b)
            for Y0, y in zip(first, Ys):
                assert np.array_equal(Y0, y.cpu().numpy()), (win, rep)
        inf = b.info()
        b.close()
        return first, inf

    base, inf0 = run(0)
    for A, X, Y in zip(mats, Xs, base):
        yref = _oracle_mm(O, A, X.cpu().numpy())
        assert np.all(np.abs(Y.astype(np.float64) - yref) <= _bound(O, A, X.cpu().numpy(), yref))
    for win in (1, 2):
        got, inf = run(win)
        _say(f"k={k} window {win}: {inf}")
        assert inf["items"] == sum(-(-h.info()["nchunks"] // 4) for h in hs)
        for Y0, Y in zip(base, got):
            assert np.array_equal(Y0, Y), win
    b16 = P.Batch(hs, 16)
    with pytest.raises(P.PavaError):
        b16.set_x_window(1)
    b16.close()
    for h in hs:
        h.close()

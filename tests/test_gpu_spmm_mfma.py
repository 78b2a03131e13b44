"""The MFMA SpMM engine (needs an MI355X): v_mfma_f64_4x4x4f64 over windows of
4 steps of 4 CVR lanes, each nonzero scaling a tile of its X row into its row
segment's accumulator.

References: k independent FP64 oracle SpMVs on the same FP32 inputs, within
one FP32 rounding plus the FP64 summation bound (as for the FMA engine); a
dense FP64 GEMM (numpy) on small dense matrices, where every element of a
4 x 4 tile is a nonzero; and the FMA engine on the same handle. The
instruction adds its four products to the accumulator in K order with one
rounding each (measured: equal to the FMA engine bit for bit on 98M
nonzeros times 56 right-hand sides), and a zero product leaves an
accumulator unchanged, so its Y equals the FMA engine's bit for bit -- which
is also the FP32 rounding of the FP64 CVR-order sum."""
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


def _bound(O, A, X, yref):
    lens = np.diff(A.rowOffsets)[:, None]
    sabs = np.stack([O.row_abs(A, np.abs(X[:, j]).astype(np.float64)) for j in range(X.shape[1])], axis=1)
    return 2.0 ** -24 * np.abs(yref) + 2 * (lens + 64) * 2.0 ** -53 * sabs + 1e-38


def _run_spmm(P, h, X, M, k, engine):
    import torch
    h.set_spmm_engine(engine)
    dX = torch.from_numpy(X).cuda()
    dY = torch.full((M, k), float("nan"), dtype=torch.float32, device="cuda")
    h.spmm(k, dX.data_ptr(), dY.data_ptr())
    torch.cuda.synchronize()
    return dY.cpu().numpy()


def _equal_to_fma(Ym, Yf):
    return np.array_equal(Ym, Yf), int(np.count_nonzero(Ym != Yf))


@pytest.mark.parametrize("bits", [16, 32])
@pytest.mark.parametrize("k", [1, 5, 8, 13, 16, 32])
def test_mfma_spmm_against_oracle_and_fma(P, O, k, bits):
    rng = np.random.default_rng(900 + k)
    mats = dict(edge_cases(P))
    mats["rmat12"] = P.rmat(12, 16, seed=5)
    for name, A in mats.items():
        if A.nnz == 0:
            continue
        A = _f32_matrix(A)
        X = rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)
        with P.CvrHip(A, steps_per_chunk=16, precision=P.PREC_F32, col_index_bits=bits, x_hot_kb=0) as h:
            assert h.info()["col_index_bits"] == bits
            Ym = _run_spmm(P, h, X, A.numRows, k, P.SPMM_MFMA)
            Yf = _run_spmm(P, h, X, A.numRows, k, P.SPMM_FMA)
        assert not np.isnan(Ym).any(), name
        yref = _oracle_mm(O, A, X)
        assert np.all(np.abs(Ym - yref) <= _bound(O, A, X, yref)), name
        ok, ndiff = _equal_to_fma(Ym, Yf)
        _say(f"mfma {name} k={k} bits={bits}: {ndiff} of {Ym.size} outputs differ from the FMA engine")
        assert ok, name


@pytest.mark.parametrize("prec", [0, 2])
def test_mfma_spmm_fp64_and_bf16_handles_with_hot_columns(P, O, prec):
    """FP64 and BF16 storage, with hot-x renumbering (hot columns mapped back
    through the hot-column list when the window's columns are staged) and
    chunks of 64 and 128 steps."""
    rng = np.random.default_rng(910 + prec)
    A = _f32_matrix(P.rmat(14, 16, seed=9))
    if prec == 2:  # values exactly representable in BF16
        v = A.vals.astype(np.float32).view(np.uint32) & np.uint32(0xFFFF0000)
        A.vals = v.view(np.float32).astype(np.float64)
    for k, S in ((8, 64), (32, 128)):
        X = rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)
        with P.CvrHip(A, steps_per_chunk=S, precision=prec, x_hot_kb=64) as h:
            assert len(h.hot_cols()) > 0
            Ym = _run_spmm(P, h, X, A.numRows, k, P.SPMM_MFMA)
            Yf = _run_spmm(P, h, X, A.numRows, k, P.SPMM_FMA)
        yref = _oracle_mm(O, A, X)
        assert np.all(np.abs(Ym - yref) <= _bound(O, A, X, yref))
        ok, ndiff = _equal_to_fma(Ym, Yf)
        _say(f"mfma rmat14 prec={prec} k={k} S={S}: {ndiff} outputs differ from the FMA engine")
        assert ok


@pytest.mark.parametrize("k", [4, 16, 32])
def test_mfma_dense_tiles_against_gemm(P, k):
    """Small dense matrices (every element a nonzero, so every 4 x 4 tile of
    the instruction is full) against a dense FP64 GEMM of the same FP32
    inputs: within one FP32 rounding plus the FP64 summation bound."""
    rng = np.random.default_rng(920 + k)
    for m, n in ((64, 48), (96, 80), (7, 130)):
        D = rng.uniform(-1, 1, (m, n)).astype(np.float32).astype(np.float64)
        A = P.CSR(m, n, np.arange(m + 1, dtype=np.int32) * n, np.tile(np.arange(n, dtype=np.int32), m), D.ravel())
        X = rng.uniform(-1, 1, (n, k)).astype(np.float32)
        with P.CvrHip(A, steps_per_chunk=8, precision=P.PREC_F32, x_hot_kb=0) as h:
            Y = _run_spmm(P, h, X, m, k, P.SPMM_MFMA).astype(np.float64)
        G = D @ X.astype(np.float64)
        bnd = 2.0 ** -24 * np.abs(G) + 2 * (n + 64) * 2.0 ** -53 * (np.abs(D) @ np.abs(X.astype(np.float64)))
        assert np.all(np.abs(Y - G) <= bnd), (m, n)


def check_queue_counts(b):
    counts, nq = b.queue_counts()
    inf = b.info()
    total, grid = inf["items"], inf["grid"]
    assert nq == min(grid, 8)
    for q in range(8):
        want = total * (q + 1) // nq - total * q // nq + grid if q < nq else 0
        assert counts[q] == want, (q, counts, total, grid)


@pytest.mark.parametrize("k", [8, 32])
def test_mfma_batch_persistent_and_graph_replay(P, O, k):
    """The batched persistent launch on the MFMA engine: every chunk handed out
    once, outputs within the oracle bound and equal to the FMA engine's batch,
    graph replays equal to the eager launch."""
    import torch
    rng = np.random.default_rng(930 + k)
    mats = []
    for i in range(12):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e4))))
        A = P.synthetic(n, P.RL_POISSON, float(rng.uniform(5, 60)), col_struct=P.COL_MIXED, half_band=1000,
                        band_frac=0.7, seed=2000 + i)
        mats.append(_f32_matrix(A))
    hs = [P.CvrHip(A, steps_per_chunk=32, precision=P.PREC_F32, col_index_bits=16 if i % 2 else 32, x_hot_kb=0)
          for i, A in enumerate(mats)]
    Xs = [torch.from_numpy(rng.uniform(-1, 1, (A.numCols, k)).astype(np.float32)).cuda() for A in mats]

    def run(engine):
        b = P.Batch(hs, k)
        b.set_engine(engine)
        Ys = [torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda") for A in mats]
        b.set_operands([x.data_ptr() for x in Xs], [y.data_ptr() for y in Ys])
        s = torch.cuda.Stream()
        b.spmm(s.cuda_stream)
        s.synchronize()
        check_queue_counts(b)
        first = [y.cpu().numpy() for y in Ys]
        b.capture(s.cuda_stream)
        for rep in range(3):
            for y in Ys:
                y.fill_(float("nan"))
            torch.cuda.synchronize()
            b.replay(s.cuda_stream)
            s.synchronize()
            check_queue_counts(b)
            for Y0, y in zip(first, Ys):
                assert np.array_equal(Y0, y.cpu().numpy()), (engine, rep)
        _say(f"batch k={k} engine={engine}: {b.info()}")
        b.close()
        return first

    ym = run(P.SPMM_MFMA)
    yf = run(P.SPMM_FMA)
    for A, X, Y, F in zip(mats, Xs, ym, yf):
        yref = _oracle_mm(O, A, X.cpu().numpy())
        assert np.all(np.abs(Y.astype(np.float64) - yref) <= _bound(O, A, X.cpu().numpy(), yref))
        assert _equal_to_fma(Y, F)[0]
    for h in hs:
        h.close()


def test_mfma_engine_arguments(P):
    """Unknown engines are rejected, and the X window (an FMA-engine staging)
    and the MFMA engine exclude each other on a batch."""
    A = _f32_matrix(P.synthetic(3000, P.RL_POISSON, 10.0, col_struct=P.COL_BANDED, half_band=100, seed=3))
    with P.CvrHip(A, steps_per_chunk=32, precision=P.PREC_F32, x_hot_kb=0) as h:
        with pytest.raises(P.PavaError):
            h.set_spmm_engine(2)
        b = P.Batch([h], 8)
        with pytest.raises(P.PavaError):
            b.set_engine(-1)
        b.set_engine(P.SPMM_MFMA)
        with pytest.raises(P.PavaError):
            b.set_x_window(1)
        b.set_engine(P.SPMM_FMA)
        b.set_x_window(1)
        with pytest.raises(P.PavaError):
            b.set_engine(P.SPMM_MFMA)
        b.close()


def test_mfma_noalloc_equals_spmm(P):
    """cvr_hip_spmm_noalloc follows the handle's engine: on the MFMA engine it
    gives cvr_hip_spmm's Y bit for bit (the same kernel; only the cross-chunk
    buffer differs)."""
    import torch
    L = P.lib()
    A = _f32_matrix(P.synthetic(20000, P.RL_POISSON, 20.0, col_struct=P.COL_MIXED, half_band=500, band_frac=0.7,
                                seed=13))
    for k in (8, 32):
        with P.CvrHip(A, precision=P.PREC_F32, steps_per_chunk=16, x_hot_kb=0) as h:
            h.set_spmm_engine(P.SPMM_MFMA)
            X = torch.rand((A.numCols, k), dtype=torch.float32, device="cuda") * 2 - 1
            Y0 = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            Y1 = torch.full_like(Y0, float("nan"))
            assert L.cvr_hip_spmm_noalloc(h.handle, k, X.data_ptr(), Y0.data_ptr(), 0) == 0, L.pava_hip_last_error()
            h.spmm(k, X.data_ptr(), Y1.data_ptr())
            torch.cuda.synchronize()
            assert not torch.isnan(Y0).any()
            assert torch.equal(Y0, Y1), k

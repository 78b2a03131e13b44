"""The batching workload at its stated size (needs an MI355X): 512 matrices
(rows log-uniform in [1e3, 2e5], about 675M nonzeros) in one persistent
launch at k = 8 and k = 32, with the tile chosen by the batched autotuner.

Every matrix's every output column is checked against k oracle SpMVs on the
same FP32 inputs: the GPU accumulates exact FP32 products in FP64 and rounds
once, so each output is within one FP32 rounding (2^-24 relative) plus the
FP64 summation bound 2 (n + 64) 2^-53 sum|a x| of the oracle. The queue
counts show every work item handed out once, and replays of the captured
graph reproduce the eager launch bit for bit."""
import sys

import numpy as np
import pytest

from test_gpu_spmm import check_queue_counts

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def family():
    from pava_amd import selector as SEL
    mats = SEL.batch_family(512)
    _say(f"batch family: {len(mats)} matrices, {sum(A.nnz for A in mats)} nonzeros")
    return mats


def _abs_csr(P, A):
    B = P.CSR(A.numRows, A.numCols, A.rowOffsets, A.cols, np.abs(A.vals))
    return B


@pytest.mark.parametrize("k", [8, 32])
def test_batch_family_full_size(P, O, family, k):
    import torch
    from pava_amd import selector as SEL
    mats = family
    tile, tune = SEL.autotune_batch(mats, k)
    _say(f"k={k}: autotuned tile {tile} from " + ", ".join(f"{t}: {v:.3f} ms" for t, v in tune.items()))
    assert np.isfinite(tune[tile]) and tune[tile] == min(tune.values())
    hs = SEL.build_batch_handles(mats, tile[0], tile[1])
    b = P.Batch(hs, k)
    try:
        b.set_engine(tile[3])
        if tile[2]:
            b.set_x_window(tile[2])
        g = torch.Generator(device="cuda")
        g.manual_seed(k)
        Xs = [torch.rand((A.numCols, k), dtype=torch.float32, device="cuda", generator=g) * 2 - 1 for A in mats]
        Ys = [torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda") for A in mats]
        b.set_operands([x.data_ptr() for x in Xs], [y.data_ptr() for y in Ys])
        _say("batch", b.info())
        s = torch.cuda.Stream()
        b.spmm(s.cuda_stream)
        s.synchronize()
        check_queue_counts(b)
        first = [y.cpu().numpy() for y in Ys]
        # every matrix, every column, within the bound of the oracle SpMVs
        threads = 16
        for i, (A, X, Y) in enumerate(zip(mats, Xs, first)):
            Xh = X.cpu().numpy().astype(np.float64)
            Aabs = _abs_csr(P, A)
            lens = np.diff(A.rowOffsets).astype(np.float64)
            for j in range(k):
                xj = np.ascontiguousarray(Xh[:, j])
                yref = O.spmv_omp(A, xj, threads)
                sabs = O.spmv_omp(Aabs, np.abs(xj), threads)
                bound = 2.0 ** -24 * np.abs(yref) + 2 * (lens + 64) * 2.0 ** -53 * sabs + 1e-38
                yj = Y[:, j].astype(np.float64)
                assert not np.isnan(yj).any(), (i, j)
                assert np.all(np.abs(yj - yref) <= bound), (i, j)
            if i % 64 == 63:
                _say(f"k={k}: {i + 1}/{len(mats)} matrices checked")
        # graph: capture once, replay with eager work in between
        b.capture(s.cuda_stream)
        for rep in range(2):
            for y in Ys:
                y.fill_(float("nan"))
            torch.cuda.synchronize()
            b.replay(s.cuda_stream)
            s.synchronize()
            check_queue_counts(b)
            for Y0, y in zip(first, Ys):
                assert np.array_equal(Y0, y.cpu().numpy()), rep
        _say(f"k={k}: graph replays bit-identical")
    finally:
        b.close()
        for h in hs:
            h.close()

"""cvr_hip_spmm_noalloc (needs an MI355X): the SpMM entry point that allocates
nothing, with its cross-chunk partial sums in the library's per-GPU scratch
array. Called through the C ABI, it must give the same Y as cvr_hip_spmm bit
for bit on two calls in a row (the second on a smaller input), leave the
library's device allocations and the handle's memory unchanged, be capturable
in a graph on a handle that has never run an SpMM, and reject inputs beyond
its scratch size before any device work."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _noalloc(P, h, k, X, Y, stream=0):
    return P.lib().cvr_hip_spmm_noalloc(h.handle, k, X.data_ptr(), Y.data_ptr(), stream)


def test_noalloc_equals_spmm_on_two_calls(P):
    import torch
    L = P.lib()
    cases = [("rmat14_fp64_k16", P.rmat(14, 16, seed=2), P.PREC_F64, 16),
             ("banded_fp32_k8", P.synthetic(20000, P.RL_POISSON, 20.0, col_struct=P.COL_MIXED, half_band=500,
                                            band_frac=0.7, seed=3), P.PREC_F32, 8)]
    for name, A, prec, k in cases:  # the second call is on a smaller input
        _say(f"noalloc {name}: rows={A.numRows} nnz={A.nnz} k={k}")
        with P.CvrHip(A, precision=prec, steps_per_chunk=16) as h:
            X = torch.rand((A.numCols, k), dtype=torch.float32, device="cuda") * 2 - 1
            Y0 = torch.full((A.numRows, k), float("nan"), dtype=torch.float32, device="cuda")
            Y1 = torch.full_like(Y0, float("nan"))
            torch.cuda.synchronize()
            live = L.pava_hip_test_live_device_allocs()
            held = h.info()["device_bytes"]
            assert _noalloc(P, h, k, X, Y0) == 0, L.pava_hip_last_error()
            torch.cuda.synchronize()
            assert L.pava_hip_test_live_device_allocs() == live
            assert h.info()["device_bytes"] == held  # the handle allocated nothing
            h.spmm(k, X.data_ptr(), Y1.data_ptr())
            torch.cuda.synchronize()
            y0, y1 = Y0.cpu().numpy(), Y1.cpu().numpy()
            assert not np.isnan(y0).any()
            assert np.array_equal(y0, y1), name
            # against k oracle-order FP64 SpMVs of the same FP32 inputs
            Xh = X.cpu().numpy().astype(np.float64)
            ref = np.stack([P.reference_spmv(A if prec == P.PREC_F64 else
                                             P.CSR(A.numRows, A.numCols, A.rowOffsets, A.cols,
                                                   A.vals.astype(np.float32).astype(np.float64)), Xh[:, j])
                            for j in range(k)], 1)
            assert np.allclose(y0, ref, rtol=1e-5, atol=1e-5)


def test_noalloc_captured_on_a_fresh_handle(P):
    import torch
    A = P.synthetic(30000, P.RL_LOGNORMAL, 12.0, shape=1.0, col_struct=P.COL_UNIFORM, seed=6)
    k = 8
    with P.CvrHip(A, precision=P.PREC_F32, steps_per_chunk=16) as h:
        X = torch.rand((A.numCols, k), dtype=torch.float32, device="cuda") * 2 - 1
        Y = torch.zeros((A.numRows, k), dtype=torch.float32, device="cuda")
        st = torch.cuda.Stream()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=st):
            assert _noalloc(P, h, k, X, Y, st.cuda_stream) == 0
        want = torch.full_like(Y, float("nan"))
        h.spmm(k, X.data_ptr(), want.data_ptr())
        torch.cuda.synchronize()
        for rep in range(3):
            Y.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(Y, want), rep


def test_noalloc_rejects_inputs_beyond_its_scratch(P):
    import torch
    # 2 * nchunks * k * 8 bytes > 16 MiB: 32,768 chunks of 8 steps at k = 32
    n = 64 * 8 * 33000
    A = P.CSR(1, 1, np.array([0, n], np.int32), np.zeros(n, np.int32), np.ones(n))
    with P.CvrHip(A, precision=P.PREC_F32, steps_per_chunk=8) as h:
        assert h.info()["nchunks"] > 32768
        X = torch.ones((1, 32), dtype=torch.float32, device="cuda")
        Y = torch.zeros((1, 32), dtype=torch.float32, device="cuda")
        assert _noalloc(P, h, 32, X, Y) == -6
        assert "scratch" in P.lib().pava_hip_last_error().decode()
        # within the limit at k = 8
        assert _noalloc(P, h, 8, X[:, :8].contiguous(), Y[:, :8].contiguous()) == 0
        torch.cuda.synchronize()

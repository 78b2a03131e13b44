"""GPU parity at the headline sizes (needs an MI355X).

R-MAT scale 22 (edge factor 16, a=0.57 b=0.19 c=0.19, seed 0: 4,194,304 rows,
about 65M nnz after duplicates are summed) and the 27-point stencil on a 128^3
grid (2,097,152 rows, 55.7M nnz), both with x ~ U[-1,1] from seed 0, through
the library defaults (automatic chunk size and hot-x staging, the LDS table on
the power-law matrix):

* y equals the oracle's layout emulation bit for bit (same summation order),
* every row is within the blocked-sum rounding bound 2 (n+64) u sum|a x| of the
  serial FMA oracle, and rows outside 1e-12 relative of it are cancelling rows
  whose result is within the same bound of the exact (double-double) row sum,
* size-independent properties: y(-x) = -y(x) and y(2x) = 2 y(x) exactly, and
  repeated launches agree,
* the FP32 / BF16 storage variants equal their emulation bit for bit, and FP32
  with FP64 recompute of cancelling rows stays within 1e-5 of the oracle.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def rmat22(P):
    A = P.rmat(22, 16, 0.57, 0.19, 0.19, seed=0)
    assert A.numRows == 1 << 22 and A.nnz > 60_000_000
    return A


@pytest.fixture(scope="module")
def stencil128(P):
    A = P.stencil27(128, 0.01, 0)
    assert A.numRows == 128 ** 3 and A.nnz == 55_742_968
    return A


def _fp64_checks(P, O, A, name):
    x = P.fill_uniform(A.numCols, 0)
    with P.CvrHip(A) as h:
        inf = h.info()
        S = inf["steps_per_chunk"]
        _say(f"{name}: S={S} hot={inf['n_hot_cols']} lds={inf['n_lds_cols']} nnz={A.nnz}")
        y = h.spmv_host(x)
        assert np.array_equal(h.spmv_host(x), y)
        assert np.array_equal(h.spmv_host(-x), -y)
        assert np.array_equal(h.spmv_host(2.0 * x), 2.0 * y)
    emu = O.cvr64_build(A, S, 0)
    assert np.array_equal(y, O.cvr64_spmv(A, emu, x)), name
    yref = O.spmv_fma(A, x)
    lens = np.diff(A.rowOffsets)
    bound = 2 * (lens + 64) * U * O.row_abs(A, x)
    assert np.all(np.abs(y - yref) <= bound), name
    out = np.abs(y - yref) > 1e-12 * np.abs(yref)
    _say(f"{name}: bit-identical rows {np.mean(y == yref):.4f}, rows beyond 1e-12 {int(out.sum())}")
    assert out.mean() <= 1e-5, name
    if out.any():
        ydd = O.spmv_dd(A, x)
        assert np.all(np.abs(y[out] - ydd[out]) <= bound[out]), name
    return inf


def test_rmat22_fp64_defaults(P, O, rmat22):
    inf = _fp64_checks(P, O, rmat22, "rmat22")
    assert inf["n_lds_cols"] > 0  # skewed column use: hot columns staged in LDS


def test_stencil128_fp64_defaults(P, O, stencil128):
    inf = _fp64_checks(P, O, stencil128, "stencil128")
    assert inf["n_hot_cols"] == 0  # even column use: no staging


@pytest.mark.parametrize("prec", [1, 2])
def test_stencil128_low_precision_equals_emulation(P, O, stencil128, prec):
    A = stencil128
    x = P.fill_uniform(A.numCols, 0)
    with P.CvrHip(A, precision=prec) as h:
        S = h.info()["steps_per_chunk"]
        y = h.spmv_host(x)
    emu = O.cvr64_build(A, S, prec)
    ye = O.cvr64_spmv(A, emu, x.astype(np.float32)).astype(np.float64)
    _say(f"stencil128 prec={prec} S={S}")
    assert np.array_equal(y, ye)


def test_stencil128_fp32_recompute_within_1e5(P, O, stencil128):
    import torch
    A = stencil128
    x = P.fill_uniform(A.numCols, 0)
    yref = O.spmv_fma(A, x)
    with P.CvrHip(A, precision=1, recompute=True) as h:
        dx32 = torch.from_numpy(x.astype(np.float32)).cuda()
        dx64 = torch.from_numpy(x).cuda()
        dy = torch.full((A.numRows,), float("nan"), dtype=torch.float32, device="cuda")
        h.spmv_mixed(dx32.data_ptr(), dy.data_ptr(), dx64.data_ptr())
        torch.cuda.synchronize()
        n = h.last_recompute_count()
        y = dy.cpu().numpy().astype(np.float64)
    rel = np.abs(y - yref) / np.maximum(np.abs(yref), 1e-300)
    _say(f"stencil128 f32+recompute: recomputed {n}, max rel {rel.max():.3e}, median {np.median(rel):.3e}")
    assert n > 0
    assert rel.max() <= 1e-5

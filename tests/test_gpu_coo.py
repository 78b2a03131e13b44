"""GPU coo_hip format through the C ABI (needs an MI355X).

coo_hip sums each 64-nonzero step of a tile with a segmented scan (a fixed
tree), the steps of a tile in order, and rows spanning tiles in tile order. So
y is deterministic and every row is within the summation bound
2 k u sum|a x| of the serial FP64 oracle, with k = 8 + ceil(n / 64) + ceil(n / 512)
for a row of n nonzeros (tree depth, steps, tiles, carries); on integer-valued
inputs every partial sum is exact and y equals the oracle exactly."""
import sys

import numpy as np
import pytest

from matrices import edge_cases

pytestmark = pytest.mark.gpu

U = 2.0 ** -53


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _run(P, A, x):
    import torch
    with P.CooHip(A) as h:
        dx = torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()
        dy = torch.full((A.numRows,), float("nan"), dtype=torch.float64, device="cuda")
        h.spmv(dx.data_ptr(), dy.data_ptr(), 0)
        torch.cuda.synchronize()
        y1 = dy.cpu().numpy()
        dy.fill_(float("nan"))
        h.spmv(dx.data_ptr(), dy.data_ptr(), 0)
        torch.cuda.synchronize()
        y2 = dy.cpu().numpy()
        info = h.info()
    assert np.array_equal(y1, y2), "coo_hip is not deterministic"
    return y1, info


def _boundary_matrix(P):
    """Rows ending exactly at step (64) and tile (512) boundaries, rows spanning
    several tiles, empty rows between them."""
    lens = [64, 0, 448, 512, 0, 0, 1, 63, 1, 1023, 2560, 0, 5, 512, 3000, 7, 0]
    rng = np.random.default_rng(4)
    n = 4096
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    return P.CSR(len(lens), n, rp, cols, rng.uniform(-1, 1, rp[-1]))


@pytest.fixture(scope="module")
def mats(P):
    m = dict(edge_cases(P))
    m["rmat13"] = P.rmat(13, 16, seed=2)
    m["stencil"] = P.stencil27(16, 0.01, 1)
    m["boundaries"] = _boundary_matrix(P)
    return m


def test_coo_within_bound_and_deterministic(P, O, mats):
    for name, A in mats.items():
        _say(f"coo {name} nnz={A.nnz}")
        x = P.fill_uniform(A.numCols, 17)
        y, info = _run(P, A, x)
        assert info["ntiles"] == -(-A.nnz // 512)
        yref = O.spmv_fma(A, x)
        n = np.diff(A.rowOffsets)
        k = 8 + -(-n // 64) + -(-n // 512)
        assert np.all(np.abs(y - yref) <= 2 * k * U * O.row_abs(A, x)), name
        assert np.all(y[n == 0] == 0.0), name


def test_coo_exact_on_integer_inputs(P, O, mats):
    for name, A in mats.items():
        B = P.CSR(A.numRows, A.numCols, A.rowOffsets.copy(), A.cols.copy(),
                  (np.arange(A.nnz) % 7 - 3).astype(np.float64))
        x = (np.arange(B.numCols) % 5 - 2).astype(np.float64)
        y, _ = _run(P, B, x)
        assert np.array_equal(y, O.spmv_fma(B, x)), name


def test_conduct_coo_hip_drop_in(P, O, capfd):
    A = P.laplacian2d(100, 90)
    x, _, _ = P.init_vectors(A.numRows, A.numCols)  # x[i] = i % 3
    y = P.conduct_coo_hip(A, x, "McRae lap100x90", 5)
    out = capfd.readouterr().out
    assert " Performance McRae lap100x90 COO_HIP " in out
    assert O.check_results(y, O.spmv_fma(A, x)) == 0
    assert np.array_equal(y, O.spmv_fma(A, x))  # integer-valued products: exact


def test_bench_file_cvr_size_gate(P, tmp_path, capfd):
    """As in pava's CVR block (main.cpp:558), the CVR_HIP block runs only for
    more than 4,096 rows and columns: a 3,000-row matrix gets the skipped-block
    line (-1 -1) and the other formats still run and pass the 1% check."""
    A = P.laplacian2d(60, 50)
    path = tmp_path / "lap3000.mtx"
    P.write_mtx(path, A)
    errors = P.lib().pava_bench_file(str(path).encode(), 2)
    out = capfd.readouterr().out
    assert errors == 0
    line = [l for l in out.splitlines() if " CVR_HIP " in l and "Performance" in l]
    assert len(line) == 1 and line[0].split()[-2:] == ["-1.000000", "-1.000000"], line
    for fmt in ("CSR5_HIP", "CSR_HIP", "COO_HIP"):
        got = [l for l in out.splitlines() if f" {fmt} " in l and "Performance" in l]
        assert len(got) == 1 and not got[0].split()[-1].startswith("-1"), (fmt, got)


def test_bench_file_runs_every_gpu_format(P, tmp_path, capfd):
    """pava_bench_file (benchFile with the GPU registry): every block reports a
    line and passes pava's 1% check."""
    A = P.laplacian2d(100, 90)
    path = tmp_path / "lap.mtx"
    P.write_mtx(path, A)
    errors = P.lib().pava_bench_file(str(path).encode(), 2)
    out = capfd.readouterr().out
    _say(out[-800:])
    assert errors == 0
    for fmt in ("CVR_HIP", "CSR5_HIP", "CSR_HIP", "COO_HIP"):
        assert f" {fmt} " in out, fmt

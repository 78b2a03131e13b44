"""GPU parity of the csr5_hip format through the C ABI (needs an MI355X).

The GPU conversion must reproduce the CPU emulation's CSR5-64 layout array for
array, and the GPU SpMV the emulation's y bit for bit (same summation order);
rows that stay inside one lane equal the serial FP64 oracle bit for bit."""
import sys

import numpy as np
import pytest

from matrices import edge_cases

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mats(P):
    m = dict(edge_cases(P))
    m["rmat13"] = P.rmat(13, 16, seed=2)
    m["stencil"] = P.stencil27(16, 0.01, 1)
    return m


@pytest.mark.parametrize("sigma", [4, 16, 32])
def test_csr5_layout_and_spmv_equal_emulation(P, O, mats, sigma):
    for name, A in mats.items():
        _say(f"csr5 {name} sigma={sigma} nnz={A.nnz}")
        x = P.fill_uniform(A.numCols, 13)
        with P.Csr5Hip(A, sigma=sigma) as h:
            lay = h.export_layout()
            y = h.spmv_host(x)
            y2 = h.spmv_host(x)
        emu = O.csr5_build(A, sigma)
        nt = emu["ntiles"]
        assert lay["ntiles"] == nt, name
        for k, size in [("vals", nt * 64 * sigma), ("cols", nt * 64 * sigma), ("flags", nt * 64),
                        ("yoff", nt * 64), ("tile_base", nt), ("fix", 2 * nt), ("cont", nt),
                        ("nerows", emu["n_nonempty"])]:
            assert np.array_equal(lay[k][:size], emu[k][:size]), (name, k)
        assert np.array_equal(y, O.csr5_spmv(A, emu, x)), name
        assert np.array_equal(y, y2), name
        rp = A.rowOffsets.astype(np.int64)
        lens = np.diff(rp)
        inside = (lens > 0) & (rp[:-1] // sigma == (rp[1:] - 1) // sigma)
        yref = O.spmv_fma(A, x)
        assert np.array_equal(y[inside], yref[inside]), name


def test_csr5_device_api_writes_every_row(P, O):
    import torch
    A = P.rmat(14, 16, seed=4)  # many empty rows
    x = P.fill_uniform(A.numCols, 1)
    dx = torch.from_numpy(x).cuda()
    dy = torch.full((A.numRows,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    with P.Csr5Hip(A) as h:
        torch.cuda.synchronize()
        h.spmv(dx.data_ptr(), dy.data_ptr(), s.cuda_stream)
        s.synchronize()
        y = dy.cpu().numpy()
        assert h.info()["sigma"] == 16
    assert not np.isnan(y).any()
    assert np.array_equal(y, O.csr5_spmv(A, O.csr5_build(A, 16), x))


def test_csr5_empty_and_invalid(P):
    A = P.CSR(7, 5, np.zeros(8, np.int32), np.zeros(0, np.int32), np.zeros(0))
    with P.Csr5Hip(A) as h:
        assert np.array_equal(h.spmv_host(np.ones(5)), np.zeros(7))
    B = P.laplacian2d(8, 8)
    for sg in (3, 64):
        with pytest.raises(P.PavaError):
            P.Csr5Hip(B, sigma=sg)


def test_conduct_csr5_hip_drop_in(P, O, capfd):
    A = P.laplacian2d(90, 70)
    x, _, _ = P.init_vectors(A.numRows, A.numCols)  # x[i] = i % 3
    y = P.conduct_csr5_hip(A, x, "McRae lap90x70", 5)
    out = capfd.readouterr().out
    assert " Performance McRae lap90x70 CSR5_HIP " in out
    assert np.array_equal(y, O.spmv_fma(A, x))  # integer-valued products: exact


def test_seeded_sweep_csr5_coo_csr(P, O):
    """24 seeded random shapes: CSR5-HIP (random sigma) equals its emulation bit
    for bit; COO-HIP and CSR-vector (random lanes per row) stay within the
    rounding bound 2 (n + 64) u sum |a x| of the serial oracle."""
    import torch
    from matrices import random_csr
    rng = np.random.default_rng(91)
    for case in range(24):
        A = random_csr(P, rng)
        x = P.fill_uniform(A.numCols, case)
        sigma = int(rng.choice([4, 8, 16, 32]))
        L = int(rng.choice([1, 2, 4, 8, 16, 32, 64]))
        _say(f"sweep {case}: M={A.numRows} N={A.numCols} nnz={A.nnz} sigma={sigma} L={L}")
        with P.Csr5Hip(A, sigma=sigma) as h:
            y5 = h.spmv_host(x)
        assert np.array_equal(y5, O.csr5_spmv(A, O.csr5_build(A, sigma), x)), case
        yref = O.spmv_fma(A, x)
        bound = 2 * (np.diff(A.rowOffsets) + 64) * 2.0 ** -53 * O.row_abs(A, x)
        dx = torch.from_numpy(x).cuda()
        for h in (P.CooHip(A), P.CsrHip(A, lanes_per_row=L)):
            dy = torch.full((A.numRows,), float("nan"), dtype=torch.float64, device="cuda")
            with h:
                torch.cuda.synchronize()
                h.spmv(dx.data_ptr(), dy.data_ptr(), 0)
                torch.cuda.synchronize()
            y = dy.cpu().numpy()
            assert not np.isnan(y).any(), (case, type(h).__name__)
            assert np.all(np.abs(y - yref) <= bound), (case, type(h).__name__)

"""GPU parity of the cvr_hip path through the C ABI (needs an MI355X).

The GPU conversion must reproduce the CPU emulation's layout array for array,
and the GPU SpMV must reproduce the emulation's y bit for bit (same summation
order); against the serial FP64 oracle, rows the layout does not split are
bit-identical and split rows stay within the rounding bound of their blocked
sum (|dy| <= 2 (n+64) u sum|a x|), with the per-row relative error reported.
"""
import sys

import numpy as np
import pytest

from cvr_check import reconstruct
from matrices import edge_cases, split_rows

pytestmark = pytest.mark.gpu

SPLIT_CAP = 72


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def mats(P):
    return edge_cases(P)


def _compare_layout(A, gpu, emu, S):
    inf = gpu["info"]
    nc = inf["nchunks"]
    assert nc == emu["nchunks"] and inf["layout_elems"] == emu["layout_elems"]
    assert np.array_equal(gpu["vals"], emu["vals"])
    assert np.array_equal(gpu["cols"], emu["cols"])
    assert np.array_equal(gpu["chunk_rows"], emu["chunk_rows"])
    if nc == 0:
        return
    assert np.array_equal(gpu["seg_count"][:nc], emu["seg_count"][:nc])
    assert np.array_equal(gpu["split_n"][:nc], emu["split_n"][:nc])
    assert np.array_equal(gpu["fix"][:2 * nc], emu["fix"][:2 * nc])
    for c in range(nc):
        sb = int(emu["chunk_rows"][c]) + 66 * c
        n = int(emu["seg_count"][c])
        assert np.array_equal(gpu["seg"][sb:sb + n], emu["seg"][sb:sb + n]), c
        k = 2 * (int(emu["split_n"][c]) + 1)
        t0 = c * 2 * SPLIT_CAP
        assert np.array_equal(gpu["split_tab"][t0:t0 + k], emu["split_tab"][t0:t0 + k]), c
    assert inf["max_slots"] == int(emu["info"][0])
    assert inf["n_fixup_rows"] == int(emu["info"][1])


@pytest.mark.parametrize("S", [8, 64, 256])
def test_gpu_layout_equals_emulation(P, O, mats, S):
    for name, A in mats.items():
        _say(f"layout {name} S={S} m={A.numRows} nnz={A.nnz}")
        with P.CvrHip(A, steps_per_chunk=S) as h:
            gpu = h.export_layout()
        emu = O.cvr64_build(A, S, 0)
        _compare_layout(A, gpu, emu, S)
        assert not reconstruct(A, gpu, S)


@pytest.mark.parametrize("S", [8, 64, 256])
def test_gpu_spmv_bit_exact_vs_emulation_and_oracle(P, O, mats, S):
    for name, A in mats.items():
        _say(f"spmv {name} S={S}")
        x = P.fill_uniform(A.numCols, 7)
        with P.CvrHip(A, steps_per_chunk=S) as h:
            y = h.spmv_host(x)
            y2 = h.spmv_host(x)
        emu = O.cvr64_build(A, S, 0)
        ye = O.cvr64_spmv(A, emu, x)
        assert np.array_equal(y, ye), name
        assert np.array_equal(y, y2), name  # deterministic
        yref = O.spmv_fma(A, x)
        split = split_rows(A, emu, S)
        mask = np.ones(A.numRows, bool)
        mask[list(split)] = False
        assert np.array_equal(y[mask], yref[mask]), name
        lens = np.diff(A.rowOffsets)
        sabs = O.row_abs(A, x)
        assert np.all(np.abs(y - yref) <= 2 * (lens + 64) * 2.0 ** -53 * sabs), name


def test_laplacian_256_config(P, O):
    """2D 5-point Laplacian 256x256 (65,536 rows, 326,656 nnz), x ~ U[-1,1] seed 0."""
    A = P.laplacian2d(256, 256)
    assert A.nnz == 326656
    x = P.fill_uniform(A.numCols, 0)
    yref = O.spmv_fma(A, x)
    with P.CvrHip(A) as h:
        y = h.spmv_host(x)
    st = O.error_stats(y, yref, O.row_abs(A, x))
    _say("laplacian256", st)
    assert st["max_rel"] <= 1e-12
    assert O.check_results(y, yref) == 0


def test_device_pointer_api_and_stream(P, O):
    import torch
    A = P.rmat(14, 16, seed=1)
    x = P.fill_uniform(A.numCols, 3)
    yref = O.spmv_fma(A, x)
    dx = torch.from_numpy(x).cuda()
    dy = torch.full((A.numRows,), float("nan"), dtype=torch.float64, device="cuda")
    s = torch.cuda.Stream()
    with P.CvrHip(A) as h:
        torch.cuda.synchronize()
        h.spmv(dx.data_ptr(), dy.data_ptr(), s.cuda_stream)
        s.synchronize()
        y = dy.cpu().numpy()
        tot, main = h.time_spmv(dx.data_ptr(), dy.data_ptr(), 2, 5)
        assert tot > 0 and main > 0
    assert not np.isnan(y).any()  # every row written, empty rows included
    sabs = O.row_abs(A, x)
    lens = np.diff(A.rowOffsets)
    assert np.all(np.abs(y - yref) <= 2 * (lens + 64) * 2.0 ** -53 * sabs)


def test_empty_matrix_writes_zeros(P):
    A = P.CSR(10, 10, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
    with P.CvrHip(A) as h:
        assert np.array_equal(h.spmv_host(np.ones(10)), np.zeros(10))


def test_conduct_cvr_hip_drop_in(P, O, capfd):
    A = P.laplacian2d(100, 90)
    x, _, _ = P.init_vectors(A.numRows, A.numCols)  # x[i] = i % 3
    y = P.conduct_cvr_hip(A, x, "McRae lap100x90", 5)
    out = capfd.readouterr().out
    assert " Performance McRae lap100x90 CVR_HIP " in out
    assert O.check_results(y, O.spmv_fma(A, x)) == 0
    assert np.array_equal(y, O.spmv_fma(A, x))  # integer-valued products: exact


@pytest.mark.parametrize("prec", [1, 2])
def test_low_precision_bit_exact_vs_emulation(P, O, mats, prec):
    for name, A in mats.items():
        _say(f"prec {prec} {name}")
        x = P.fill_uniform(A.numCols, 5)
        with P.CvrHip(A, steps_per_chunk=64, precision=prec) as h:
            gpu = h.export_layout()
            y = h.spmv_host(x)  # x rounded to FP32 inside, y widened from FP32
        emu = O.cvr64_build(A, 64, prec)
        assert np.array_equal(gpu["vals"], emu["vals"]), name
        ye = O.cvr64_spmv(A, emu, x.astype(np.float32)).astype(np.float64)
        assert np.array_equal(y, ye), name


def test_invalid_options_rejected(P):
    A = P.laplacian2d(10, 10)
    for kw in [dict(steps_per_chunk=7), dict(steps_per_chunk=8192), dict(precision=5),
               dict(col_index_bits=16)]:
        with pytest.raises(P.PavaError):
            P.CvrHip(A, **kw)

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
        with P.CvrHip(A, steps_per_chunk=S, x_hot_kb=0, col_index_bits=32) as h:
            gpu = h.export_layout()
        emu = O.cvr64_build(A, S, 0)
        _compare_layout(A, gpu, emu, S)
        assert not reconstruct(A, gpu, S)


@pytest.mark.parametrize("hot_kb,lds_kb", [(1, 0), (64, 1), (2048, -1), (2048, 0)])
def test_hot_x_staging_decodes_and_is_bit_identical(P, O, mats, hot_kb, lds_kb):
    for name, A in mats.items():
        _say(f"hot-x {name} hot_kb={hot_kb} lds_kb={lds_kb}")
        S = 64
        x = P.fill_uniform(A.numCols, 11)
        with P.CvrHip(A, steps_per_chunk=S, x_hot_kb=hot_kb, x_lds_kb=lds_kb) as h:
            lay = h.export_layout()
            hot = h.hot_cols()
            y = h.spmv_host(x)
            n_lds = h.info()["n_lds_cols"]
        if lds_kb == 0 or len(hot) == 0:
            assert n_lds == 0
        elif lds_kb > 0:
            assert n_lds == min(len(hot), lds_kb * 1024 // 8)
        emu = O.cvr64_build(A, S, 0)
        cols = lay["cols"].astype(np.int64)
        is_hot = (cols & 0x40000000) != 0
        idx = cols & 0x3FFFFFFF
        dec = np.where(is_hot, hot[np.minimum(idx, max(len(hot) - 1, 0))] if len(hot) else idx, idx)
        assert np.array_equal(dec, emu["cols"] & 0x7FFFFFFF), name
        assert np.array_equal(cols < 0, emu["cols"] < 0), name
        assert len(hot) <= hot_kb * 1024 // 8
        if len(hot):
            counts = np.bincount(A.cols, minlength=A.numCols)
            # hot numbers follow use count (descending), ties by column
            key = np.stack([-counts[hot], hot])
            assert np.all((np.diff(key[0]) > 0) | ((np.diff(key[0]) == 0) & (np.diff(key[1]) > 0)))
            assert counts[hot].min() >= 2
            # the hot set is the top of the ranking: nothing left out beats its last entry
            rest = np.setdiff1d(np.nonzero(counts >= 2)[0], hot)
            if len(rest):
                last = (counts[hot[-1]], hot[-1])
                assert not np.any((counts[rest] > last[0]) | ((counts[rest] == last[0]) & (rest < last[1])))
        assert np.array_equal(y, O.cvr64_spmv(A, emu, x)), name


def _check_hot_ranking(A, hot, budget):
    counts = np.bincount(A.cols, minlength=A.numCols)
    assert len(hot) <= budget
    key = np.stack([-counts[hot], hot])
    assert np.all((np.diff(key[0]) > 0) | ((np.diff(key[0]) == 0) & (np.diff(key[1]) > 0)))
    assert counts[hot].min() >= 2
    rest = np.setdiff1d(np.nonzero(counts >= 2)[0], hot)
    if len(rest):
        last = (counts[hot[-1]], hot[-1])
        assert not np.any((counts[rest] > last[0]) | ((counts[rest] == last[0]) & (rest < last[1])))
    if len(hot) < budget:  # a short selection takes every column used twice or more
        assert len(rest) == 0


@pytest.mark.parametrize("hot_kb", [64, 2048])
def test_hot_x_selection_on_device(P, O, hot_kb):
    """Matrices with >= 65,536 columns select the hot columns from a device
    histogram of the counts: same ranking contract, including columns used
    65,535 times or more (ordered by their own counts) and a budget that cuts
    through a count value."""
    rng = np.random.default_rng(21)
    M, N = 72000, 100000
    # columns 0, 1, 2 in most rows; then k = 0..len-1 distinct columns
    # 3 + (start + 4732 k) mod (N - 3) with a random start per row
    lens = rng.integers(3, 12, M)
    start = rng.integers(0, N - 3, M)
    r = np.arange(M)
    fixed = np.stack([np.zeros(M), np.where(r < 69000, 1, -1), np.where(r % 12 < 11, 2, -1)], 1)
    k = np.arange(12)
    rest = 3 + (start[:, None] + 4732 * k[None, :]) % (N - 3)
    rest = np.where(k[None, :] < lens[:, None], rest, -1)
    allc = np.sort(np.concatenate([fixed, rest], 1).astype(np.int64), 1)
    keep = allc >= 0
    rp = np.concatenate([[0], np.cumsum(keep.sum(1))]).astype(np.int32)
    cols = allc[keep].astype(np.int32)
    A = P.CSR(M, N, rp, cols, rng.uniform(-1, 1, len(cols)))
    big = np.bincount(cols, minlength=N)[:3]
    assert np.all(big >= 65535)
    for name, B in [("synthetic", A), ("rmat17", P.rmat(17, 16, seed=8))]:
        x = P.fill_uniform(B.numCols, 4)
        with P.CvrHip(B, steps_per_chunk=64, x_hot_kb=hot_kb) as h:
            hot = h.hot_cols()
            y = h.spmv_host(x)
        _say(f"device hot selection {name} hot_kb={hot_kb}: {len(hot)} hot columns")
        assert len(hot) > 0
        _check_hot_ranking(B, hot, hot_kb * 1024 // 8)
        if name == "synthetic":
            assert list(hot[:3]) == [0, 1, 2]
        assert np.array_equal(y, O.cvr64_spmv(B, O.cvr64_build(B, 64, 0), x)), name


def test_automatic_hot_x_only_on_skewed_column_use(P, O):
    """x_hot_kb = -1 stages the hot columns of a power-law matrix and nothing
    for meshes whose columns are used evenly; y is the emulation's either way."""
    cases = {"rmat": (P.rmat(14, 16, seed=6), True), "stencil": (P.stencil27(24, 0.01, 2), False),
             "laplacian": (P.laplacian2d(200, 150), False)}
    for name, (A, skewed) in cases.items():
        x = P.fill_uniform(A.numCols, 3)
        with P.CvrHip(A, steps_per_chunk=64, x_hot_kb=-1, col_index_bits=32) as h:
            n_hot = h.info()["n_hot_cols"]
            y = h.spmv_host(x)
        counts = np.bincount(A.cols, minlength=A.numCols)
        _say(f"auto hot-x {name}: hot columns {n_hot}, max/mean use {counts.max() / counts[counts > 0].mean():.1f}")
        assert (n_hot > 0) == skewed, name
        if skewed:
            assert n_hot <= 2048 * 1024 // 8
        assert np.array_equal(y, O.cvr64_spmv(A, O.cvr64_build(A, 64, 0), x)), name


@pytest.mark.parametrize("prec,recompute", [(0, False), (1, False), (1, True), (2, False)])
def test_lds_hot_table_bit_identical(P, O, prec, recompute):
    """The LDS hot-x variant (one workgroup per CU walking a chunk range) gives
    the same y, bit for bit, as the plain kernel on a power-law matrix."""
    A = P.rmat(16, 16, seed=3)
    x = P.fill_uniform(A.numCols, 9)
    _say(f"lds-hot prec={prec} recompute={recompute} nnz={A.nnz}")
    kw = dict(steps_per_chunk=64, precision=prec, recompute=recompute, keep_fp64=recompute, col_index_bits=32)
    with P.CvrHip(A, x_hot_kb=256, **kw) as h, P.CvrHip(A, x_hot_kb=0, **kw) as h0:
        assert h.info()["n_lds_cols"] >= 1024
        y, y0 = h.spmv_host(x), h0.spmv_host(x)
        if recompute:
            assert h.last_recompute_count() == h0.last_recompute_count()
        # repeated launches give the same y
        for _ in range(3):
            assert np.array_equal(h.spmv_host(x), y)
    assert np.array_equal(y, y0)
    if prec == 0:
        assert np.array_equal(y, O.cvr64_spmv(A, O.cvr64_build(A, 64, 0), x))


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


def test_automatic_chunk_size(P, O):
    """steps_per_chunk = 0: about four chunks per resident wave slot (16 per CU),
    a multiple of 8 in [8, 256]; two per slot with FP64 parity on a 16-bit
    layout. The layout is the emulation's at that size."""
    import torch
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    for A in (P.stencil27(64, 0.01, 0), P.laplacian2d(40, 40)):
        for tol in (0.0, None):
            with P.CvrHip(A, parity_tol=tol) as h:
                inf = h.info()
                S = inf["steps_per_chunk"]
                x = P.fill_uniform(A.numCols, 1)
                y = h.spmv_host(x)
            rounds = 2 if inf["parity_tol"] > 0 and inf["col_index_bits"] == 16 else 4
            steps = -(-A.nnz // (64 * rounds * 16 * ncu))  # ceil: chunks per wave slot
            want = min(max(-(-steps // 8) * 8, 8), 256)
            assert S == want, (S, want, tol)
            lay = O.cvr64_build(A, S, 0)
            assert np.array_equal(y, O.cvr64_spmv(A, lay, x, parity_tol=inf["parity_tol"]))


def test_invalid_options_rejected(P):
    A = P.laplacian2d(10, 10)
    for kw in [dict(steps_per_chunk=7), dict(steps_per_chunk=8192), dict(precision=5),
               dict(col_index_bits=12)]:
        with pytest.raises(P.PavaError):
            P.CvrHip(A, **kw)


def _mixed_run(P, A, x, prec, recompute, kappa=64.0, col_bits=32, S=64):
    import torch
    with P.CvrHip(A, steps_per_chunk=S, precision=prec, recompute=recompute, kappa_threshold=kappa,
                  col_index_bits=col_bits) as h:
        dx32 = torch.from_numpy(x.astype(np.float32)).cuda()
        dx64 = torch.from_numpy(x).cuda()
        dy = torch.full((A.numRows,), float("nan"), dtype=torch.float32, device="cuda")
        h.spmv_mixed(dx32.data_ptr(), dy.data_ptr(), dx64.data_ptr() if recompute else 0)
        torch.cuda.synchronize()
        y = dy.cpu().numpy().astype(np.float64)
        n = h.last_recompute_count()
        # second call: flags were cleared, same result
        h.spmv_mixed(dx32.data_ptr(), dy.data_ptr(), dx64.data_ptr() if recompute else 0)
        torch.cuda.synchronize()
        assert np.array_equal(y, dy.cpu().numpy().astype(np.float64))
    return y, n


def test_fp32_recompute_meets_1e5_on_cancelling_rows(P, O):
    """27-point stencil with x = 1: interior rows cancel (26 - sum of 26 values ~ -1)."""
    A = P.stencil27(20, 0.01, 0)
    x = np.ones(A.numCols)
    yref = O.spmv_fma(A, x)
    sabs = O.row_abs(A, x)
    kappa = sabs / np.maximum(np.abs(yref), 1e-300)
    y_plain, n0 = _mixed_run(P, A, x, 1, False)
    y_rc, n1 = _mixed_run(P, A, x, 1, True)
    assert n0 == 0 and n1 > 0
    rel_plain = np.abs(y_plain - yref) / np.abs(yref)
    rel_rc = np.abs(y_rc - yref) / np.abs(yref)
    _say("kappa max", kappa.max(), "rows > 64:", int((kappa > 64).sum()), "recomputed", n1,
         "max rel plain", rel_plain.max(), "with recompute", rel_rc.max())
    assert rel_rc.max() <= 1e-5
    # rows certainly flagged (rows spanning chunks included) are recomputed in
    # the serial FMA order: the FP32 rounding of the oracle
    sure = np.nonzero(kappa > 1024)[0]
    assert sure.size > 0
    assert np.array_equal(y_rc[sure], yref[sure].astype(np.float32).astype(np.float64))
    # rows with kappa well below the threshold are left to the FP32 path and its bound
    low = kappa < 32
    assert np.all(np.abs(y_rc[low] - yref[low]) <= 2.0 ** -22 * sabs[low] + 2.0 ** -23 * np.abs(yref[low]))


@pytest.mark.parametrize("S", [8, 64])
def test_fp32_recompute_every_row_flagged(P, O, S):
    """Every row cancels (values +-1 alternating, x = 1 + tiny): all rows are
    flagged, including rows spanning chunks and rows recomputed in several
    pieces. Rows of at most 128 elements are the serial FMA sum rounded to FP32
    exactly; longer rows (pieces of 512 summed by whole waves, then added in
    order) are within FP64 rounding of it."""
    rng = np.random.default_rng(11)
    lens = np.concatenate([rng.integers(2, 40, 3000), rng.integers(100, 160, 300), [700, 2000, 5000]])
    rng.shuffle(lens)
    M = lens.size
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    n = 8192
    cols = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    vals = np.concatenate([np.where(np.arange(l) % 2 == 0, 1.0, -1.0) * (1 + 0.01 * rng.random(l)) for l in lens])
    A = P.CSR(M, n, rp, cols, vals)
    x = 1.0 + 1e-3 * rng.random(n)
    yref = O.spmv_fma(A, x)
    kappa = O.row_abs(A, x) / np.maximum(np.abs(yref), 1e-300)
    y, cnt = _mixed_run(P, A, x, 1, True, kappa=1.0, S=S)
    _say(f"all-flagged S={S}: rows {M}, recomputed {cnt}, kappa min {kappa.min():.1f}")
    assert cnt == M
    short = lens <= 128
    assert np.array_equal(y[short], yref[short].astype(np.float32).astype(np.float64))
    rel = np.abs(y[~short] - yref[~short]) / np.abs(yref[~short])
    assert rel.max() <= 2.0 ** -24 + 1e-12 * kappa[~short].max()


def test_fp32_recompute_random_x_and_rmat(P, O):
    for name, A in [("stencil", P.stencil27(24, 0.01, 3)), ("rmat", P.rmat(15, 16, seed=2))]:
        x = P.fill_uniform(A.numCols, 0)
        yref = O.spmv_fma(A, x)
        y, n = _mixed_run(P, A, x, 1, True, S=256)
        rel = np.abs(y - yref) / np.maximum(np.abs(yref), 1e-300)
        _say(name, "recomputed", n, "max rel", rel.max(), "median", np.median(rel))
        assert rel.max() <= 1e-5


def test_col16_layout_and_spmv(P, O):
    A = P.stencil27(24, 0.01, 1)
    x = P.fill_uniform(A.numCols, 0)
    y16, _ = _mixed_run(P, A, x, 1, False, col_bits=16, S=64)
    y32, _ = _mixed_run(P, A, x, 1, False, col_bits=32, S=64)
    assert np.array_equal(y16, y32)  # same layout order, only the column encoding differs
    with P.CvrHip(A, steps_per_chunk=64, precision=1, col_index_bits=16) as h:
        lay = h.export_layout()
    emu = O.cvr64_build(A, 64, 1)
    nc = lay["info"]["nchunks"]
    # rebuild 32-bit flagged columns from base + offset + end mask and compare with the emulation
    base = np.repeat(lay["col_base"][:nc], 64 * 64)[: lay["cols"].size]
    full = base + lay["cols"].astype(np.int64)
    ends = np.unpackbits(lay["end_mask"].view(np.uint8), bitorder="little").astype(bool)[: full.size]
    # padding elements (value 0, sums discarded) point at the chunk's base column here
    real = ~((emu["vals"] == 0) & ((emu["cols"] & 0x7FFFFFFF) == 0))
    assert np.array_equal(full[real], (emu["cols"] & 0x7FFFFFFF)[real])
    assert np.array_equal(ends, emu["cols"] < 0)


def test_fp64_col16_equals_32bit_layout(P, O, mats):
    """FP64 storage with 16-bit chunk-relative columns: same layout order, so
    the same y bit for bit as the 32-bit layout; a chunk spanning 65,536
    columns or more is refused with PAVA_HIP_ERANGE."""
    cases = dict(mats)
    cases["stencil"] = P.stencil27(24, 0.01, 3)
    cases["laplacian"] = P.laplacian2d(300, 200)
    done = 0
    for name, A in cases.items():
        x = P.fill_uniform(A.numCols, 6)
        for S in (8, 64):
            with P.CvrHip(A, steps_per_chunk=S, x_hot_kb=0, col_index_bits=32) as h32:
                y32 = h32.spmv_host(x)
            try:
                h16 = P.CvrHip(A, steps_per_chunk=S, col_index_bits=16)
            except P.PavaError as e:
                assert e.code == -6, (name, S)
                continue
            with h16:
                assert h16.info()["col_index_bits"] == 16
                y16 = h16.spmv_host(x)
            assert np.array_equal(y16, y32), (name, S)
            done += 1
    wide = P.CSR(2, 200000, np.array([0, 2, 3]), np.array([0, 150000, 7]), np.ones(3))
    with pytest.raises(P.PavaError) as e:
        P.CvrHip(wide, steps_per_chunk=8, col_index_bits=16)
    assert e.value.code == -6
    assert done >= 4


def test_automatic_column_width(P, O):
    """col_index_bits = 0 (the default) takes 16-bit columns when every chunk's
    nonzeros span fewer than 65,536 columns (always when numCols <= 65,536),
    32-bit otherwise or when hot-x staging is requested explicitly; y is the
    32-bit layout's either way."""
    rng = np.random.default_rng(12)
    n = 300000
    rows = np.arange(n)
    def csr(other):  # row r: columns r and other[r]
        c = np.sort(np.stack([rows, other], 1), 1)
        two = c[:, 1] != c[:, 0]
        cols = np.stack([c[:, 0], np.where(two, c[:, 1], -1)], 1).reshape(-1)
        rp = np.concatenate([[0], np.cumsum(1 + two)]).astype(np.int32)
        return P.CSR(n, n, rp, cols[cols >= 0].astype(np.int32), rng.uniform(-1, 1, int(rp[-1])))
    band = csr(np.minimum(rows + 1000, n - 1))     # a chunk spans its rows plus 1,000 columns
    far = csr((rows * 7919 + 12345) % n)           # a chunk spans the whole matrix
    cases = [("banded", band, 16, {}), ("scattered", far, 32, {}),
             ("small rmat", P.rmat(14, 16, seed=2), 16, {}), ("explicit hot-x", P.rmat(14, 16, seed=2), 32,
                                                              dict(x_hot_kb=64))]
    for name, A, bits, kw in cases:
        x = P.fill_uniform(A.numCols, 1)
        with P.CvrHip(A, steps_per_chunk=64, **kw) as h:
            got = h.info()["col_index_bits"]
            y = h.spmv_host(x)
        with P.CvrHip(A, steps_per_chunk=64, x_hot_kb=0, col_index_bits=32) as h:
            y32 = h.spmv_host(x)
        _say(f"auto column width {name}: {got} bits")
        assert got == bits, name
        assert np.array_equal(y, y32), name


@pytest.mark.parametrize("span,bits", [(65535, 16), (65536, 32)])
def test_column_width_boundary(P, span, bits):
    """A chunk whose columns span exactly 65,535 (largest 16-bit offset)
    takes 16-bit columns automatically; one more column forces 32 bits, and an
    explicit 16-bit request is then refused."""
    n = 200000
    rows = 600
    lens = np.full(rows, 2)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = np.stack([np.arange(rows) + 10, np.arange(rows) + 20], 1).reshape(-1)
    cols[1] = 10 + span  # row 0: columns 10 and 10 + span, in the first chunk
    A = P.CSR(rows, n, rp, cols.astype(np.int32), np.random.default_rng(1).uniform(-1, 1, rp[-1]))
    x = P.fill_uniform(n, 2)
    with P.CvrHip(A, steps_per_chunk=8) as h:
        assert h.info()["col_index_bits"] == bits
        y = h.spmv_host(x)
    with P.CvrHip(A, steps_per_chunk=8, col_index_bits=32, x_hot_kb=0) as h:
        assert np.array_equal(y, h.spmv_host(x))
    if bits == 32:
        with pytest.raises(P.PavaError) as e:
            P.CvrHip(A, steps_per_chunk=8, col_index_bits=16)
        assert e.value.code == -6


def test_spmv_captured_in_a_graph(P):
    """cvr_hip_spmv and cvr_hip_spmv_mixed enqueue kernels only (no memset, copy,
    allocation or synchronisation), so a caller can capture them in a graph.
    Replays with new x values in the same buffers, with eager work and
    synchronisations between them, equal eager calls bit for bit. Covers hot-x
    staging with the LDS table, 16-bit columns, FP32 with the recompute pass
    (long rows included) and SpMM of a matrix without nonzeros."""
    import torch
    rng = np.random.default_rng(5)
    lens = np.concatenate([rng.integers(1, 30, 4000), [300, 900]])
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    cols = np.concatenate([np.sort(rng.choice(6000, l, replace=False)) for l in lens]).astype(np.int32)
    cancel = P.CSR(lens.size, 6000, rp, cols, rng.choice([-1.0, 1.0], rp[-1]) * (1 + 0.01 * rng.random(rp[-1])))
    cases = [("rmat hot-x + LDS", P.rmat(16, 16, seed=3), dict(x_hot_kb=256, col_index_bits=32)),
             ("stencil 16-bit", P.stencil27(32, 0.01, 1), dict(col_index_bits=16)),
             ("f32 recompute", cancel, dict(precision=1, recompute=True, kappa_threshold=4.0))]
    for name, A, kw in cases:
        prec = kw.get("precision", 0)
        with P.CvrHip(A, steps_per_chunk=64, **kw) as h:
            if name.startswith("rmat"):
                assert h.info()["n_lds_cols"] > 0
            dx64 = torch.empty(A.numCols, dtype=torch.float64, device="cuda")
            dx32 = torch.empty(A.numCols, dtype=torch.float32, device="cuda")
            dy = torch.empty(A.numRows, dtype=torch.float64 if prec == 0 else torch.float32, device="cuda")

            def call(stream):
                if prec == 0:
                    h.spmv(dx64.data_ptr(), dy.data_ptr(), stream)
                else:
                    h.spmv_mixed(dx32.data_ptr(), dy.data_ptr(), dx64.data_ptr(), stream)

            def set_x(seed):
                dx64.copy_(torch.from_numpy(P.fill_uniform(A.numCols, seed)))
                dx32.copy_(dx64)
                torch.cuda.synchronize()

            set_x(0)
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                call(s.cuda_stream)
            s.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                call(torch.cuda.current_stream().cuda_stream)
            torch.cuda.synchronize()
            for rep in range(4):
                set_x(10 + rep)
                call(0)
                torch.cuda.synchronize()
                want = dy.clone()
                if prec:
                    assert h.last_recompute_count() > 0
                dy.fill_(float("nan"))
                torch.cuda.synchronize()
                g.replay()
                torch.cuda.synchronize()
                _say(f"graph {name}: replay {rep}")
                assert torch.equal(dy, want), (name, rep)
            del g
    # SpMM of a matrix without nonzeros: a fill kernel, capturable as well
    E = P.CSR(5, 5, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
    with P.CvrHip(E, precision=1) as h:
        X = torch.ones((5, 8), dtype=torch.float32, device="cuda")
        Y = torch.full((5, 8), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            h.spmm(8, X.data_ptr(), Y.data_ptr(), torch.cuda.current_stream().cuda_stream)
        for rep in range(3):
            Y.fill_(float("nan"))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(Y, torch.zeros_like(Y))
        del g


def test_seeded_sweep_of_shapes_and_options(P, O):
    """40 seeded random matrices (empty rows, single-element rows, rows longer
    than a chunk, ragged lengths, banded or scattered columns) under random
    chunk sizes, column widths and hot-x settings: y equals the emulation of the
    same chunk size bit for bit, and the FP32 layout its FP32 emulation."""
    rng = np.random.default_rng(2024)
    for case in range(40):
        M = int(rng.integers(1, 20000))
        N = int(rng.choice([M, int(rng.integers(1, 200000))]))
        kind = case % 4
        if kind == 0:
            lens = rng.integers(0, 40, M)
        elif kind == 1:
            lens = rng.choice([0, 1, 2, 3, 500], M, p=[0.2, 0.3, 0.2, 0.29, 0.01])
        elif kind == 2:
            lens = np.minimum(rng.zipf(1.8, M), 3000)
        else:
            lens = np.full(M, int(rng.integers(1, 30)))
        lens = np.minimum(lens, N)
        rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        banded = rng.random() < 0.5
        cols = []
        for r, l in enumerate(lens):
            if l == 0:
                continue
            if banded:
                lo = max(0, min(N - l, r * N // max(M, 1) - 200))
                c = lo + np.sort(rng.choice(min(N - lo, max(l, 400)), l, replace=False))
            else:
                c = np.sort(rng.choice(N, l, replace=False))
            cols.append(c)
        ci = (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)
        A = P.CSR(M, N, rp, ci, rng.uniform(-1, 1, int(rp[-1])))
        S = int(rng.integers(1, 33)) * 8
        bits = int(rng.choice([0, 32]))
        hot = int(rng.choice([-1, 0, 64]))
        prec = int(rng.choice([0, 0, 1]))
        _say(f"sweep {case}: M={M} N={N} nnz={A.nnz} S={S} bits={bits} hot={hot} prec={prec}")
        x = P.fill_uniform(N, case)
        with P.CvrHip(A, steps_per_chunk=S, col_index_bits=bits, x_hot_kb=hot, precision=prec) as h:
            y = h.spmv_host(x)
        emu = O.cvr64_build(A, S, prec)
        if prec == 0:
            want = O.cvr64_spmv(A, emu, x)
        else:
            want = O.cvr64_spmv(A, emu, x.astype(np.float32)).astype(np.float64)
        assert np.array_equal(y, want), case


def test_unsorted_and_repeated_columns(P, O):
    """A CSR whose rows list columns out of order and repeat some: the layout
    keeps each row's element order, so y equals the emulation bit for bit
    (both column widths) and the serial oracle within the bound."""
    from matrices import random_csr
    rng = np.random.default_rng(31)
    A = random_csr(P, rng, max_rows=6000)
    cols = A.cols.copy()
    for r in range(A.numRows):
        b, e = A.rowOffsets[r], A.rowOffsets[r + 1]
        if e - b > 1:
            seg = cols[b:e]
            rng.shuffle(seg)
            if e - b > 3:
                seg[-1] = seg[0]  # a repeated column
    B = P.CSR(A.numRows, A.numCols, A.rowOffsets, cols, A.vals)
    x = P.fill_uniform(B.numCols, 4)
    want = O.cvr64_spmv(B, O.cvr64_build(B, 16, 0), x)
    for bits in (16, 32):
        try:
            h = P.CvrHip(B, steps_per_chunk=16, col_index_bits=bits, x_hot_kb=0)
        except P.PavaError as e:
            assert bits == 16 and e.code == -6
            continue
        with h:
            assert np.array_equal(h.spmv_host(x), want), bits
    yref = O.spmv_fma(B, x)
    lens = np.diff(B.rowOffsets)
    assert np.all(np.abs(want - yref) <= 2 * (lens + 64) * 2.0 ** -53 * O.row_abs(B, x))


def test_extreme_shapes(P, O):
    """One row of 3,000,000 elements (hundreds of chunks, thousands of
    recompute pieces), a single column of 100,000 rows, a million rows with
    only the last one non-empty, 200,000 rows of which every 1,000th is
    non-empty, and 1 x 1: FP64 equals the emulation bit for
    bit (both column widths). With FP32 storage and a cancellation threshold
    of 1, every row stays within FP32 input rounding of the oracle, and the long
    row (random signs, so certainly flagged) is recomputed in FP64: the serial
    sum rounded to FP32, within FP64 rounding of it."""
    rng = np.random.default_rng(8)
    n = 3_000_000
    shapes = {
        "one_long_row": P.CSR(1, n, np.array([0, n], np.int32), np.arange(n, dtype=np.int32),
                              rng.uniform(-1, 1, n)),
        "one_column": P.CSR(100000, 1, np.arange(100001, dtype=np.int32), np.zeros(100000, np.int32),
                            rng.uniform(-1, 1, 100000)),
        "last_row_only": P.CSR(1_000_000, 5000, np.concatenate([np.zeros(1_000_000, np.int32), [100]]).astype(np.int32),
                               np.sort(rng.choice(5000, 100, replace=False)).astype(np.int32), rng.uniform(-1, 1, 100)),
        "one_by_one": P.CSR(1, 1, np.array([0, 1], np.int32), np.array([0], np.int32), np.array([-2.5])),
        # every 1,000th of 200,000 rows holds 90 elements: 36 chunks at the
        # automatic size, each zeroing ~174 words of the empty-row bitmap (more
        # than the 64 a wave loads with its chunk's first loads)
        "mostly_empty": P.CSR(200_000, 5000,
                              np.concatenate([[0], np.cumsum(np.where(np.arange(200_000) % 1000 == 0, 90, 0))]).astype(np.int32),
                              np.concatenate([np.sort(rng.choice(5000, 90, replace=False)) for _ in range(200)]).astype(np.int32),
                              rng.uniform(-1, 1, 200 * 90)),
    }
    for name, A in shapes.items():
        x = P.fill_uniform(A.numCols, 3)
        for bits in (0, 32):
            with P.CvrHip(A, col_index_bits=bits) as h:
                S = h.info()["steps_per_chunk"]
                y = h.spmv_host(x)
            _say(f"extreme {name}: S={S} bits={bits}")
            assert np.array_equal(y, O.cvr64_spmv(A, O.cvr64_build(A, S, 0), x)), (name, bits)
        yref = O.spmv_fma(A, x)
        sabs = O.row_abs(A, x)
        with P.CvrHip(A, precision=1, recompute=True, kappa_threshold=1.0) as h:
            y32 = h.spmv_host(x)
            n_rc = h.last_recompute_count()
        assert np.all(np.abs(y32 - yref) <= 4 * 2.0 ** -24 * sabs), name
        if name == "one_long_row":
            assert n_rc == 1
            assert abs(y32[0] - yref[0]) <= 2.0 ** -24 * abs(yref[0]) + 2 * (n + 64) * 2.0 ** -53 * sabs[0]


def test_recompute_needs_fp64_x(P):
    import torch
    A = P.laplacian2d(30, 30)
    with P.CvrHip(A, precision=1, recompute=True) as h:
        dx = torch.zeros(A.numCols, dtype=torch.float32, device="cuda")
        dy = torch.zeros(A.numRows, dtype=torch.float32, device="cuda")
        with pytest.raises(P.PavaError):
            h.spmv_mixed(dx.data_ptr(), dy.data_ptr(), 0)


@pytest.mark.parametrize("L", [1, 4, 16, 64])
def test_csr_hip_format(P, O, mats, L):
    import torch
    for name, A in mats.items():
        x = P.fill_uniform(A.numCols, 9)
        yref = O.spmv_fma(A, x)
        lens = np.diff(A.rowOffsets)
        sabs = O.row_abs(A, x)
        with P.CsrHip(A, lanes_per_row=L) as h:
            dx = torch.from_numpy(x).cuda()
            dy = torch.full((A.numRows,), float("nan"), dtype=torch.float64, device="cuda")
            h.spmv(dx.data_ptr(), dy.data_ptr())
            torch.cuda.synchronize()
            y = dy.cpu().numpy()
        if L == 1:
            assert np.array_equal(y, yref), name  # one lane per row: the oracle's order
        assert np.all(np.abs(y - yref) <= 2 * (lens + 64) * 2.0 ** -53 * sabs), name


def test_time_spmv_stats(P):
    """hipEvent timing: the main kernel's min <= median <= max, its mean
    inside that range, and a whole SpMV at least as long as the kernel."""
    import torch
    A = P.stencil27(48, 0.01, 0)
    x = torch.from_numpy(P.fill_uniform(A.numCols, 1)).cuda()
    y = torch.empty(A.numRows, dtype=torch.float64, device="cuda")
    with P.CvrHip(A) as h:
        st = h.time_spmv_stats(x.data_ptr(), y.data_ptr(), 3, 101)
        tot, main = h.time_spmv(x.data_ptr(), y.data_ptr(), 3, 50)
    _say("timing", st, tot, main)
    assert 0 < st["main_min"] <= st["main_median"] <= st["main_max"]
    assert st["main_min"] <= st["main_mean"] <= st["main_max"]
    assert st["ms_total"] >= 0.5 * st["main_mean"] and tot > 0 and main > 0

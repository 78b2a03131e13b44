"""The CSR5-64 CPU emulation (oracle/oracle.c orc_csr5_*), pinned on the CPU.

Layout: every array is recomputed here from its definition by brute force
(transposed tiles, row-start flags, per-lane row-start offsets, per-tile bases,
rows spanning tiles). SpMV: rows wholly inside one lane are summed exactly as
the serial oracle (bit-identical); every row is within the blocked-sum bound;
integer-valued inputs give exact results."""
import numpy as np
import pytest

from matrices import edge_cases


def _brute_layout(A, sigma):
    T = 64 * sigma
    nnz = A.nnz
    nt = -(-nnz // T) if nnz else 0
    rp = A.rowOffsets.astype(np.int64)
    starts = {int(rp[r]): r for r in range(A.numRows) if rp[r + 1] > rp[r]}
    vals = np.zeros(max(nt * T, 1))
    cols = np.zeros(max(nt * T, 1), np.int32)
    flags = np.zeros(max(nt * 64, 1), np.uint32)
    yoff = np.zeros(max(nt * 64, 1), np.int32)
    base = np.zeros(max(nt, 1), np.int32)
    fix = np.full(max(2 * nt, 2), -1, np.int32)
    cont = np.zeros(max(nt, 1), np.uint8)
    row_of = np.repeat(np.arange(A.numRows), np.diff(rp))  # row of each element
    b = 0
    for t in range(nt):
        base[t] = b
        acc = 0
        for l in range(64):
            f = 0
            yoff[t * 64 + l] = acc
            for i in range(sigma):
                e = t * T + l * sigma + i
                if e < nnz:
                    vals[t * T + i * 64 + l] = A.vals[e]
                    cols[t * T + i * 64 + l] = A.cols[e]
                    if e in starts:
                        f |= 1 << i
                        acc += 1
            flags[t * 64 + l] = f
        b += acc
        e0, e1 = t * T, min((t + 1) * T, nnz)
        rf = row_of[e0]
        if rp[rf] < e0 and rp[rf + 1] <= e1:
            fix[2 * t], fix[2 * t + 1] = rf, rp[rf] // T
        rl = row_of[e1 - 1]
        cont[t] = rp[rl + 1] > e1
    return dict(vals=vals, cols=cols, flags=flags, yoff=yoff, tile_base=base, fix=fix, cont=cont, ntiles=nt)


@pytest.fixture(scope="module")
def mats(P):
    m = dict(edge_cases(P))
    m["rmat10"] = P.rmat(10, 16, seed=7)
    m["lap"] = P.laplacian2d(20, 30)
    return m


@pytest.mark.parametrize("sigma", [1, 3, 16, 32])
def test_csr5_layout_matches_definition(P, O, mats, sigma):
    for name, A in mats.items():
        lay = O.csr5_build(A, sigma)
        want = _brute_layout(A, sigma)
        assert lay["ntiles"] == want["ntiles"], name
        n = lay["ntiles"]
        for k, size in [("vals", n * 64 * sigma), ("cols", n * 64 * sigma), ("flags", n * 64), ("yoff", n * 64),
                        ("tile_base", n), ("fix", 2 * n), ("cont", n)]:
            assert np.array_equal(lay[k][:size], want[k][:size]), (name, k)
        ne = np.nonzero(np.diff(A.rowOffsets))[0]
        assert np.array_equal(lay["nerows"][: len(ne)], ne), name
        # every non-empty row starts exactly once
        assert sum(bin(int(f)).count("1") for f in lay["flags"][: n * 64]) == len(ne)


@pytest.mark.parametrize("sigma", [1, 4, 16, 32])
def test_csr5_spmv_order_and_bounds(P, O, mats, sigma):
    for name, A in mats.items():
        lay = O.csr5_build(A, sigma)
        x = P.fill_uniform(A.numCols, 5)
        y = O.csr5_spmv(A, lay, x)
        yref = O.spmv_fma(A, x)
        rp = A.rowOffsets.astype(np.int64)
        # rows wholly inside one lane (first and last element in the same lane) are the serial sum
        T = 64 * sigma
        lens = np.diff(rp)
        nz = lens > 0
        lane_first = rp[:-1] // sigma
        lane_last = (rp[1:] - 1) // sigma
        inside = nz & (lane_first == lane_last)
        assert np.array_equal(y[inside], yref[inside]), name
        assert np.all(y[~nz] == 0.0), name
        sabs = O.row_abs(A, x)
        bound = 2 * (lens + 64 + lens // T + 2) * 2.0 ** -53 * sabs
        assert np.all(np.abs(y - yref) <= bound), name


def test_csr5_exact_on_integer_inputs(P, O, mats):
    """Small integers: every partial sum is exact, so any order gives the oracle's y."""
    for name, A in mats.items():
        B = P.CSR(A.numRows, A.numCols, A.rowOffsets.copy(), A.cols.copy(),
                  (np.arange(A.nnz) % 7 - 3).astype(np.float64))
        x = (np.arange(B.numCols) % 3).astype(np.float64)
        for sigma in (2, 16):
            assert np.array_equal(O.csr5_spmv(B, O.csr5_build(B, sigma), x), O.spmv_fma(B, x)), name


def test_csr5_long_row_across_tiles(P, O):
    """One row spanning several tiles, neighbours on both sides: tail + whole
    tiles + head in tile order."""
    n = 3000
    lens = np.ones(10, np.int64)
    lens[4] = n
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    rng = np.random.default_rng(3)
    A = P.CSR(10, 64, rp, rng.integers(0, 64, rp[-1]).astype(np.int32), rng.uniform(-1, 1, rp[-1]))
    x = P.fill_uniform(64, 2)
    lay = O.csr5_build(A, 4)  # tiles of 256: the long row spans 12 tiles
    assert (lay["fix"][0::2] == 4).sum() == 1
    y = O.csr5_spmv(A, lay, x)
    # the long row by hand: lane sums -> lane order within tiles -> tile order
    T, s = 256, 4
    vals, cols = A.vals, A.cols
    b, e = rp[4], rp[5]
    parts = []
    for t0 in range(b // T, (e - 1) // T + 1):
        acc_t = None
        for l in range(64):
            lo, hi = max(b, t0 * T + l * s), min(e, t0 * T + (l + 1) * s)
            if lo >= hi:
                continue
            a = 0.0
            for k in range(lo, hi):  # (unfused here: the tolerance below covers it)
                a = a + vals[k] * x[cols[k]]
            acc_t = a if acc_t is None else acc_t + a
        parts.append(acc_t)
    want = parts[0]
    for p in parts[1:]:
        want = want + p
    assert abs(y[4] - want) <= 1e-15 * np.abs(vals[b:e] * x[cols[b:e]]).sum()

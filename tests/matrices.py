"""Seeded test matrices covering the layout's edge cases (test infrastructure)."""
import numpy as np


def csr_from_dense_pattern(P, m, n, rows_cols, seed=0):
    """rows_cols: list (per row) of column lists."""
    rng = np.random.default_rng(seed)
    rp = np.concatenate([[0], np.cumsum([len(c) for c in rows_cols])]).astype(np.int32)
    ci = np.concatenate([np.sort(np.asarray(c, dtype=np.int32)) for c in rows_cols]) if rp[-1] else np.zeros(0, np.int32)
    va = rng.uniform(-1, 1, int(rp[-1]))
    return P.CSR(m, n, rp, ci, va)


def edge_cases(P):
    """name -> CSR."""
    out = {}
    rng = np.random.default_rng(11)
    out["lap_40x30"] = P.laplacian2d(40, 30)
    out["rmat_12"] = P.rmat(12, 8, seed=3)              # empty rows, long rows
    out["one_dense_row"] = csr_from_dense_pattern(P, 50, 5000, [[] for _ in range(20)] + [list(range(5000))] + [[] for _ in range(29)])
    out["identity_777"] = P.CSR(777, 777, np.arange(778), np.arange(777), np.ones(777))
    out["single_entry"] = P.CSR(3, 3, np.array([0, 0, 1, 1]), np.array([2]), np.array([2.5]))
    out["empty_edges"] = csr_from_dense_pattern(
        P, 300, 300, [[] for _ in range(40)] + [list(rng.choice(300, rng.integers(1, 40), replace=False)) for _ in range(220)] + [[] for _ in range(40)])
    out["few_rows_long"] = csr_from_dense_pattern(P, 5, 4000, [list(rng.choice(4000, L, replace=False)) for L in (1500, 7, 2500, 3, 999)])
    out["nnz_not_mult_64"] = csr_from_dense_pattern(P, 97, 97, [list(rng.choice(97, rng.integers(0, 9), replace=False)) for _ in range(97)])
    out["banded_mixed"] = P.synthetic(3000, P.RL_POISSON, 20.0, col_struct=P.COL_MIXED, half_band=100, band_frac=0.7, seed=9)
    out["powerlaw_rows"] = P.synthetic(4000, P.RL_POWERLAW, 2.0, shape=1.9, col_struct=P.COL_UNIFORM, max_len=3000, seed=4)
    out["empty_matrix"] = P.CSR(10, 10, np.zeros(11, np.int32), np.zeros(0, np.int32), np.zeros(0))
    out["rect_wide"] = csr_from_dense_pattern(P, 20, 10000, [list(rng.choice(10000, 300, replace=False)) for _ in range(20)])
    return out


def split_rows(A, lay, S):
    """Rows whose sum the layout splits into several partial sums."""
    import bisect
    rp = [int(v) for v in A.rowOffsets]
    nc = int(lay["nchunks"]) if "nchunks" in lay else int(lay["info"]["nchunks"])
    out = set()
    C = 64 * S
    for c in range(nc):
        b, e = c * C, min((c + 1) * C, A.nnz)
        tab = lay["split_tab"][c * 144:(c + 1) * 144]
        for g in range(int(lay["split_n"][c])):
            code = int(tab[2 * g])
            if code >= 0:
                out.add(code)
            elif code == -1:
                out.add(bisect.bisect_right(rp, b) - 1)
            elif code == -2:
                out.add(bisect.bisect_right(rp, e - 1) - 1)
    return out


def random_csr(P, rng, max_rows=20000):
    """A seeded random CSR: ragged / empty / single-element / long rows,
    banded or scattered sorted columns, values U[-1, 1]."""
    M = int(rng.integers(1, max_rows))
    N = int(rng.choice([M, int(rng.integers(1, 200000))]))
    lens = np.minimum(rng.choice([0, 1, 3, 17, 60, 900], M, p=[0.1, 0.2, 0.3, 0.3, 0.095, 0.005]), N)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    banded = rng.random() < 0.5
    cols = []
    for r, l in enumerate(lens):
        if l == 0:
            continue
        if banded:
            lo = max(0, min(N - l, r * N // M - 150))
            cols.append(lo + np.sort(rng.choice(min(N - lo, max(l, 400)), l, replace=False)))
        else:
            cols.append(np.sort(rng.choice(N, l, replace=False)))
    ci = (np.concatenate(cols) if cols else np.zeros(0)).astype(np.int32)
    return P.CSR(M, N, rp, ci, rng.uniform(-1, 1, int(rp[-1])))

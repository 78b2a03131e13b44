"""Pins the CPU MLP oracle (oracle/mlp_ref.py) by finite differences, and
checks the selector's pure-host helpers (CPU only)."""
import numpy as np

from oracle import mlp_ref


def test_mlp_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    I, H, O, B = 5, 7, 3, 11
    n = H * I + H + H * H + H + O * H + O
    p = rng.normal(scale=0.5, size=n)
    X = rng.normal(size=(B, I))
    T = rng.normal(size=(B, O))
    M = (rng.random((B, O)) < 0.8).astype(float)
    L, g = mlp_ref.loss_grad(p, X, T, M, I, H, O)
    assert abs(L - mlp_ref.loss(p, X, T, M, I, H, O)) < 1e-12
    eps = 1e-6
    for i in rng.choice(n, 40, replace=False):
        e = np.zeros(n)
        e[i] = eps
        fd = (mlp_ref.loss(p + e, X, T, M, I, H, O) - mlp_ref.loss(p - e, X, T, M, I, H, O)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_mlp_oracle_layout_sizes():
    I, H, O = 16, 64, 13
    n = H * I + H + H * H + H + O * H + O
    parts = mlp_ref.unpack(np.arange(n, dtype=float), I, H, O)
    assert [a.shape for a in parts] == [(H, I), (H,), (H, H), (H,), (O, H), (O,)]
    assert parts[1][0] == H * I  # b1 follows W1


def test_targets_and_slowdowns():
    from pava_amd import selector as S
    t = np.array([[2.0, 1.0, np.inf], [4.0, 8.0, 2.0]])
    T, M = S.targets(t)
    assert np.allclose(T, [[1, 0, 0], [1, 2, 0]]) and np.array_equal(M, [[1, 1, 0], [1, 1, 1]])
    assert np.allclose(S.slowdowns(t, np.array([0, 2])), [2.0, 1.0])


def test_features_closed_forms(P):
    from pava_amd import selector as S
    A = P.laplacian2d(64, 64)
    f = S.features(A)
    assert abs(f[0] - 12) < 1e-6 and abs(f[15] - 12) < 1e-6       # log2 rows / cols
    assert abs(f[1] - np.log2(A.nnz)) < 1e-5
    mean = A.nnz / A.numRows
    assert abs(f[2] - np.log2(1 + mean)) < 1e-5
    assert abs(f[4] - np.log2(6)) < 1e-6                          # max row length 5
    assert f[6] == 0 and abs(f[7] - 4 / 4096) < 1e-7               # no empty rows; 4 corners of length 3
    assert f[10] == 1.0                                            # all within +-1024 of the diagonal
    E = P.CSR(4, 4, np.array([0, 0, 2, 2, 3]), np.array([0, 3, 1]), np.ones(3))
    g = S.features(E)
    assert abs(g[6] - 0.5) < 1e-6 and abs(g[7] - 1.0) < 1e-6       # 2 of 4 rows empty, all < 4


def test_config_space():
    from pava_amd import selector as S
    names = [c.name for c in S.CONFIGS]
    assert len(names) == len(set(names)) == 21
    assert S.DEFAULT in S.CONFIGS
    # every axis the autotuner searches is present: format, precision, tile, x staging
    assert {c.fmt for c in S.CONFIGS} == {"cvr", "csr5", "csr"}
    assert {c.prec for c in S.CONFIGS} == {"f64", "f32rc", "f32rc16"}
    assert {c.xs for c in S.CONFIGS} == {-1, 0, 2048}
    assert len({c.tile for c in S.CONFIGS if c.fmt == "cvr"}) >= 3


def test_committed_selector_dataset_is_consistent():
    """The autotune dataset matches the config space and the seeded family:
    every index once, features recomputed on the CPU for a few matrices agree."""
    from pathlib import Path
    from pava_amd import selector as S
    from pava_amd import selector_data as D
    d = np.load(Path(S.__file__).parent / "data" / "selector_dataset.npz", allow_pickle=False)
    assert [str(n) for n in d["configs"]] == [c.name for c in S.CONFIGS]
    n = D.N_TRAIN + D.N_HELD + D.N_LARGE + D.N_LARGE_TRAIN
    assert np.array_equal(d["index"], np.arange(n))
    # training covers rows 1e4 to 5e6: the large training matrices are in
    assert d["rows"][D.LARGE0 + D.N_LARGE:].min() >= 1_000_000 and d["rows"].max() <= 5_000_000
    assert d["features"].shape == (n, S.N_FEATURES) and d["ms"].shape == (n, len(S.CONFIGS))
    assert np.all(np.isfinite(d["ms"]).sum(1) >= 12) and np.all(d["ms"][np.isfinite(d["ms"])] > 0)
    for i in (0, 7, 2001):
        A = D.matrix(i)
        assert A.numRows == d["rows"][i] and A.nnz == d["nnz"][i]
        assert np.allclose(S.features(A), d["features"][i], rtol=1e-5, atol=1e-6)


def test_selector_covers_only_the_training_range():
    """Outside the per-feature training range (widened by 5% of its width) the
    selector reports no cover, and its caller autotunes instead."""
    from pava_amd import selector as S
    sel = S.Selector(None, np.zeros(3), np.ones(3), lo=np.zeros(3), hi=np.array([1.0, 2.0, 4.0]))
    F = np.array([[0.5, 1.0, 2.0], [1.06, 1.0, 2.0], [1.04, 2.09, -0.19], [0.5, 1.0, 4.21]])
    assert sel.covers(F).tolist() == [True, False, True, False]

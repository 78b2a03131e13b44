"""GPU tests of the selector: MLP kernels vs the CPU oracle, training, and the
learned choice vs the fixed default on held-out matrices (needs an MI355X)."""
import sys

import numpy as np
import pytest

from oracle import mlp_ref

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_mlp_forward_and_gradient_match_oracle():
    from pava_amd import selector as S
    rng = np.random.default_rng(3)
    I, H, O, B = 16, 64, 13, 97
    m = S.Mlp(I, H, O, seed=5)
    p = m.params()
    X = rng.normal(size=(B, I)).astype(np.float32)
    T = rng.normal(size=(B, O)).astype(np.float32)
    M = (rng.random((B, O)) < 0.85).astype(np.float32)
    Yg = m.predict(X)
    Yr, _ = mlp_ref.forward(p, X, I, H, O)
    assert np.allclose(Yg, Yr, rtol=1e-4, atol=1e-4)
    Lg, gg = m.loss_grad(X, T, M)
    Lr, gr = mlp_ref.loss_grad(p, X, T, M, I, H, O)
    assert abs(Lg - Lr) <= 1e-4 * abs(Lr)
    scale = np.abs(gr).max()
    assert np.max(np.abs(gg - gr)) <= 1e-4 * scale
    m.close()


def test_mlp_training_reduces_loss():
    from pava_amd import selector as S
    rng = np.random.default_rng(4)
    X = rng.normal(size=(512, 16)).astype(np.float32)
    W = rng.normal(size=(16, 13))
    T = np.tanh(X @ W / 4).astype(np.float32)
    M = np.ones_like(T)
    m = S.Mlp(n_out=13, seed=1)
    hist = m.train(X, T, M, epochs=60, batch=64, lr=3e-3)
    _say("loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.25 * hist[0]
    m.close()


def test_autotune_best_config_reproducible():
    """Exhaustive autotuning of seeded matrices of the selector family, run
    twice: where the first run's best config beats the runner-up by more than
    30% the second run picks it again, and everywhere the first choice is
    within 10% of the second run's best. Configs that cannot apply (16-bit
    columns on chunks spanning more than 65,536 columns) are inf in both."""
    import pava_amd as P
    from pava_amd import selector as S
    from pava_amd import selector_data as D
    clear = 0
    for i in (3, 17, 93, 143, 777, 1650):  # close and clear winners in the committed dataset
        A = D.matrix(i)
        x = P.fill_uniform(A.numCols, i)
        t1 = S.autotune(A, x, iters=10)
        t2 = S.autotune(A, x, iters=10)
        assert np.array_equal(np.isfinite(t1), np.isfinite(t2))
        b1, b2 = int(np.argmin(t1)), int(np.argmin(t2))
        runner_up = np.partition(t1[np.isfinite(t1)], 1)[1]
        _say(f"matrix {i}: rows {A.numRows} nnz {A.nnz}: best {S.CONFIGS[b1].name} then {S.CONFIGS[b2].name}, "
             f"margin {runner_up / t1[b1]:.2f}")
        assert t2[b1] <= 1.10 * t2[b2]
        if runner_up > 1.30 * t1[b1]:
            clear += 1
            assert b1 == b2
    assert clear >= 1


def test_learned_selector_beats_fixed_default():
    """Autotune small seeded synthetic matrices, train on 3/4, evaluate the
    chosen config's measured slowdown vs the exhaustive best on the rest."""
    import pava_amd as P
    from pava_amd import selector as S
    rng = np.random.default_rng(2026)
    F, TT = [], []
    for i in range(64):
        prm = S.gen_params(rng, max_rows=60_000)
        A = P.synthetic(**prm)
        x = P.fill_uniform(A.numCols, i)
        F.append(S.features(A))
        TT.append(S.autotune(A, x, iters=3))
    F = np.array(F)
    TT = np.array(TT)
    tr, te = slice(0, 48), slice(48, 64)
    sel, hist = S.Selector.fit(F[tr], TT[tr], epochs=300)
    choice = sel.choose(F[te], valid=np.isfinite(TT[te]))
    sd_model = S.slowdowns(TT[te], choice)
    d = S.CONFIGS.index(S.DEFAULT)
    sd_default = S.slowdowns(TT[te], np.full(16, d))
    gm = lambda v: float(np.exp(np.mean(np.log(v))))
    _say("held-out geo-mean slowdown: model", gm(sd_model), "default", gm(sd_default),
         "loss", hist[0], "->", hist[-1])
    assert np.all(np.isfinite(sd_model))
    assert gm(sd_model) <= gm(sd_default)


def test_trained_selector_gain_on_committed_dataset():
    """The MLP trained on the GPU on the committed autotune dataset (2,000
    matrices) picks configs whose measured times on the 200 held-out matrices
    beat the library's default configuration and come close to the exhaustive
    best."""
    from pathlib import Path
    from pava_amd import selector as S
    from pava_amd import selector_data as D
    d = np.load(Path(S.__file__).parent / "data" / "selector_dataset.npz", allow_pickle=False)
    idx, F, T = d["index"], d["features"], d["ms"]
    tr = np.asarray([D.is_train(int(i)) for i in idx])
    ho = (idx >= D.N_TRAIN) & (idx < D.LARGE0)
    sel, hist = S.Selector.fit(F[tr], T[tr], epochs=200, seed=0)
    choice = sel.choose(F[ho], valid=np.isfinite(T[ho]))
    sel.mlp.close()
    gm = lambda v: float(np.exp(np.mean(np.log(v))))
    sd_model = S.slowdowns(T[ho], choice)
    sd_default = S.slowdowns(T[ho], np.full(int(ho.sum()), S.CONFIGS.index(S.DEFAULT)))
    _say("held-out geo-mean slowdown vs exhaustive best: model", gm(sd_model), "default", gm(sd_default),
         "loss", hist[0], "->", hist[-1])
    assert hist[-1] < 0.5 * hist[0]
    # the default is the library's own (automatic chunk size); the model must
    # recover most of what separates it from the exhaustive best
    assert gm(sd_model) - 1 < 0.4 * (gm(sd_default) - 1)
    assert gm(sd_model) < 1.1


def test_selector_gain_on_fresh_held_out_timings():
    """Held-out matrices re-timed in this run: 12 seeded picks of the 200
    held-out matrices (1e4-1e6 rows) and the two smallest of the 20 large
    held-out ones (1e6-5e6 rows) are autotuned over every config now. The MLP
    trained on the committed dataset's training matrices chooses a config for
    each from its features; on these fresh timings its geo-mean slowdown
    against the exhaustive best beats the library default's and stays within
    15% of the best."""
    from pathlib import Path
    import pava_amd as P
    from pava_amd import selector as S
    from pava_amd import selector_data as D
    d = np.load(Path(S.__file__).parent / "data" / "selector_dataset.npz", allow_pickle=False)
    idx = d["index"]
    tr = np.asarray([D.is_train(int(i)) for i in idx])
    sel, hist = S.Selector.fit(d["features"][tr], d["ms"][tr], epochs=400, seed=0)
    rng = np.random.default_rng(7)
    picks = sorted(int(i) for i in rng.choice(np.arange(D.N_TRAIN, D.LARGE0), 12, replace=False))
    large = [int(i) for i in range(D.LARGE0, D.LARGE0 + D.N_LARGE)]
    picks += sorted(large, key=lambda i: int(d["nnz"][i]))[:2]
    F, T = [], []
    for i in picks:
        A = D.matrix(i)
        F.append(S.features(A))
        T.append(S.autotune(A, P.fill_uniform(A.numCols, i), iters=5))
        _say(f"matrix {i}: {A.numRows} rows {A.nnz} nnz, best {S.CONFIGS[int(np.argmin(T[-1]))].name}")
    F, T = np.array(F), np.array(T)
    choice = sel.choose(F, valid=np.isfinite(T))
    sel.mlp.close()
    gm = lambda v: float(np.exp(np.mean(np.log(v))))
    sd_model = S.slowdowns(T, choice)
    sd_default = S.slowdowns(T, np.full(len(picks), S.CONFIGS.index(S.DEFAULT)))
    _say("fresh held-out geo-mean slowdown vs exhaustive best: model", gm(sd_model), "default", gm(sd_default),
         "choices", [S.CONFIGS[int(c)].name for c in choice])
    assert gm(sd_model) < gm(sd_default)
    assert gm(sd_model) < 1.15

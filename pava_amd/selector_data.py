"""Selector dataset and evaluation (command line; needs the GPU for ``gen`` and
``eval``).

Matrix i of the family is ``synthetic(**gen_params(default_rng([SEED, i])))``:
indices [0, N_TRAIN) train, [N_TRAIN, N_TRAIN + N_HELD) are held out, and
[N_TRAIN + N_HELD, N_TRAIN + N_HELD + N_LARGE) is a further held-out set of
large matrices (1e6 to 5e6 rows); the N_LARGE_TRAIN matrices after those are
large training matrices. The other sets use rows 1e4 to 1e6, so training
covers rows 1e4 to 5e6.

  python -m pava_amd.selector_data gen --first 0 --count 500 --out DIR
      autotunes matrices [first, first+count) over every config and writes
      DIR/shard_<first>_<count>.npz (features, per-config ms, sizes, params)
  python -m pava_amd.selector_data retime --config NAME --first F --count C --out DIR
      re-times one config (after a change to it) for matrices [F, F+C):
      DIR/retime_<NAME>_<F>_<C>.npz
  python -m pava_amd.selector_data patch --data FILE.npz --config NAME [--replaces OLD] RETIME.npz ...
      replaces that config's column of the dataset with the re-timed values
      (with --replaces: the column of config OLD, which becomes NAME)
  python -m pava_amd.selector_data merge --out FILE.npz SHARD.npz ...
  python -m pava_amd.selector_data eval --data FILE.npz --out REPORT.json
      trains the MLP on the GPU (the library's own kernels) on the training
      set and reports, on each held-out set, the measured slowdown of the
      chosen config against the exhaustive best and of the fixed default
      against the same best, the throughput of the choice against plain FP64
      CVR-HIP, and the chosen config's error against the FP64 reference SpMV.
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

import numpy as np

SEED = 20260
N_TRAIN, N_HELD, N_LARGE = 2000, 200, 20
N_LARGE_TRAIN = 300
LARGE0 = N_TRAIN + N_HELD  # first index of the large (1e6-5e6 rows) matrices


def params_for(i: int) -> dict:
    from . import selector as S
    rng = np.random.default_rng([SEED, i])
    if i >= LARGE0:
        return S.gen_params(rng, max_rows=5_000_000, min_rows=1_000_000)
    return S.gen_params(rng, max_rows=1_000_000)


def is_train(i: int) -> bool:
    return i < N_TRAIN or i >= LARGE0 + N_LARGE


def is_held_out_large(i: int) -> bool:
    return LARGE0 <= i < LARGE0 + N_LARGE


def matrix(i: int):
    import pava_amd as P
    return P.synthetic(**params_for(i))


def gen(first: int, count: int, out: Path, iters: int = 5) -> Path:
    import pava_amd as P
    from . import selector as S
    F, T, rows, nnz, idx = [], [], [], [], []
    t0 = time.perf_counter()
    for i in range(first, first + count):
        A = matrix(i)
        x = P.fill_uniform(A.numCols, i)
        F.append(S.features(A))
        T.append(S.autotune(A, x, iters=iters))
        rows.append(A.numRows)
        nnz.append(A.nnz)
        idx.append(i)
        if (i - first) % 10 == 9 or i == first + count - 1 or A.nnz > 20_000_000:
            print(f"[selector] {i - first + 1}/{count} matrices ({time.perf_counter() - t0:.0f}s, last: "
                  f"{A.numRows} rows {A.nnz} nnz)", flush=True)
    out.mkdir(parents=True, exist_ok=True)
    path = out / f"shard_{first}_{count}.npz"
    np.savez_compressed(path, index=np.array(idx), features=np.array(F, np.float32), ms=np.array(T),
                        rows=np.array(rows), nnz=np.array(nnz), configs=np.array([c.name for c in S.CONFIGS]))
    return path


def retime(name: str, first: int, count: int, out: Path, iters: int = 5) -> Path:
    import pava_amd as P
    from . import selector as S
    cfg = [c for c in S.CONFIGS if c.name == name]
    if not cfg:
        raise SystemExit(f"no config named {name}")
    idx, ms = [], []
    t0 = time.perf_counter()
    for i in range(first, first + count):
        A = matrix(i)
        ms.append(S.autotune(A, P.fill_uniform(A.numCols, i), iters=iters, configs=cfg)[0])
        idx.append(i)
        if (i - first) % 50 == 49 or i == first + count - 1:
            print(f"[selector] retime {name}: {i - first + 1}/{count} ({time.perf_counter() - t0:.0f}s)", flush=True)
    out.mkdir(parents=True, exist_ok=True)
    path = out / f"retime_{name}_{first}_{count}.npz"
    np.savez_compressed(path, index=np.array(idx), ms=np.array(ms), config=np.array(name))
    return path


def patch(data: Path, name: str, parts: list[Path], replaces: str | None = None) -> None:
    d = dict(np.load(data, allow_pickle=False))
    names = [str(n) for n in d["configs"]]
    j = names.index(replaces or name)
    names[j] = name
    d["configs"] = np.array(names)
    pos = {int(i): k for k, i in enumerate(d["index"])}
    ms = d["ms"].copy()
    n = 0
    for p in parts:
        q = np.load(p, allow_pickle=False)
        assert str(q["config"]) == name, (p, str(q["config"]))
        for i, v in zip(q["index"], q["ms"]):
            ms[pos[int(i)], j] = v
            n += 1
    d["ms"] = ms
    np.savez_compressed(data, **d)
    print(f"[selector] patched {n} timings of {name}")


def merge(out: Path, shards: list[Path]) -> None:
    parts = [np.load(p, allow_pickle=False) for p in shards]
    names = parts[0]["configs"]
    for q in parts:
        assert np.array_equal(q["configs"], names), "shards from different config spaces"
    cat = {k: np.concatenate([q[k] for q in parts]) for k in ("index", "features", "ms", "rows", "nnz")}
    order = np.argsort(cat["index"])
    cat = {k: v[order] for k, v in cat.items()}
    assert len(np.unique(cat["index"])) == len(cat["index"]), "duplicate matrices"
    out.parent.mkdir(parents=True, exist_ok=True)
    np.savez_compressed(out, configs=names, seed=np.array(SEED), **cat)


def _gm(v) -> float:
    return float(np.exp(np.mean(np.log(v))))


def _summary(sd: np.ndarray) -> dict:
    return {"geo_mean": _gm(sd), "mean": float(sd.mean()), "median": float(np.median(sd)),
            "p90": float(np.percentile(sd, 90)), "max": float(sd.max()),
            "within_10pct": float(np.mean(sd <= 1.1))}


def _chosen_error(i: int, cfg) -> float:
    """Max relative row error of config `cfg` on matrix i against the FP64
    reference SpMV (the driver's verification routine)."""
    import torch
    import pava_amd as P
    from . import selector as S
    A = matrix(i)
    x = P.fill_uniform(A.numCols, i)
    yref = P.reference_spmv(A, x)
    h = S.build(A, cfg)
    try:
        dev = torch.device("cuda:0")
        x64 = torch.from_numpy(x).to(dev)
        if cfg.fmt in ("csr", "csr5") or cfg.prec == "f64":
            y = torch.empty(A.numRows, dtype=torch.float64, device=dev)
            h.spmv(x64.data_ptr(), y.data_ptr(), 0)
        else:
            y = torch.empty(A.numRows, dtype=torch.float32, device=dev)
            h.spmv_mixed(x64.float().data_ptr(), y.data_ptr(), x64.data_ptr(), 0)
        torch.cuda.synchronize()
        yy = y.double().cpu().numpy()
    finally:
        h.close()
    d = np.abs(yy - yref)
    den = np.abs(yref)
    rel = np.where(den > 0, d / np.where(den > 0, den, 1.0), d)
    return float(rel.max()) if rel.size else 0.0


def evaluate(data: Path, epochs: int = 400, seed: int = 0, check_errors: int = 200) -> dict:
    from . import selector as S
    d = np.load(data, allow_pickle=False)
    names = [str(n) for n in d["configs"]]
    assert names == [c.name for c in S.CONFIGS], "dataset built for another config space"
    idx, F, T = d["index"], d["features"], d["ms"]
    tr = np.asarray([is_train(int(i)) for i in idx])
    sets = {"held_out": (idx >= N_TRAIN) & (idx < LARGE0),
            "held_out_large": np.asarray([is_held_out_large(int(i)) for i in idx])}
    t0 = time.perf_counter()
    sel, hist = S.Selector.fit(F[tr], T[tr], epochs=epochs, seed=seed)
    report = {"configs": names, "n_train": int(tr.sum()), "n_train_large": int(np.sum(idx >= LARGE0 + N_LARGE)),
              "epochs": epochs,
              "train_loss_first_last": [float(hist[0]), float(hist[-1])],
              "train_s": round(time.perf_counter() - t0, 2)}
    dflt = S.CONFIGS.index(S.DEFAULT)
    plain = dflt  # plain FP64 CVR-HIP: the library's default configuration
    for name, m in sets.items():
        if not m.any():
            continue
        Tm = T[m]
        valid = np.isfinite(Tm)
        choice = sel.choose(F[m], valid=valid)
        sd_model = S.slowdowns(Tm, choice)
        sd_default = S.slowdowns(Tm, np.full(len(Tm), dflt))
        speed = Tm[:, plain] / Tm[np.arange(len(Tm)), choice]   # nnz/s of the choice over plain FP64 CVR-HIP
        best_speed = Tm[:, plain] / np.where(valid, Tm, np.inf).min(1)
        hist_choice = {names[k]: int(np.sum(choice == k)) for k in np.unique(choice)}
        errs = [_chosen_error(int(i), S.CONFIGS[int(c)]) for i, c in list(zip(idx[m], choice))[:check_errors]]
        report[name] = {"n": int(m.sum()), "slowdown_model_vs_best": _summary(sd_model),
                        "slowdown_default_vs_best": _summary(sd_default),
                        "speedup_model_vs_plain_fp64_cvr": {"geo_mean": _gm(speed), "median": float(np.median(speed)),
                                                            "min": float(speed.min()), "max": float(speed.max())},
                        "speedup_best_vs_plain_fp64_cvr_geo_mean": _gm(best_speed),
                        "choices": hist_choice,
                        "chosen_config_max_rel_error_vs_fp64": {"max": float(max(errs)),
                                                                "median": float(np.median(errs)),
                                                                "matrices_checked": len(errs)}}
        print(f"[selector] {name}: model geo-mean slowdown {_gm(sd_model):.3f}, default {_gm(sd_default):.3f}, "
              f"speedup vs plain FP64 {_gm(speed):.2f}x, max error {max(errs):.2e}", flush=True)
    sel.mlp.close()
    return report


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pava_amd.selector_data")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen")
    g.add_argument("--first", type=int, required=True)
    g.add_argument("--count", type=int, required=True)
    g.add_argument("--out", type=Path, required=True)
    g.add_argument("--iters", type=int, default=5)
    rt = sub.add_parser("retime")
    rt.add_argument("--config", required=True)
    rt.add_argument("--first", type=int, required=True)
    rt.add_argument("--count", type=int, required=True)
    rt.add_argument("--out", type=Path, required=True)
    pt = sub.add_parser("patch")
    pt.add_argument("--data", type=Path, required=True)
    pt.add_argument("--config", required=True)
    pt.add_argument("--replaces", default=None)
    pt.add_argument("parts", type=Path, nargs="+")
    m = sub.add_parser("merge")
    m.add_argument("--out", type=Path, required=True)
    m.add_argument("shards", type=Path, nargs="+")
    e = sub.add_parser("eval")
    e.add_argument("--data", type=Path, required=True)
    e.add_argument("--out", type=Path, required=True)
    e.add_argument("--epochs", type=int, default=400)
    e.add_argument("--check-errors", type=int, default=200)
    a = ap.parse_args(argv)
    if a.cmd == "gen":
        print(gen(a.first, a.count, a.out, a.iters))
    elif a.cmd == "retime":
        print(retime(a.config, a.first, a.count, a.out))
    elif a.cmd == "patch":
        patch(a.data, a.config, a.parts, a.replaces)
    elif a.cmd == "merge":
        merge(a.out, a.shards)
    else:
        rep = evaluate(a.data, a.epochs, check_errors=a.check_errors)
        a.out.parent.mkdir(parents=True, exist_ok=True)
        a.out.write_text(json.dumps(rep, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())

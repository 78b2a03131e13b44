#!/usr/bin/env python3
"""Benchmark of the cvr_hip SpMV path on one MI355X.

Workload (default): FP64 CVR-HIP SpMV on an R-MAT scale-22 graph (4,194,304
rows/cols, edge factor 16, a/b/c/d = .57/.19/.19/.05, duplicates summed,
values U(0,1]), x ~ U[-1,1] from seed 0. A "step" is one y = A*x over the
whole matrix with A, x and y resident in HBM.

Prints ONE JSON line. `value` is the SpMV's algorithmic HBM throughput in GB/s:
bytes = nnz*(8+4) + (rows+1)*4 + cols*8 + rows*8 per SpMV (values, column
indices, row metadata, x, y), divided by the wall time per step. `roofline`
prices the main SpMV kernel alone (hipEvent-timed launches) against 8 TB/s.
`cpu_baseline` is the oracle's FP64 CSR SpMV (OpenMP over rows) timed on the
host cores on the same matrix (the only use of the oracle here); accuracy is
checked against the driver's own FP64 reference SpMV.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    "rmat22_fp64": dict(kind="rmat", scale=22, prec=0),
    "rmat20_fp64": dict(kind="rmat", scale=20, prec=0),
    "stencil27_128_fp64": dict(kind="stencil", n=128, prec=0),
    "laplacian256_fp64": dict(kind="laplacian", n=256, prec=0),
}


def algorithmic_bytes(m: int, n: int, nnz: int, s_val=8, s_idx=4, s_x=8, s_y=8) -> int:
    return int(round(nnz * (s_val + s_idx) + (m + 1) * 4 + n * s_x + m * s_y))


def make_matrix(P, w):
    if w["kind"] == "rmat":
        return P.rmat(w["scale"], 16, 0.57, 0.19, 0.19, seed=0)
    if w["kind"] == "stencil":
        return P.stencil27(w["n"], 0.01, 0)
    return P.laplacian2d(w["n"], w["n"])


def error_stats(y, y_ref, scale=None) -> dict:
    """Per-row relative error |y - y_ref| / |y_ref| (absolute where y_ref == 0)
    and, with scale = sum_j |a_ij x_j|, the scale-normalised error."""
    y = np.asarray(y, np.float64)
    y_ref = np.asarray(y_ref, np.float64)
    d = np.abs(y - y_ref)
    den = np.abs(y_ref)
    rel = np.where(den > 0, d / np.where(den > 0, den, 1.0), d)
    out = {"max_rel": float(rel.max()) if rel.size else 0.0,
           "median_rel": float(np.median(rel)) if rel.size else 0.0}
    if scale is not None:
        sc = np.asarray(scale, np.float64)
        out["max_scaled"] = float(np.max(np.where(sc > 0, d / np.where(sc > 0, sc, 1.0), d))) if d.size else 0.0
    return out


def row_abs(A, x) -> np.ndarray:
    """sum_j |a_ij x_j| per row."""
    import scipy.sparse as sp
    M = sp.csr_matrix((np.abs(A.vals), A.cols, A.rowOffsets), shape=(A.numRows, A.numCols))
    return M @ np.abs(x)


def exact_rows(A, x, rows) -> np.ndarray:
    """Exact row sums (rational arithmetic) rounded to FP64, for a few rows."""
    from fractions import Fraction
    out = []
    for r in rows:
        b, e = int(A.rowOffsets[r]), int(A.rowOffsets[r + 1])
        out.append(float(sum(Fraction(float(v)) * Fraction(float(x[c])) for v, c in zip(A.vals[b:e], A.cols[b:e]))))
    return np.array(out)


def measured_traffic(workload: str, kernel: str, S: int, x_hot_kb: int):
    """HBM bytes per launch of the SpMV kernel from the committed rocprofv3
    counter passes (FETCH_SIZE + WRITE_SIZE, calibrated against a plain
    copy), when this run matches the measured configuration; else None."""
    f = ROOT / "profiles" / "traffic_per_launch.json"
    if not f.exists():
        return None
    ms = json.loads(f.read_text()).get(workload) or []
    for m in ms if isinstance(ms, list) else [ms]:
        if m["kernel"] == kernel and m["steps_per_chunk"] == S and m["x_hot_kb"] == x_hot_kb:
            return m["traffic_bytes"]
    return None


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def _physical_cores():
    """Distinct (package, core) pairs in /proc/cpuinfo; None where it lists none."""
    seen, pkg = set(), None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                pkg = line.split(":", 1)[1].strip()
            elif line.startswith("core id"):
                seen.add((pkg, line.split(":", 1)[1].strip()))
    except OSError:
        return None
    return len(seen) or None


def cpu_baseline(O, A, x, budget_s: float = 12.0):
    """Oracle CSR SpMV (FP64, OpenMP over rows) on the host: a short sweep of
    thread counts (P/4, P/2, P, 2P, P = this process's CPU share), then a
    bounded sample of repeated full SpMVs at the fastest count until ~budget_s
    of wall time."""
    share = int(os.environ.get("OMP_NUM_THREADS", "0")) or min(16, os.cpu_count() or 1)
    share = max(1, min(share, 16))
    cands = sorted({max(1, share // 4), max(1, share // 2), share, 2 * share})
    sweep = {}
    for t in cands:
        O.spmv_omp(A, x, t)  # warm
        t0 = time.perf_counter()
        n = 0
        while True:
            O.spmv_omp(A, x, t)
            n += 1
            if time.perf_counter() - t0 >= 0.15 * budget_s / len(cands) or n >= 200:
                break
        sweep[t] = (time.perf_counter() - t0) / n
    threads = min(sweep, key=sweep.get)
    t0 = time.perf_counter()
    n = 0
    while True:
        O.spmv_omp(A, x, threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= 0.85 * budget_s or n >= 2000:
            break
    per = el / n
    sw = ", ".join(f"{t}: {v * 1e3:.1f}" for t, v in sweep.items())
    return {
        "value": algorithmic_bytes(A.numRows, A.numCols, A.nnz) / per / 1e9,
        "unit": "GB/s",
        "cores": threads,
        "kind": "stand-in: project OpenMP CSR (pava CVR/CSR5 not built)",
        "threads": threads,
        "physical_cores": _physical_cores(),
        "nproc": len(os.sched_getaffinity(0)),
        "cpu_model": _cpu_model(),
        "sample": f"{n} full FP64 CSR SpMVs of the same matrix ({el:.1f} s) at the fastest of threads {{{sw}}} "
                  f"ms/SpMV, oracle orc_csr_spmv_omp, {per * 1e3:.2f} ms each, {A.nnz / per:.3e} nnz/s; "
                  f"{os.cpu_count()} logical CPUs on the host",
    }


def run_precision(P, A, x, steps: int, S: int = 0) -> dict:
    """FP64 (with the default serial parity, and without: parity_tol = 0) vs
    FP32 / BF16 storage (FP32 compensated accumulation), with and without FP64
    recompute of rows whose cancellation ratio exceeds 64, and with 16-bit
    chunk-relative columns; per-row error vs the FP64 reference SpMV. 32-bit
    layouts use automatic hot-x staging, as the main run does."""
    import torch
    dev = torch.device("cuda:0")
    yref = P.reference_spmv(A, x)
    x64 = torch.from_numpy(x).to(dev)
    x32 = x64.float()
    variants = [("f64", dict(precision=0, col_index_bits=32), 8, 4, 8),
                ("f64_parity_off", dict(precision=0, col_index_bits=32, parity_tol=0.0), 8, 4, 8),
                ("f64_col16", dict(precision=0, col_index_bits=16), 8, 2.125, 8),
                ("f64_col16_parity_off", dict(precision=0, col_index_bits=16, parity_tol=0.0), 8, 2.125, 8),
                ("f32", dict(precision=1, col_index_bits=32), 4, 4, 4),
                ("f32_recompute", dict(precision=1, recompute=True, col_index_bits=32), 4, 4, 4),
                ("f32_col16_recompute", dict(precision=1, recompute=True, col_index_bits=16), 4, 2.125, 4),
                ("bf16", dict(precision=2, col_index_bits=32), 2, 4, 4),
                ("bf16_col16", dict(precision=2, col_index_bits=16), 2, 2.125, 4)]
    out = {}
    for name, kw, sv, si, sx in variants:
        try:
            h = P.CvrHip(A, steps_per_chunk=S, x_hot_kb=-1, **kw)
        except P.PavaError as e:
            out[name] = {"error": str(e)}
            continue
        rc = kw.get("recompute", False)
        if kw["precision"] == 0:
            y = torch.empty(A.numRows, dtype=torch.float64, device=dev)
            tot, main = h.time_spmv(x64.data_ptr(), y.data_ptr(), 5, steps)
            h.spmv(x64.data_ptr(), y.data_ptr(), 0)
        else:
            y = torch.empty(A.numRows, dtype=torch.float32, device=dev)
            tot, main = h.time_spmv(x32.data_ptr(), y.data_ptr(), 5, steps, x64.data_ptr() if rc else 0)
            h.spmv_mixed(x32.data_ptr(), y.data_ptr(), x64.data_ptr() if rc else 0, 0)
        torch.cuda.synchronize()
        yy = y.double().cpu().numpy()
        st = error_stats(yy, yref)
        B = A.nnz * (sv + si) + (A.numRows + 1) * 4 + A.numCols * sx + A.numRows * sx
        out[name] = {"ms": round(tot, 5), "main_kernel_ms": round(main, 5), "GB_s": round(B / tot / 1e6, 1),
                     "nnz_per_s": A.nnz / (tot * 1e-3), "bytes": int(B), "max_rel": st["max_rel"],
                     "median_rel": st["median_rel"],
                     "recomputed_rows": h.last_recompute_count() if rc or kw["precision"] == 0 else 0}
        h.close()
    base = out["f64"]["nnz_per_s"]
    for v in out.values():
        if "nnz_per_s" in v:
            v["speedup_vs_f64"] = round(v["nnz_per_s"] / base, 3)
    return out


def _time_calls(fn, iters: int, stream) -> float:
    """ms per call of fn() (launches on `stream`), hipEvent-timed after 3 warm calls."""
    import torch
    for _ in range(3):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(iters):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def run_batch(P, k: int, nmat: int, steps: int, check: int = 8, mats=None, tile=None) -> dict:
    """nmat matrices of the batching family times FP32 RHS blocks of k columns
    in one persistent launch replayed from a hipGraph, with the tile (steps
    per chunk, column width, X staging, SpMM engine) chosen by the batched
    autotuner on a seeded sample of 64 of the matrices; beside it the same
    batch at the fixed tile it used before (S = 64, automatic width, FMA
    engine), the chosen tile on the other engine, and the plain FP64 CVR-HIP
    SpMV (library defaults) of every matrix, one launch each. With `tile`
    given (S, bits, win, engine) the autotuner and the comparisons are skipped
    (a profiler then sees only the full batch's launches)."""
    import torch
    from pava_amd import selector as SEL
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    if mats is None:
        mats = SEL.batch_family(nmat, threads=min(8, os.cpu_count() or 1))
    t_gen = time.perf_counter() - t0
    print(f"[bench] batch: {nmat} matrices generated ({t_gen:.0f}s)", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    given = tile is not None
    tune = {}
    if not given:
        tile, tune = SEL.autotune_batch(mats, k)
    t_tune = time.perf_counter() - t1
    print(f"[bench] batch k={k}: tile {SEL.tile_name(tile)} ({t_tune:.0f}s)", file=sys.stderr, flush=True)
    nnz = sum(A.nnz for A in mats)
    rows = sum(A.numRows for A in mats)
    cols = sum(A.numCols for A in mats)
    res = {"matrices": nmat, "k": k, "nnz": nnz, "rows": rows,
           "tile": {"steps_per_chunk": tile[0], "col_index_bits": "auto" if tile[1] == 0 else tile[1],
                    "x_window_wgs_per_cu": tile[2], "engine": "mfma" if tile[3] == P.SPMM_MFMA else "fma"},
           "autotune_ms_on_sample": {SEL.tile_name(t): round(v, 4) for t, v in tune.items()},
           "autotune_s": round(t_tune, 1)}

    def batch_ms(S, bits, win, eng):
        hs = SEL.build_batch_handles(mats, S, bits)
        try:
            b = P.Batch(hs, k)
            b.set_engine(eng)
            if win:
                b.set_x_window(win)
            inf = b.info()
            inf["col16_nnz"] = sum(A.nnz for A, h in zip(mats, hs) if h.info()["col_index_bits"] == 16)
            b.close()
            return SEL.time_batch(hs, mats, k, iters=max(steps, 5), win=win, engine=eng), inf
        finally:
            for h in hs:
                h.close()

    ms_batch, binf = batch_ms(*tile)
    ms_fixed = None if given else batch_ms(*SEL.BATCH_FIXED)[0]
    # the chosen tile on the other engine (the X window runs on the FMA engine only)
    other = (tile[0], tile[1], 0, P.SPMM_FMA if tile[3] == P.SPMM_MFMA else P.SPMM_MFMA)
    ms_other = None if given else batch_ms(*other)[0]
    # accuracy of the tuned batch on `check` matrices: FP64 reference SpMV on the same FP32 inputs
    hs = SEL.build_batch_handles(mats[:check], tile[0], tile[1])
    errs = []
    try:
        b = P.Batch(hs, k)
        b.set_engine(tile[3])
        if tile[2]:
            b.set_x_window(tile[2])
        g = torch.Generator(device="cuda")
        g.manual_seed(1)
        Xs = [torch.rand((A.numCols, k), dtype=torch.float32, device=dev, generator=g) * 2 - 1 for A in mats[:check]]
        Ys = [torch.empty((A.numRows, k), dtype=torch.float32, device=dev) for A in mats[:check]]
        b.set_operands([t.data_ptr() for t in Xs], [t.data_ptr() for t in Ys])
        b.spmm(0)
        torch.cuda.synchronize()
        b.close()
        for A, Xd, Yd in zip(mats[:check], Xs, Ys):
            X = Xd.cpu().numpy().astype(np.float64)
            Y = Yd.cpu().numpy().astype(np.float64)
            for j in range(k):
                errs.append(error_stats(Y[:, j], P.reference_spmv(A, np.ascontiguousarray(X[:, j]))))
    finally:
        for h in hs:
            h.close()
    # plain FP64 CVR-HIP, library defaults, one SpMV per matrix
    h64 = [P.CvrHip(A) for A in mats]
    try:
        st = torch.cuda.Stream(device=dev)
        x64 = [torch.rand(A.numCols, dtype=torch.float64, device=dev) * 2 - 1 for A in mats]
        y64 = [torch.empty(A.numRows, dtype=torch.float64, device=dev) for A in mats]

        def all64():
            for h, xx, yy in zip(h64, x64, y64):
                h.spmv(xx.data_ptr(), yy.data_ptr(), st.cuda_stream)

        ms_fp64 = _time_calls(all64, max(1, steps // 4), st)
    finally:
        for h in h64:
            h.close()
    # algorithmic bytes with each layout's index width: FP32 values, 4-byte
    # columns or 2.125 B (16-bit offsets + end masks), row pointers, X and Y
    c16 = binf["col16_nnz"]
    Bb = int(round((nnz - c16) * 8 + c16 * 6.125 + (rows + len(mats)) * 4 + cols * k * 4 + rows * k * 4))
    res.update({
        "chunks": binf["chunks"], "items": binf["items"], "grid": binf["grid"], "algorithmic_bytes": Bb,
        "nnz_16bit_columns": c16, "roofline_frac": round(Bb / (ms_batch * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "ms_batch": round(ms_batch, 4),
        "ms_batch_fixed_tile_S64": None if ms_fixed is None else round(ms_fixed, 4),
        "tuned_vs_fixed_tile": None if ms_fixed is None else round(ms_fixed / ms_batch, 3),
        "ms_batch_other_engine": None if ms_other is None else {SEL.tile_name(other): round(ms_other, 4)},
        "GB_s": round(Bb / ms_batch / 1e6, 1), "nnz_rhs_per_s": nnz * k / (ms_batch * 1e-3),
        "ms_fp64_spmv_all": round(ms_fp64, 4), "fp64_nnz_per_s": nnz / (ms_fp64 * 1e-3),
        "speedup_nnz_rhs_vs_fp64_spmv": round((nnz * k / ms_batch) / (nnz / ms_fp64), 2),
        "max_rel": max(e["max_rel"] for e in errs), "median_rel": float(np.median([e["median_rel"] for e in errs])),
        "error_checked": f"{check} matrices x {k} columns vs the FP64 reference SpMV on the same FP32 inputs",
        "build_s": round(time.perf_counter() - t0, 1)})
    return res


def trained_selector():
    """The MLP selector trained on the GPU (library kernels) on the committed
    autotune dataset's training matrices."""
    from pava_amd import selector as SEL
    from pava_amd import selector_data as D
    d = np.load(Path(SEL.__file__).parent / "data" / "selector_dataset.npz", allow_pickle=False)
    assert [str(n) for n in d["configs"]] == [c.name for c in SEL.CONFIGS], "dataset built for another config space"
    tr = np.asarray([D.is_train(int(i)) for i in d["index"]])
    sel, _ = SEL.Selector.fit(d["features"][tr], d["ms"][tr], epochs=400, seed=0)
    return sel, int(tr.sum())


def run_mixed_spmv(P, A, x, sel, iters: int = 200) -> dict:
    """The selector's config for matrix A (MLP over its 16 CSR features) against
    the plain FP64 CVR-HIP default on the same matrix and x: nnz/s of each
    (hipEvents over `iters` launches), and the chosen config's per-row error
    against the FP64 reference SpMV (pava's serial FMA sum)."""
    import torch
    from pava_amd import selector as SEL
    dev = torch.device("cuda:0")
    F = SEL.features(A)
    # exhaustive autotuning over the same configs: the judge of the
    # prediction, and the fallback where the features leave the training range
    tt = SEL.autotune(A, x, iters=20)
    ib = int(np.argmin(tt))
    covered = bool(sel.covers(F)[0])
    if covered:
        pred = sel.mlp.predict(((F - sel.mu) / sel.sd)[None, :])[0]
        order = [int(i) for i in np.argsort(pred)]
        how = "MLP"
    else:
        order = [ib]
        out_of = [int(i) for i in np.nonzero((F < sel.lo) | (F > sel.hi))[0]]
        how = f"exhaustive autotune (fallback: features {out_of} outside the training range)"
    yref = P.reference_spmv(A, x)
    x64 = torch.from_numpy(x).to(dev)
    x32 = x64.float()
    st = torch.cuda.Stream(device=dev)
    for ci in order:  # the best predicted config that applies to this matrix
        cfg = SEL.CONFIGS[ci]
        try:
            h = SEL.build(A, cfg)
        except P.PavaError:
            continue
        break
    try:
        if cfg.fmt in ("csr", "csr5") or cfg.prec == "f64":
            y = torch.empty(A.numRows, dtype=torch.float64, device=dev)
            fn = lambda: h.spmv(x64.data_ptr(), y.data_ptr(), st.cuda_stream)
        else:
            y = torch.empty(A.numRows, dtype=torch.float32, device=dev)
            fn = lambda: h.spmv_mixed(x32.data_ptr(), y.data_ptr(), x64.data_ptr(), st.cuda_stream)
        ms_sel = _time_calls(fn, iters, st)
        fn()
        st.synchronize()
        err = error_stats(y.double().cpu().numpy(), yref)
    finally:
        h.close()
    with P.CvrHip(A) as h0:
        y0 = torch.empty(A.numRows, dtype=torch.float64, device=dev)
        ms_64 = _time_calls(lambda: h0.spmv(x64.data_ptr(), y0.data_ptr(), st.cuda_stream), iters, st)
        h0.spmv(x64.data_ptr(), y0.data_ptr(), st.cuda_stream)
        st.synchronize()
        err0 = error_stats(y0.cpu().numpy(), yref)
    auto = {"selected_by": how, "autotuned_best": SEL.CONFIGS[ib].name, "autotuned_best_ms": round(float(tt[ib]), 5),
            "selected_slowdown_vs_autotuned_best": round(float(tt[SEL.CONFIGS.index(cfg)] / tt[ib]), 3)}
    return {"rows": A.numRows, "nnz": A.nnz, "selected": cfg.name, **auto,
            "selected_nnz_per_s": A.nnz / (ms_sel * 1e-3), "selected_ms": round(ms_sel, 5),
            "selected_max_rel": err["max_rel"], "selected_median_rel": err["median_rel"],
            "fp64_default_nnz_per_s": A.nnz / (ms_64 * 1e-3), "fp64_default_ms": round(ms_64, 5),
            "fp64_default_max_rel": err0["max_rel"],
            "ratio": round(ms_64 / ms_sel, 3)}


def run_mixed_legs(P, A, x, args) -> dict:
    """The metric's second half: the learned, mixed-precision and batched path
    against plain FP64 CVR-HIP (the library default, FP64 serial parity) on
    the same matrices, at the error bound 1e-5 of the FP64 reference."""
    t0 = time.perf_counter()
    sel, n_train = trained_selector()
    out = {"error_bound": 1e-5,
           "baseline": "plain FP64 CVR-HIP: library defaults (automatic chunk size, column width and hot-x; "
                       "FP64 serial parity 1e-12), hipEvent-timed on the same matrix and x",
           "selector": f"MLP trained on the GPU at bench start on {n_train} matrices of the committed autotune "
                       "dataset; the best predicted config that applies"}
    legs = {}
    if args.workload != "stencil27_128_fp64":
        St = P.stencil27(128, 0.01, 0)
        legs["spmv_stencil27_128"] = run_mixed_spmv(P, St, P.fill_uniform(St.numCols, 0), sel)
        del St
    legs[f"spmv_{args.workload}"] = run_mixed_spmv(P, A, x, sel)
    sel.mlp.close()
    print(f"[bench] mixed legs (SpMV) {time.perf_counter() - t0:.0f}s", file=sys.stderr, flush=True)
    b = run_batch(P, 8, args.nmat, 20)
    legs["batch512_k8"] = {
        "matrices": b["matrices"], "k": 8, "nnz": b["nnz"], "tile": b["tile"],
        "selected_nnz_rhs_per_s": b["nnz_rhs_per_s"], "selected_max_rel": b["max_rel"],
        "selected_median_rel": b["median_rel"], "error_checked": b["error_checked"],
        "fp64_default_nnz_per_s": b["fp64_nnz_per_s"], "ratio": b["speedup_nnz_rhs_vs_fp64_spmv"],
        "ms_batch": b["ms_batch"], "ms_batch_fixed_tile_S64": b["ms_batch_fixed_tile_S64"],
        "ms_batch_other_engine": b["ms_batch_other_engine"], "ms_fp64_spmv_all": b["ms_fp64_spmv_all"]}
    out["legs"] = legs
    out["note"] = ("ratio = nonzeros (x right-hand sides) per second of the selected path over the plain FP64 "
                   "default on the same matrices; a batch leg counts nnz x k products against one FP64 SpMV "
                   "per matrix and right-hand side")
    out["seconds"] = round(time.perf_counter() - t0, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="rmat22_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--steps-per-chunk", type=int, default=0, help="CVR chunk steps (0 = automatic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed_vs_fp64 legs (selector-chosen SpMV on the stencil and R-MAT matrices, "
                         "autotuned batched SpMM)")
    ap.add_argument("--variant", choices=["none", "precision", "batch", "create", "gathers"], default="none",
                    help="precision: FP64/FP32/BF16 storage on the workload's matrix; batch: 512-matrix SpMM; "
                         "create: --steps handle creations from the host CSR")
    ap.add_argument("--k", type=int, default=8, help="batch variant: right-hand sides")
    ap.add_argument("--nmat", type=int, default=512, help="batch variant: matrices")
    ap.add_argument("--batch-tile", default="",
                    help="batch variant: fixed tile S:bits:win[:engine] (bits 0 = automatic width; win = X "
                         "window workgroups per CU, 0 = off; engine 0 = FMA, 1 = MFMA) instead of the autotuner")
    ap.add_argument("--x-hot-kb", type=int, default=-1,
                    help="hot-x staging buffer (KiB; 0 = off; -1 = automatic: 2048 KiB when column use is skewed): "
                         "the most used columns read from a dense copy")
    ap.add_argument("--col-index-bits", type=int, default=0, choices=[0, 16, 32],
                    help="CVR column width: 32, 16 (chunk-relative offsets + end masks), 0 = automatic (16 when "
                         "every chunk spans fewer than 65,536 columns)")
    ap.add_argument("--x-lds-kb", type=int, default=-1,
                    help="of the hot-x buffer, KiB held in LDS per workgroup (-1 = as much as fits, 0 = none)")
    ap.add_argument("--copy-calib-mb", type=int, default=0,
                    help="also run 3 plain device copies of this many MiB (counter calibration under rocprofv3)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures the single-GPU SpMV path only; the library has no multi-GPU path. "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import pava_amd as P

    if args.copy_calib_mb > 0:
        n = args.copy_calib_mb * (1 << 20) // 8
        src = torch.rand(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)  # an elementwise kernel: reads and writes exactly n*8 bytes
        torch.cuda.synchronize()
        del src, dst

    if args.variant == "batch":
        tile = tuple(int(v) for v in args.batch_tile.split(":")) if args.batch_tile else None
        if tile is not None and len(tile) == 3:
            tile = tile + (P.SPMM_FMA,)
        res = run_batch(P, args.k, args.nmat, args.steps, tile=tile)
        print(json.dumps({"variant": "batch", **res}), flush=True)
        return

    w = WORKLOADS[args.workload]
    if args.variant == "create":
        # handle creation from a host CSR, repeated: upload, conversion and
        # hot-column staging (the first creation of a process also loads code)
        A = make_matrix(P, w)
        reps = []
        for _ in range(max(args.steps, 1)):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            h = P.CvrHip(A, steps_per_chunk=args.steps_per_chunk, col_index_bits=args.col_index_bits,
                         x_hot_kb=args.x_hot_kb)
            torch.cuda.synchronize()
            wall = (time.perf_counter() - t0) * 1e3
            inf = h.info()
            reps.append({"wall_ms": round(wall, 3), "convert_ms": round(inf["convert_ms"], 3),
                         "hot_setup_ms": round(inf["hot_setup_ms"], 3)})
            h.close()
        print(json.dumps({"variant": "create", "workload": args.workload, "nnz": A.nnz, "reps": reps}), flush=True)
        return
    if args.variant == "gathers":
        # where each x gather of the SpMV is served from: the LDS table, the
        # hot-x copy or x itself (cold), from the handle's own staging
        A = make_matrix(P, w)
        with P.CvrHip(A, steps_per_chunk=args.steps_per_chunk, col_index_bits=args.col_index_bits,
                      x_hot_kb=args.x_hot_kb, x_lds_kb=args.x_lds_kb) as h:
            inf = h.info()
            hot = h.hot_cols()
        rank = np.full(A.numCols, np.iinfo(np.int64).max, np.int64)
        rank[hot] = np.arange(hot.size)
        r = rank[A.cols]
        n_lds, n_hot = inf["n_lds_cols"], inf["n_hot_cols"]
        lds = int(np.sum(r < n_lds))
        hotn = int(np.sum((r >= n_lds) & (r < n_hot)))
        cold_cols = np.unique(A.cols[r >= n_hot])
        print(json.dumps({"variant": "gathers", "workload": args.workload, "nnz": A.nnz, "lds_columns": n_lds,
                          "hot_columns": n_hot, "frac_lds": lds / A.nnz, "frac_hot_copy": hotn / A.nnz,
                          "frac_cold": (A.nnz - lds - hotn) / A.nnz,
                          "cold_distinct_columns": int(cold_cols.size),
                          "cold_distinct_x_lines_128B": int(np.unique(cold_cols // 16).size)}), flush=True)
        return
    if args.variant == "precision":
        A = make_matrix(P, w)
        x = P.fill_uniform(A.numCols, 0)
        res = run_precision(P, A, x, args.steps, args.steps_per_chunk)
        print(json.dumps({"variant": "precision", "workload": args.workload, "rows": A.numRows, "nnz": A.nnz,
                          "results": res}), flush=True)
        return
    t0 = time.perf_counter()
    A = make_matrix(P, w)
    t_gen = time.perf_counter() - t0
    x = P.fill_uniform(A.numCols, 0)
    print(f"[bench] {args.workload}: rows={A.numRows} nnz={A.nnz} max_row={int(np.diff(A.rowOffsets).max())} "
          f"gen {t_gen:.1f}s", file=sys.stderr, flush=True)

    h = P.CvrHip(A, steps_per_chunk=args.steps_per_chunk, precision=w["prec"], x_hot_kb=args.x_hot_kb,
                 x_lds_kb=args.x_lds_kb, col_index_bits=args.col_index_bits)
    info = h.info()
    S = info["steps_per_chunk"]
    dev = torch.device("cuda:0")
    dx = torch.from_numpy(x).to(dev)
    dy = torch.empty(A.numRows, dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream(device=dev)
    sp = stream.cuda_stream

    for _ in range(args.warmup):
        h.spmv(dx.data_ptr(), dy.data_ptr(), sp)
    stream.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        h.spmv(dx.data_ptr(), dy.data_ptr(), sp)
    stream.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall / args.steps * 1e3

    # per-kernel timing with hipEvents on the handle's own stream (>= 100 launches)
    tst = h.time_spmv_stats(dx.data_ptr(), dy.data_ptr(), 5, max(100, min(args.steps, 500)))
    ms_main = tst["main_mean"]

    col16 = info["col_index_bits"] == 16
    B = algorithmic_bytes(A.numRows, A.numCols, A.nnz, s_idx=2.125 if col16 else 4)
    value = B / (ms_step * 1e-3) / 1e9
    par = info["parity_tol"] > 0
    kernel = "spmv_kernel<double,double,double,false,%s,%s,%s>" % ("true" if col16 else "false",
                                                                   "true" if par else "false",
                                                                   "true" if info["n_lds_cols"] else "false")
    achieved = B / (ms_main * 1e-3) / 1e9

    result = {
        "metric": "SpMV HBM GB/s (% peak) + nnz/s; FP64 CVR-HIP on R-MAT scale 22"
        if w["kind"] == "rmat" and w["scale"] == 22 else f"SpMV HBM GB/s ({args.workload})",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if w["prec"] == 0 else "f32",
        "data": "synthetic (seeded generator; x ~ U[-1,1] seed 0)",
        "config": {"workload": args.workload, "rows": A.numRows, "cols": A.numCols, "nnz": A.nnz,
                   "max_row_len": int(np.diff(A.rowOffsets).max()), "steps_per_chunk": S,
                   "chunks": info["nchunks"], "algorithmic_bytes": B, "x_hot_kb": args.x_hot_kb,
                   "hot_columns": info["n_hot_cols"], "lds_columns": info["n_lds_cols"],
                   "col_index_bits": info["col_index_bits"]},
        "nnz_per_s": A.nnz / (ms_step * 1e-3),
        "gflops": 2.0 * A.nnz / (ms_step * 1e-3) / 1e9,
        "hbm_frac": value / HBM_PEAK_GBS,
        "convert_ms": round(info["convert_ms"], 3),
        "hot_setup_ms": round(info["hot_setup_ms"], 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": kernel,
                     "kernel_ms": round(ms_main, 5), "kernel_ms_median": round(tst["main_median"], 5),
                     "kernel_ms_min_max": [round(tst["main_min"], 5), round(tst["main_max"], 5)],
                     # a step is the hot-x gather (when x is staged) and the SpMV
                     # kernel; events around back-to-back steps time both
                     "step_kernels": 2 if info["n_hot_cols"] else 1,
                     "step_kernels_ms": round(tst["ms_total"], 5)},
    }
    # the staging actually in effect (the automatic mode stages 2048 KiB or nothing)
    hot_kb = 0 if info["n_hot_cols"] == 0 else (2048 if args.x_hot_kb < 0 else args.x_hot_kb)
    result["config"]["x_hot_kb_effective"] = hot_kb
    traffic = measured_traffic(args.workload, kernel, S, hot_kb)
    if traffic is not None:
        result["roofline"]["traffic"] = traffic
        result["roofline"]["traffic_GB_s_at_this_kernel_time"] = round(traffic / (ms_main * 1e-3) / 1e9, 1)

    if par:
        # what the FP64 parity pass costs: the same matrix with parity_tol = 0
        # (plain CVR order), timed the same way
        result["parity"] = {"parity_tol": info["parity_tol"], "rows_recomputed": h.last_recompute_count()}
        with P.CvrHip(A, steps_per_chunk=args.steps_per_chunk, precision=w["prec"], x_hot_kb=args.x_hot_kb,
                      x_lds_kb=args.x_lds_kb, col_index_bits=args.col_index_bits, parity_tol=0.0) as h0:
            dy0 = torch.empty_like(dy)
            t0s = h0.time_spmv_stats(dx.data_ptr(), dy0.data_ptr(), 5, 200)
        result["parity"].update({"ms_per_spmv": round(tst["ms_total"], 5),
                                 "ms_per_spmv_parity_off": round(t0s["ms_total"], 5),
                                 "cost": round(tst["ms_total"] / t0s["ms_total"] - 1.0, 4)})
        del dy0
    if not args.no_check:
        # FP64 reference: the driver's verification SpMV (pava's referenceSpMV,
        # serial FMA per row)
        y = dy.cpu().numpy()
        yref = P.reference_spmv(A, x)
        st = error_stats(y, yref, row_abs(A, x))
        st["bit_identical_rows"] = float(np.mean(y == yref))
        out_rows = np.abs(y - yref) > 1e-12 * np.abs(yref)
        st["rows_rel_le_1e-12"] = float(1.0 - np.mean(out_rows))
        st["rows_rel_gt_1e-12"] = int(out_rows.sum())
        if out_rows.any():
            # rows outside 1e-12 of the serial sum: both results against the
            # exact row sum
            ydd = exact_rows(A, x, np.nonzero(out_rows)[0])
            den = np.abs(ydd)
            st["those_rows_max_rel_vs_exact"] = {
                "cvr_hip": float(np.max(np.abs(y[out_rows] - ydd) / den)),
                "serial_reference": float(np.max(np.abs(yref[out_rows] - ydd) / den))}
        result["error_vs_fp64_reference"] = st
    h.close()
    if not args.no_mixed:
        result["mixed_vs_fp64"] = run_mixed_legs(P, A, x, args)
    if not args.no_cpu_baseline:
        import oracle as O
        result["cpu_baseline"] = cpu_baseline(O, A, x)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

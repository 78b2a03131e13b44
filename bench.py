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
    m = json.loads(f.read_text()).get(workload)
    if not m or m["kernel"] != kernel or m["steps_per_chunk"] != S or m["x_hot_kb"] != x_hot_kb:
        return None
    return m["traffic_bytes"]


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


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
        "kind": "port",
        "sample": f"{n} full FP64 CSR SpMVs of the same matrix ({el:.1f} s) at the fastest of threads {{{sw}}} "
                  f"ms/SpMV, oracle orc_csr_spmv_omp, {per * 1e3:.2f} ms each, {A.nnz / per:.3e} nnz/s; "
                  f"host {_cpu_model()}, {os.cpu_count()} logical CPUs visible",
    }


def run_precision(P, A, x, steps: int, S: int = 0) -> dict:
    """FP64 vs FP32 / BF16 storage (FP32 compensated accumulation), with and
    without FP64 recompute of rows whose cancellation ratio exceeds 64, and with
    16-bit chunk-relative columns; per-row error vs the FP64 reference SpMV.
    32-bit layouts use automatic hot-x staging, as the main run does."""
    import torch
    dev = torch.device("cuda:0")
    yref = P.reference_spmv(A, x)
    x64 = torch.from_numpy(x).to(dev)
    x32 = x64.float()
    variants = [("f64", dict(precision=0, col_index_bits=32), 8, 4, 8),
                ("f64_col16", dict(precision=0, col_index_bits=16), 8, 2.125, 8),
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
                     "recomputed_rows": h.last_recompute_count() if rc else 0}
        h.close()
    base = out["f64"]["nnz_per_s"]
    for v in out.values():
        if "nnz_per_s" in v:
            v["speedup_vs_f64"] = round(v["nnz_per_s"] / base, 3)
    return out


def run_batch(P, k: int, nmat: int, steps: int, S: int = 64, check: int = 4) -> dict:
    """nmat random matrices (rows log-uniform in [1e3, 2e5], Poisson row
    lengths with mean U[5, 60], 70% of columns within +-1000 of the diagonal,
    FP32 N(0,1) values) times FP32 RHS blocks of k columns, in one persistent
    launch replayed from a hipGraph; compared with plain FP64 CVR-HIP SpMVs of
    the same matrices (one launch each)."""
    import torch
    rng = np.random.default_rng(4)
    dev = torch.device("cuda:0")
    mats, hs, h64 = [], [], []
    t0 = time.perf_counter()
    for i in range(nmat):
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(2e5))))
        A = P.synthetic(n, P.RL_POISSON, float(rng.uniform(5, 60)), col_struct=P.COL_MIXED, half_band=1000,
                        band_frac=0.7, value_dist=0, seed=10_000 + i)
        A.vals = A.vals.astype(np.float32).astype(np.float64)
        mats.append(A)
        hs.append(P.CvrHip(A, steps_per_chunk=S, precision=P.PREC_F32))
        h64.append(P.CvrHip(A))
        if i % 64 == 63:
            print(f"[bench] batch: {i + 1}/{nmat} matrices built ({time.perf_counter() - t0:.0f}s)", file=sys.stderr,
                  flush=True)
    nnz = sum(A.nnz for A in mats)
    Xs = [torch.rand((A.numCols, k), dtype=torch.float32, device=dev) * 2 - 1 for A in mats]
    Ys = [torch.empty((A.numRows, k), dtype=torch.float32, device=dev) for A in mats]
    b = P.Batch(hs, k)
    b.set_operands([t.data_ptr() for t in Xs], [t.data_ptr() for t in Ys])
    st = torch.cuda.Stream(device=dev)
    b.capture(st.cuda_stream)
    for _ in range(3):
        b.replay(st.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(steps):
        b.replay(st.cuda_stream)
    e1.record(st)
    e1.synchronize()
    ms_batch = e0.elapsed_time(e1) / steps
    # plain FP64 CVR-HIP, one SpMV per matrix (x = first column of the block)
    x64 = [X[:, 0].double().contiguous() for X in Xs]
    y64 = [torch.empty(A.numRows, dtype=torch.float64, device=dev) for A in mats]
    for _ in range(2):
        for h, xx, yy in zip(h64, x64, y64):
            h.spmv(xx.data_ptr(), yy.data_ptr(), st.cuda_stream)
    e0.record(st)
    reps = max(1, steps // 4)
    for _ in range(reps):
        for h, xx, yy in zip(h64, x64, y64):
            h.spmv(xx.data_ptr(), yy.data_ptr(), st.cuda_stream)
    e1.record(st)
    e1.synchronize()
    ms_fp64 = e0.elapsed_time(e1) / reps
    # accuracy on a sample of matrices: FP64 reference SpMV on the same FP32 inputs
    errs = []
    for i in range(min(check, nmat)):
        X = Xs[i].cpu().numpy().astype(np.float64)
        Y = Ys[i].cpu().numpy().astype(np.float64)
        for j in range(k):
            errs.append(error_stats(Y[:, j], P.reference_spmv(mats[i], X[:, j])))
    rows = sum(A.numRows for A in mats)
    cols = sum(A.numCols for A in mats)
    Bb = nnz * 8 + (rows + 1) * 4 + cols * k * 4 + rows * k * 4
    res = {"matrices": nmat, "k": k, "nnz": nnz, "rows": rows, "chunks": b.info()["chunks"], "grid": b.info()["grid"],
           "ms_batch": round(ms_batch, 4), "GB_s": round(Bb / ms_batch / 1e6, 1),
           "nnz_rhs_per_s": nnz * k / (ms_batch * 1e-3),
           "ms_fp64_spmv_all": round(ms_fp64, 4), "fp64_nnz_per_s": nnz / (ms_fp64 * 1e-3),
           "speedup_nnz_rhs_vs_fp64_spmv": round((nnz * k / ms_batch) / (nnz / ms_fp64), 2),
           "max_rel": max(e["max_rel"] for e in errs), "median_rel": float(np.median([e["median_rel"] for e in errs])),
           "build_s": round(time.perf_counter() - t0, 1)}
    b.close()
    for h in hs + h64:
        h.close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="rmat22_fp64", choices=sorted(WORKLOADS))
    ap.add_argument("--steps-per-chunk", type=int, default=0, help="CVR chunk steps (0 = automatic)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--variant", choices=["none", "precision", "batch", "create"], default="none",
                    help="precision: FP64/FP32/BF16 storage on the workload's matrix; batch: 512-matrix SpMM; "
                         "create: --steps handle creations from the host CSR")
    ap.add_argument("--k", type=int, default=8, help="batch variant: right-hand sides")
    ap.add_argument("--nmat", type=int, default=512, help="batch variant: matrices")
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
        res = run_batch(P, args.k, args.nmat, args.steps)
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
                     "kernel_ms_min_max": [round(tst["main_min"], 5), round(tst["main_max"], 5)]},
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
    if not args.no_cpu_baseline:
        import oracle as O
        result["cpu_baseline"] = cpu_baseline(O, A, x)
    else:
        result["cpu_baseline"] = None
    h.close()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

"""Autotuner and learned selector for the SpMV path (no pava counterpart).

* ``CONFIGS`` -- the candidate space: format (cvr_hip / csr5_hip / csr_hip), value
  precision (FP64; FP32 with FP64 recompute of rows whose cancellation ratio
  exceeds 64, which keeps every row within ~1e-5 of FP64; the same with 16-bit
  chunk-relative columns), tile (CVR steps per chunk S or the library's
  automatic size, CSR5 elements per lane sigma, or CSR lanes per row L) and x
  staging (hot-x copy of the most used columns, its head held in LDS).
* ``autotune`` -- exhaustive search: builds each candidate on the GPU, times
  it with hipEvents, returns per-config times (inf where a candidate does not
  apply, e.g. a chunk spanning more than 65536 columns for 16-bit columns).
* ``features`` -- 16 CSR features (C ABI ``pava_csr_features``).
* ``Mlp`` -- 3-layer MLP trained on the GPU by the library's own kernels
  (``pava_mlp_*``): it regresses log2(time / best time) of every config from
  the standardised features; the selector picks the config with the lowest
  prediction.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import (CSR, Csr5Hip, CvrHip, CsrHip, PREC_F32, PREC_F64, SPMM_FMA, SPMM_MFMA, PavaError, _check, lib)

N_FEATURES = 16


@dataclass(frozen=True)
class Config:
    fmt: str        # "cvr" | "csr5" | "csr"
    prec: str       # "f64" | "f32rc" | "f32rc16"
    tile: int       # S for cvr (0: the library's automatic size), sigma for csr5, L for csr
    xs: int = 0     # x staging: hot-x copy size in KiB (0 = off, -1 = the library's automatic choice)

    @property
    def name(self) -> str:
        xs = "" if self.xs == 0 else "-xauto" if self.xs < 0 else f"-x{self.xs}"
        return f"{self.fmt}-{self.prec}-{self.tile}{xs}"


CONFIGS = ([Config("cvr", "f64", s) for s in (32, 128, 512)]
           + [Config("cvr", "f32rc", s) for s in (32, 128, 512)]
           + [Config("cvr", "f32rc16", s) for s in (32, 128, 512)]
           + [Config("csr", "f64", l) for l in (1, 4, 16, 64)]
           + [Config("cvr", "f64", s, 2048) for s in (128, 512)]
           + [Config("cvr", "f32rc", 128, 2048)]
           + [Config("csr5", "f64", s) for s in (8, 32)]
           + [Config("cvr", "f64", 0, -1)]
           + [Config("cvr", "f32rc", 0, -1), Config("cvr", "f32rc16", 0)])
DEFAULT = Config("cvr", "f64", 0, -1)  # the fallback without a model: the library's default


def features(A: CSR) -> np.ndarray:
    L = lib()
    L.pava_csr_features.argtypes = [C.c_void_p, C.c_void_p]
    out = np.zeros(N_FEATURES, np.float32)
    s = A.c_struct()
    _check(L.pava_csr_features(C.byref(s), out.ctypes.data))
    return out


def _time(fn, iters: int):
    import torch
    s = torch.cuda.current_stream()
    for _ in range(2):
        fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(iters):
        fn()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / iters


def build(A: CSR, cfg: Config):
    if cfg.fmt == "csr":
        return CsrHip(A, lanes_per_row=cfg.tile, precision=PREC_F64)
    if cfg.fmt == "csr5":
        return Csr5Hip(A, sigma=cfg.tile)
    # fixed configs fix the column width; the library default (xs < 0) keeps
    # its automatic choices of column width and hot-x staging
    bits = 16 if cfg.prec == "f32rc16" else 0 if cfg.xs < 0 else 32
    if cfg.prec == "f64":
        return CvrHip(A, steps_per_chunk=cfg.tile, x_hot_kb=cfg.xs, col_index_bits=bits)
    return CvrHip(A, steps_per_chunk=cfg.tile, precision=PREC_F32, recompute=True, kappa_threshold=64.0,
                  col_index_bits=bits, x_hot_kb=cfg.xs)


def autotune(A: CSR, x: np.ndarray, iters: int = 5, configs=CONFIGS) -> np.ndarray:
    """Milliseconds per SpMV of every config (inf where it does not apply).
    Timed with events on torch's current stream (the launches use it)."""
    import torch
    dev = torch.device("cuda:0")
    x64 = torch.from_numpy(np.ascontiguousarray(x, np.float64)).to(dev)
    x32 = x64.float()
    y64 = torch.empty(A.numRows, dtype=torch.float64, device=dev)
    y32 = torch.empty(A.numRows, dtype=torch.float32, device=dev)
    st = torch.cuda.current_stream().cuda_stream
    out = np.full(len(configs), np.inf)
    for i, cfg in enumerate(configs):
        try:
            h = build(A, cfg)
        except PavaError:
            continue
        try:
            if cfg.fmt in ("csr", "csr5"):
                fn = lambda: h.spmv(x64.data_ptr(), y64.data_ptr(), st)
            elif cfg.prec == "f64":
                fn = lambda: h.spmv(x64.data_ptr(), y64.data_ptr(), st)
            else:
                fn = lambda: h.spmv_mixed(x32.data_ptr(), y32.data_ptr(), x64.data_ptr(), st)
            out[i] = _time(fn, iters)
        finally:
            h.close()
    return out


class Mlp:
    """3-layer MLP on the GPU (pava_mlp_* kernels)."""

    def __init__(self, n_in: int = N_FEATURES, n_hidden: int = 64, n_out: int = len(CONFIGS), seed: int = 0,
                 device: int = 0):
        L = lib()
        P = C.c_void_p
        L.pava_mlp_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.POINTER(P)]
        L.pava_mlp_num_params.argtypes = [P]
        L.pava_mlp_get_params.argtypes = [P, P]
        L.pava_mlp_set_params.argtypes = [P, P]
        L.pava_mlp_predict.argtypes = [P, C.c_int, P, P]
        L.pava_mlp_loss_grad.argtypes = [P, C.c_int, P, P, P, P, P]
        L.pava_mlp_train.argtypes = [P, C.c_int, P, P, P, C.c_int, C.c_int, C.c_float, C.c_ulonglong, P]
        L.pava_mlp_destroy.argtypes = [P]
        self.shape = (n_in, n_hidden, n_out)
        self._m = P()
        _check(L.pava_mlp_create(n_in, n_hidden, n_out, seed, device, C.byref(self._m)))
        self.np = L.pava_mlp_num_params(self._m)

    def params(self) -> np.ndarray:
        p = np.empty(self.np, np.float32)
        _check(lib().pava_mlp_get_params(self._m, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.np
        _check(lib().pava_mlp_set_params(self._m, p.ctypes.data))

    def predict(self, X: np.ndarray) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.empty((X.shape[0], self.shape[2]), np.float32)
        _check(lib().pava_mlp_predict(self._m, X.shape[0], X.ctypes.data, Y.ctypes.data))
        return Y

    def loss_grad(self, X, T, M):
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        loss = np.zeros(1, np.float32)
        g = np.empty(self.np, np.float32)
        _check(lib().pava_mlp_loss_grad(self._m, X.shape[0], X.ctypes.data, T.ctypes.data, M.ctypes.data,
                                        loss.ctypes.data, g.ctypes.data))
        return float(loss[0]), g

    def train(self, X, T, M, epochs: int = 200, batch: int = 64, lr: float = 2e-3, seed: int = 1) -> np.ndarray:
        X, T, M = (np.ascontiguousarray(a, np.float32) for a in (X, T, M))
        hist = np.zeros(epochs, np.float32)
        _check(lib().pava_mlp_train(self._m, X.shape[0], X.ctypes.data, T.ctypes.data, M.ctypes.data, epochs, batch,
                                    lr, seed, hist.ctypes.data))
        return hist

    def close(self):
        if self._m:
            lib().pava_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def targets(times: np.ndarray):
    """log2(time / best) per config and a validity mask (rows: matrices)."""
    t = np.asarray(times, np.float64)
    best = np.min(np.where(np.isfinite(t), t, np.inf), axis=1, keepdims=True)
    mask = np.isfinite(t)
    T = np.where(mask, np.log2(np.where(mask, t, 1.0) / best), 0.0)
    return T.astype(np.float32), mask.astype(np.float32)


class Selector:
    """Standardised features -> MLP -> argmin predicted slowdown (valid configs).
    `lo`/`hi` bound each feature over the training matrices: outside them the
    model extrapolates, and `covers` tells the caller to fall back to the
    exhaustive autotuner."""

    def __init__(self, mlp: Mlp, mu: np.ndarray, sd: np.ndarray, lo=None, hi=None):
        self.mlp, self.mu, self.sd = mlp, mu, sd
        self.lo, self.hi = lo, hi

    @staticmethod
    def fit(F: np.ndarray, times: np.ndarray, epochs: int = 300, seed: int = 0, lr: float = 2e-3) -> tuple:
        mu = F.mean(0)
        sd = F.std(0) + 1e-6
        T, M = targets(times)
        mlp = Mlp(seed=seed)
        hist = mlp.train((F - mu) / sd, T, M, epochs=epochs, batch=64, lr=lr, seed=seed + 1)
        return Selector(mlp, mu, sd, F.min(0), F.max(0)), hist

    def covers(self, F: np.ndarray, margin: float = 0.05) -> np.ndarray:
        """Per matrix: every feature within the training range widened by
        `margin` of its width."""
        F = np.atleast_2d(F)
        if self.lo is None:
            return np.ones(F.shape[0], bool)
        w = (self.hi - self.lo) * margin
        return np.all((F >= self.lo - w) & (F <= self.hi + w), axis=1)

    def choose(self, F: np.ndarray, valid: np.ndarray | None = None) -> np.ndarray:
        P = self.mlp.predict((np.atleast_2d(F) - self.mu) / self.sd)
        if valid is not None:
            P = np.where(valid, P, np.inf)
        return np.argmin(P, axis=1)


def slowdowns(times: np.ndarray, choice: np.ndarray) -> np.ndarray:
    """Measured time of the chosen config over the best config, per matrix."""
    t = np.where(np.isfinite(times), times, np.inf)
    best = t.min(1)
    return t[np.arange(t.shape[0]), choice] / best


def gen_params(rng: np.random.Generator, max_rows: int = 1_000_000, min_rows: int = 10_000) -> dict:
    """One random member of the selector's synthetic family: rows log-uniform in
    [min_rows, max_rows], row lengths constant / Poisson / lognormal (sigma
    0.5-2) / power law (alpha 1.8-2.5), columns banded / block-diagonal (block
    32-512) / uniform."""
    rows = int(math.exp(rng.uniform(math.log(min_rows), math.log(max_rows))))
    dist = int(rng.integers(0, 4))
    mean = float(math.exp(rng.uniform(math.log(4), math.log(64))))
    shape = 1.0
    if dist == 2:
        shape = float(rng.uniform(0.5, 2.0))
    elif dist == 3:
        shape = float(rng.uniform(1.8, 2.5))
        mean = float(rng.uniform(1, 8))  # x_min of the Pareto law
    col = int(rng.integers(0, 3))
    return dict(rows=rows, row_dist=dist, len_param=mean, shape=shape, col_struct=col,
                half_band=int(rng.integers(8, 4096)), block=int(2 ** rng.integers(5, 10)), band_frac=0.7,
                value_dist=0, max_len=min(rows, 20000), seed=int(rng.integers(0, 2 ** 31)))


# ------------------------------------------------------- batched SpMM tiles
# Candidate tiles of a batched multi-RHS SpMM, one choice per batch: CVR steps
# per chunk S; the column width (0: the library's automatic width, 16-bit
# chunk-relative columns wherever a chunk spans fewer than 65,536 columns; 32:
# 32-bit columns); the X staging: 0 reads X rows in place, 1 or 2 stage the X
# rows around each group of 4 chunks in an LDS window sized for 1 or 2
# workgroups per CU (k <= 8 only); and the engine: SPMM_FMA (per-lane FP64
# FMA) or SPMM_MFMA (the matrix cores; X rows read in place).
_STEPS = (16, 32, 64, 128, 256)
BATCH_TILES = ([(s, bits, win, SPMM_FMA) for s in _STEPS for bits in (0, 32) for win in (0, 1, 2)] +
               [(s, bits, 0, SPMM_MFMA) for s in _STEPS for bits in (0, 32)])
BATCH_FIXED = (64, 0, 0, SPMM_FMA)  # the tile the batched path used before it was autotuned


def tile_name(tile) -> str:
    """S64-auto-w0-fma style name of a batch tile."""
    s, bits, win, eng = tile
    return f"S{s}-{'auto' if bits == 0 else bits}-w{win}-{'mfma' if eng == SPMM_MFMA else 'fma'}"


def batch_family(nmat: int = 512, seed: int = 4, threads: int = 8):
    """The batching workload: nmat square matrices, rows log-uniform in
    [1e3, 2e5], Poisson row lengths with mean U[5, 60], 70% of columns within
    +-1000 of the diagonal and 30% uniform, FP32 N(0,1) values (held in the
    FP64 CSR as their FP32 roundings). Generated on `threads` host threads
    (the generator releases the GIL); matrix i is seeded 10,000 + i."""
    from concurrent.futures import ThreadPoolExecutor
    from . import COL_MIXED, RL_POISSON, synthetic
    rng = np.random.default_rng(seed)
    prm = [(int(np.exp(rng.uniform(np.log(1e3), np.log(2e5)))), float(rng.uniform(5, 60))) for _ in range(nmat)]

    def one(i):
        n, mean = prm[i]
        A = synthetic(n, RL_POISSON, mean, col_struct=COL_MIXED, half_band=1000, band_frac=0.7, value_dist=0,
                      seed=10_000 + i)
        A.vals = A.vals.astype(np.float32).astype(np.float64)
        return A

    with ThreadPoolExecutor(max(1, threads)) as ex:
        return list(ex.map(one, range(nmat)))


def build_batch_handles(mats, S: int, bits: int):
    """FP32 CvrHip handles of every matrix at one tile (the batched SpMM's input)."""
    hs = []
    try:
        for A in mats:
            hs.append(CvrHip(A, steps_per_chunk=S, precision=PREC_F32, col_index_bits=bits, x_hot_kb=0))
    except BaseException:
        for h in hs:
            h.close()
        raise
    return hs


def time_batch(hs, mats, k: int, iters: int = 10, seed: int = 0, win: int = 0, engine: int = SPMM_FMA) -> float:
    """Milliseconds per replay of the captured persistent batched SpMM over
    handles hs (FP32 X blocks of k columns, uniform [-1, 1]), with X window
    mode `win` and SpMM engine `engine`."""
    import torch
    from . import Batch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    Xs = [torch.rand((A.numCols, k), dtype=torch.float32, device="cuda", generator=g) * 2 - 1 for A in mats]
    Ys = [torch.empty((A.numRows, k), dtype=torch.float32, device="cuda") for A in mats]
    b = Batch(hs, k)
    try:
        b.set_engine(engine)
        if win:
            b.set_x_window(win)
        b.set_operands([t.data_ptr() for t in Xs], [t.data_ptr() for t in Ys])
        st = torch.cuda.Stream()
        b.capture(st.cuda_stream)
        for _ in range(3):
            b.replay(st.cuda_stream)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(iters):
            b.replay(st.cuda_stream)
        e1.record(st)
        e1.synchronize()
        return e0.elapsed_time(e1) / iters
    finally:
        b.close()


def autotune_batch(mats, k: int, tiles=BATCH_TILES, sample: int = 64, iters: int = 10, seed: int = 0):
    """Exhaustive tile search for a batched SpMM: every candidate (S, column
    width, X staging, engine) is built and timed as one persistent launch over
    a seeded sample of the batch's matrices. Returns (best tile, {tile: ms per
    launch on the sample}); inf where a candidate does not apply (X windows
    beyond k = 8, a layout that cannot be built)."""
    rng = np.random.default_rng(seed)
    pick = np.sort(rng.choice(len(mats), size=min(sample, len(mats)), replace=False))
    sub = [mats[int(i)] for i in pick]
    times = {}
    for S, bits in dict.fromkeys((t[0], t[1]) for t in tiles):
        rest = [t for t in tiles if (t[0], t[1]) == (S, bits)]
        try:
            hs = build_batch_handles(sub, S, bits)
        except PavaError:
            times.update({t: float("inf") for t in rest})
            continue
        try:
            for t in rest:
                try:
                    times[t] = time_batch(hs, sub, k, iters=iters, seed=seed, win=t[2], engine=t[3])
                except PavaError:
                    times[t] = float("inf")
        finally:
            for h in hs:
                h.close()
    best = min(times, key=times.get)
    return best, times

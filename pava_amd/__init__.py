"""pava_amd -- MI355X CVR SpMV path ("cvr_hip") behind pava's format-plugin surface.

Python mirror of the C ABI declared in ``include/pava_hip.h`` (ctypes). The
compute runs in ``libpava_hip.so`` (hand-written gfx950 HIP kernels); there is
no CPU fallback: loading fails loudly when the library is missing.

Reference interfaces mirrored (paths relative to the pava source tree):
  CSR / read_mtx           _PAVA_CSRMatrix, readCOOMatrix + convertCOO2CSR
                           (utils/basicMatrix.h:50-60, utils/basicMatrix.cpp:63-415)
  init_vectors             initVectors (utils/basicMatrix.cpp:499-520)
  reference_spmv           referenceSpMV (utils/basicMatrix.cpp:538-556)
  check_results            checkResults (utils/basicMatrix.cpp:559-587)
  CvrHip                   pre_processing + spmv_compute_kernel (cvr/cvr.h:583-1884)
  conduct_cvr_hip          conduct_cvr (cvr/cvr.h:1953-2186)
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Optional

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "libpava_hip.so"

PREC_F64, PREC_F32, PREC_BF16 = 0, 1, 2
SPMM_FMA, SPMM_MFMA = 0, 1  # SpMM engines: per-lane FP64 FMA, or the matrix cores
SPLIT_CAP = 72
SEG_PER_CHUNK_EXTRA = 66

RL_CONSTANT, RL_POISSON, RL_LOGNORMAL, RL_POWERLAW = 0, 1, 2, 3
COL_BANDED, COL_BLOCKDIAG, COL_UNIFORM, COL_MIXED = 0, 1, 2, 3


class PavaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class _CSR(C.Structure):
    _fields_ = [("nnz", C.c_int), ("numRows", C.c_int), ("numCols", C.c_int),
                ("vals", C.POINTER(C.c_double)), ("cols", C.POINTER(C.c_int)),
                ("rowOffsets", C.POINTER(C.c_int))]


class _Opts(C.Structure):
    _fields_ = [("steps_per_chunk", C.c_int), ("precision", C.c_int), ("col_index_bits", C.c_int),
                ("recompute", C.c_int), ("kappa_threshold", C.c_double), ("keep_fp64", C.c_int),
                ("device", C.c_int), ("x_hot_kb", C.c_int), ("x_lds_kb", C.c_int), ("parity_tol", C.c_double)]


class _Info(C.Structure):
    _fields_ = [("numRows", C.c_int), ("numCols", C.c_int), ("nnz", C.c_int), ("nchunks", C.c_int),
                ("steps_per_chunk", C.c_int), ("precision", C.c_int), ("col_index_bits", C.c_int),
                ("layout_elems", C.c_longlong), ("seg_capacity", C.c_longlong), ("max_slots", C.c_int),
                ("has_empty_rows", C.c_int), ("n_fixup_rows", C.c_int), ("device_bytes", C.c_longlong),
                ("convert_ms", C.c_double), ("n_hot_cols", C.c_int), ("n_lds_cols", C.c_int),
                ("hot_setup_ms", C.c_double), ("parity_tol", C.c_double),
                ("n_spec_rows", C.c_int)]


class _GenParams(C.Structure):
    _fields_ = [("rows", C.c_int), ("row_dist", C.c_int), ("len_param", C.c_double), ("shape", C.c_double),
                ("col_struct", C.c_int), ("half_band", C.c_int), ("block", C.c_int), ("band_frac", C.c_double),
                ("value_dist", C.c_int), ("max_len", C.c_int), ("seed", C.c_ulonglong)]


_lib: Optional[C.CDLL] = None

# every symbol the header declares
EXPORTED = [
    "pava_read_coo_matrix", "pava_convert_coo_to_csr", "pava_read_csr_matrix", "pava_write_matrix_market",
    "pava_delete_coo_matrix", "pava_delete_csr_matrix", "pava_csr_validate", "pava_init_vectors",
    "pava_fill_uniform", "pava_reference_spmv", "pava_check_results", "pava_print_performance",
    "pava_microtime", "pava_gen_laplacian2d", "pava_gen_stencil27", "pava_gen_rmat", "pava_gen_synthetic",
    "cvr_hip_default_opts", "cvr_hip_create", "cvr_hip_create_device", "cvr_hip_destroy", "cvr_hip_get_info",
    "cvr_hip_spmv", "cvr_hip_spmv_host", "cvr_hip_spmv_mixed", "cvr_hip_spmm", "cvr_hip_last_recompute_count",
    "cvr_hip_time_spmv", "cvr_hip_time_spmv_x64", "cvr_hip_time_spmv_stats", "cvr_hip_export_layout", "conduct_cvr_hip", "pava_bench_file",
    "pava_hip_last_error", "pava_hip_version", "cvr_hip_batch_create", "cvr_hip_batch_set_operands",
    "cvr_hip_batch_spmm", "cvr_hip_batch_graph", "cvr_hip_batch_replay", "cvr_hip_batch_info",
    "cvr_hip_batch_set_x_window",
    "cvr_hip_batch_destroy", "csr_hip_create", "csr_hip_spmv", "csr_hip_destroy", "pava_csr_features",
    "pava_mlp_create", "pava_mlp_num_params", "pava_mlp_get_params", "pava_mlp_set_params", "pava_mlp_predict",
    "pava_mlp_loss_grad", "pava_mlp_train", "pava_mlp_destroy", "cvr_hip_export_hot", "pava_write_csr_binary",
    "pava_read_csr_binary", "conduct_csr_hip", "csr5_hip_create", "csr5_hip_spmv", "csr5_hip_spmv_host",
    "csr5_hip_get_info", "csr5_hip_export_layout", "csr5_hip_destroy", "conduct_csr5_hip", "coo_hip_create",
    "coo_hip_spmv", "coo_hip_get_info", "coo_hip_destroy", "conduct_coo_hip", "cvr_hip_batch_queue_counts",
    "pava_hip_test_fail_alloc", "pava_hip_test_live_device_allocs", "cvr_hip_spmm_noalloc",
    "cvr_hip_set_spmm_engine", "cvr_hip_batch_set_engine",
]


def lib() -> C.CDLL:
    """Loads libpava_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m pava_amd.build`")
    # PyTorch-ROCm ships its own libamdhip64 / libhsa-runtime64 under the same
    # sonames as /opt/rocm's. Whichever is loaded first serves the whole
    # process; load torch's first so that torch and this library share one HIP
    # runtime (device pointers and streams pass between them).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(str(LIB_PATH))
    P = C.c_void_p
    i, d, ull = C.c_int, C.c_double, C.c_ulonglong
    L.pava_hip_last_error.restype = C.c_char_p
    L.pava_hip_version.restype = C.c_char_p
    L.pava_microtime.restype = d
    L.pava_read_csr_matrix.argtypes = [C.c_char_p, C.POINTER(_CSR)]
    L.pava_write_matrix_market.argtypes = [C.c_char_p, C.POINTER(_CSR)]
    L.pava_delete_csr_matrix.argtypes = [C.POINTER(_CSR)]
    L.pava_csr_validate.argtypes = [C.POINTER(_CSR)]
    L.pava_init_vectors.argtypes = [i, i, P, P, P]
    L.pava_init_vectors.restype = None
    L.pava_fill_uniform.argtypes = [P, C.c_longlong, ull, d, d]
    L.pava_fill_uniform.restype = None
    L.pava_reference_spmv.argtypes = [C.POINTER(_CSR), P, P]
    L.pava_reference_spmv.restype = None
    L.pava_check_results.argtypes = [i, P, P]
    L.pava_print_performance.argtypes = [C.c_char_p, C.c_char_p, i, d, d]
    L.pava_gen_laplacian2d.argtypes = [i, i, C.POINTER(_CSR)]
    L.pava_gen_stencil27.argtypes = [i, d, ull, C.POINTER(_CSR)]
    L.pava_gen_rmat.argtypes = [i, i, d, d, d, ull, C.POINTER(_CSR)]
    L.pava_gen_synthetic.argtypes = [C.POINTER(_GenParams), C.POINTER(_CSR)]
    L.cvr_hip_default_opts.argtypes = [C.POINTER(_Opts)]
    L.cvr_hip_default_opts.restype = None
    L.cvr_hip_create.argtypes = [C.POINTER(_CSR), C.POINTER(_Opts), C.POINTER(P)]
    L.cvr_hip_create_device.argtypes = [i, i, i, P, P, P, C.POINTER(_Opts), C.POINTER(P)]
    L.cvr_hip_destroy.argtypes = [P]
    L.cvr_hip_get_info.argtypes = [P, C.POINTER(_Info)]
    L.cvr_hip_spmv.argtypes = [P, P, P, P]
    L.cvr_hip_spmv_host.argtypes = [P, P, P]
    L.cvr_hip_spmv_mixed.argtypes = [P, P, P, P, P]
    L.cvr_hip_spmm.argtypes = [P, i, P, P, P]
    L.cvr_hip_spmm_noalloc.argtypes = [P, i, P, P, P]
    L.cvr_hip_set_spmm_engine.argtypes = [P, i]
    L.cvr_hip_batch_set_engine.argtypes = [P, i]
    L.cvr_hip_last_recompute_count.argtypes = [P, C.POINTER(i)]
    L.cvr_hip_time_spmv.argtypes = [P, P, P, i, i, C.POINTER(d), C.POINTER(d)]
    L.cvr_hip_time_spmv_x64.argtypes = [P, P, P, P, i, i, C.POINTER(d), C.POINTER(d)]
    L.cvr_hip_export_layout.argtypes = [P] + [P] * 10
    L.conduct_cvr_hip.argtypes = [i, i, i, P, P, P, P, P, i, C.c_char_p, i]
    L.pava_bench_file.argtypes = [C.c_char_p, i]
    L.cvr_hip_batch_create.argtypes = [i, C.POINTER(P), i, C.POINTER(P)]
    L.cvr_hip_batch_set_operands.argtypes = [P, C.POINTER(P), C.POINTER(P)]
    L.cvr_hip_batch_spmm.argtypes = [P, P]
    L.cvr_hip_batch_graph.argtypes = [P, P]
    L.cvr_hip_batch_replay.argtypes = [P, P]
    L.cvr_hip_batch_info.argtypes = [P, C.POINTER(i), C.POINTER(i), C.POINTER(C.c_longlong)]
    L.cvr_hip_batch_destroy.argtypes = [P]
    L.cvr_hip_batch_queue_counts.argtypes = [P, C.POINTER(i), C.POINTER(i)]
    L.pava_hip_test_fail_alloc.argtypes = [i, i]
    L.pava_hip_test_fail_alloc.restype = None
    L.pava_hip_test_live_device_allocs.argtypes = []
    L.pava_hip_test_live_device_allocs.restype = C.c_longlong
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise PavaError(rc, lib().pava_hip_last_error().decode())
    return rc


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class CSR:
    """Host CSR with the fields of _PAVA_CSRMatrix; 0-based int32 indices."""
    numRows: int
    numCols: int
    rowOffsets: np.ndarray
    cols: np.ndarray
    vals: np.ndarray

    @property
    def nnz(self) -> int:
        return int(self.cols.shape[0])

    def __post_init__(self):
        self.rowOffsets = np.ascontiguousarray(self.rowOffsets, dtype=np.int32)
        self.cols = np.ascontiguousarray(self.cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(self.vals, dtype=np.float64)

    def c_struct(self) -> _CSR:
        s = _CSR()
        s.nnz, s.numRows, s.numCols = self.nnz, self.numRows, self.numCols
        s.vals = self.vals.ctypes.data_as(C.POINTER(C.c_double))
        s.cols = self.cols.ctypes.data_as(C.POINTER(C.c_int))
        s.rowOffsets = self.rowOffsets.ctypes.data_as(C.POINTER(C.c_int))
        return s

    @staticmethod
    def _from_c(s: _CSR) -> "CSR":
        m, n = s.numRows, s.numCols
        rp = np.ctypeslib.as_array(s.rowOffsets, shape=(m + 1,)).copy()
        ci = np.ctypeslib.as_array(s.cols, shape=(s.nnz,)).copy() if s.nnz else np.zeros(0, np.int32)
        va = np.ctypeslib.as_array(s.vals, shape=(s.nnz,)).copy() if s.nnz else np.zeros(0)
        lib().pava_delete_csr_matrix(C.byref(s))
        return CSR(m, n, rp, ci, va)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.cols, self.rowOffsets), shape=(self.numRows, self.numCols))


def _gen(fn, *args) -> CSR:
    s = _CSR()
    _check(fn(*args, C.byref(s)))
    return CSR._from_c(s)


def read_mtx(path: str | os.PathLike) -> CSR:
    return _gen(lib().pava_read_csr_matrix, str(path).encode())


def write_csr_binary(path: str | os.PathLike, A: CSR) -> None:
    lib().pava_write_csr_binary.argtypes = [C.c_char_p, C.POINTER(_CSR)]
    s = A.c_struct()
    _check(lib().pava_write_csr_binary(str(path).encode(), C.byref(s)))


def read_csr_binary(path: str | os.PathLike) -> CSR:
    lib().pava_read_csr_binary.argtypes = [C.c_char_p, C.POINTER(_CSR)]
    return _gen(lib().pava_read_csr_binary, str(path).encode())


def write_mtx(path: str | os.PathLike, A: CSR) -> None:
    s = A.c_struct()
    _check(lib().pava_write_matrix_market(str(path).encode(), C.byref(s)))


def validate(A: CSR) -> None:
    s = A.c_struct()
    _check(lib().pava_csr_validate(C.byref(s)))


def laplacian2d(nx: int, ny: int) -> CSR:
    return _gen(lib().pava_gen_laplacian2d, nx, ny)


def stencil27(n: int, jitter: float = 0.01, seed: int = 0) -> CSR:
    return _gen(lib().pava_gen_stencil27, n, jitter, seed)


def rmat(scale: int, edge_factor: int = 16, a: float = 0.57, b: float = 0.19, c: float = 0.19,
         seed: int = 0) -> CSR:
    return _gen(lib().pava_gen_rmat, scale, edge_factor, a, b, c, seed)


def synthetic(rows: int, row_dist: int = RL_POISSON, len_param: float = 10.0, shape: float = 1.0,
              col_struct: int = COL_MIXED, half_band: int = 1000, block: int = 64, band_frac: float = 0.7,
              value_dist: int = 0, max_len: int = 0, seed: int = 0) -> CSR:
    p = _GenParams(rows, row_dist, len_param, shape, col_struct, half_band, block, band_frac, value_dist,
                   max_len, seed)
    return _gen(lib().pava_gen_synthetic, C.byref(p))


def fill_uniform(n: int, seed: int = 0, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    x = np.empty(n, dtype=np.float64)
    lib().pava_fill_uniform(_ptr(x), n, seed, lo, hi)
    return x


def init_vectors(numRows: int, numCols: int):
    x = np.empty(numCols)
    y = np.empty(numRows)
    yr = np.empty(numRows)
    lib().pava_init_vectors(numRows, numCols, _ptr(x), _ptr(y), _ptr(yr))
    return x, y, yr


def reference_spmv(A: CSR, x: np.ndarray) -> np.ndarray:
    """The driver's verification SpMV (main.cpp:318), not the hot path."""
    y = np.empty(A.numRows)
    s = A.c_struct()
    xx = np.ascontiguousarray(x, np.float64)  # kept alive across the call
    lib().pava_reference_spmv(C.byref(s), _ptr(xx), _ptr(y))
    return y


def check_results(y: np.ndarray, y_ref: np.ndarray) -> int:
    y = np.ascontiguousarray(y, np.float64)
    y_ref = np.ascontiguousarray(y_ref, np.float64)
    return lib().pava_check_results(int(y.shape[0]), _ptr(y), _ptr(y_ref))


class CvrHip:
    """A CVR-64 matrix resident on one GPU (the cvr_hip format plugin)."""

    def __init__(self, A: CSR, steps_per_chunk: int = 0, precision: int = PREC_F64, col_index_bits: int = 0,
                 recompute: bool = False, kappa_threshold: float = 64.0, keep_fp64: bool = False, device: int = 0,
                 x_hot_kb: int = -1, x_lds_kb: int = -1, parity_tol: Optional[float] = None):
        L = lib()
        o = self._opts(steps_per_chunk, precision, col_index_bits, recompute, kappa_threshold, keep_fp64, device,
                       x_hot_kb, x_lds_kb, parity_tol)
        self._h = C.c_void_p()
        s = A.c_struct()
        _check(L.cvr_hip_create(C.byref(s), C.byref(o), C.byref(self._h)))
        self.A_shape = (A.numRows, A.numCols)
        self.precision = precision

    @staticmethod
    def _opts(steps_per_chunk=0, precision=PREC_F64, col_index_bits=0, recompute=False, kappa_threshold=64.0,
              keep_fp64=False, device=0, x_hot_kb=-1, x_lds_kb=-1, parity_tol=None) -> _Opts:
        o = _Opts()
        lib().cvr_hip_default_opts(C.byref(o))
        if parity_tol is not None:
            o.parity_tol = parity_tol
        o.steps_per_chunk, o.precision, o.col_index_bits = steps_per_chunk, precision, col_index_bits
        o.recompute, o.kappa_threshold, o.keep_fp64, o.device = int(recompute), kappa_threshold, int(keep_fp64), device
        o.x_hot_kb, o.x_lds_kb = x_hot_kb, x_lds_kb
        return o

    @classmethod
    def from_device(cls, numRows: int, numCols: int, nnz: int, rp_ptr: int, ci_ptr: int, va_ptr: int,
                    **kw) -> "CvrHip":
        """cvr_hip_create_device: a CSR already in device memory (complete and
        valid when called; not retained)."""
        self = cls.__new__(cls)
        self._h = C.c_void_p()
        o = cls._opts(**kw)
        _check(lib().cvr_hip_create_device(numRows, numCols, nnz, rp_ptr, ci_ptr, va_ptr, C.byref(o),
                                           C.byref(self._h)))
        self.A_shape = (numRows, numCols)
        self.precision = o.precision
        return self

    def close(self) -> None:
        if self._h:
            lib().cvr_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self) -> int:
        return self._h.value

    def info(self) -> dict:
        i = _Info()
        _check(lib().cvr_hip_get_info(self._h, C.byref(i)))
        return {f: getattr(i, f) for f, _ in _Info._fields_}

    def spmv_host(self, x: np.ndarray) -> np.ndarray:
        """y = A x with host arrays (copies in and out, synchronous)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.A_shape[0])
        _check(lib().cvr_hip_spmv_host(self._h, _ptr(x), _ptr(y)))
        return y

    def spmv(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        """y = A x on device pointers (FP64 handle), asynchronous on `stream`."""
        _check(lib().cvr_hip_spmv(self._h, x_ptr, y_ptr, stream))

    def spmv_mixed(self, x32_ptr: int, y_ptr: int, x64_ptr: int = 0, stream: int = 0) -> None:
        _check(lib().cvr_hip_spmv_mixed(self._h, x32_ptr, x64_ptr, y_ptr, stream))

    def spmm(self, k: int, X_ptr: int, Y_ptr: int, stream: int = 0) -> None:
        _check(lib().cvr_hip_spmm(self._h, k, X_ptr, Y_ptr, stream))

    def set_spmm_engine(self, engine: int) -> None:
        """SPMM_FMA (per-lane FP64 FMA, the default) or SPMM_MFMA (matrix cores)."""
        _check(lib().cvr_hip_set_spmm_engine(self._h, engine))

    def time_spmv(self, x_ptr: int, y_ptr: int, warmup: int = 10, iters: int = 100, x64_ptr: int = 0):
        """(ms per SpMV, ms of the main kernel alone), hipEvents on the handle's stream."""
        tot, main = C.c_double(), C.c_double()
        _check(lib().cvr_hip_time_spmv_x64(self._h, x_ptr, x64_ptr or None, y_ptr, warmup, iters, C.byref(tot),
                                           C.byref(main)))
        return tot.value, main.value

    def time_spmv_stats(self, x_ptr: int, y_ptr: int, warmup: int = 10, iters: int = 100, x64_ptr: int = 0) -> dict:
        """ms per SpMV (all kernels) and the main kernel's mean / median / min / max ms."""
        st = (C.c_double * 5)()
        L = lib()
        L.cvr_hip_time_spmv_stats.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int, C.c_void_p]
        _check(L.cvr_hip_time_spmv_stats(self._h, x_ptr, x64_ptr or None, y_ptr, warmup, iters, st))
        return {"ms_total": st[0], "main_mean": st[1], "main_median": st[2], "main_min": st[3], "main_max": st[4]}

    def last_recompute_count(self) -> int:
        n = C.c_int()
        _check(lib().cvr_hip_last_recompute_count(self._h, C.byref(n)))
        return n.value

    def hot_cols(self) -> np.ndarray:
        """Columns staged in the hot-x copy (index = hot number)."""
        n = self.info()["n_hot_cols"]
        out = np.zeros(max(n, 1), np.int32)
        lib().cvr_hip_export_hot.argtypes = [C.c_void_p, C.c_void_p]
        _check(lib().cvr_hip_export_hot(self._h, _ptr(out)))
        return out[:n]

    def export_layout(self) -> dict:
        inf = self.info()
        nc, L, segcap = inf["nchunks"], inf["layout_elems"], inf["seg_capacity"]
        vdt = {PREC_F64: np.float64, PREC_F32: np.float32, PREC_BF16: np.uint16}[inf["precision"]]
        col16 = inf["col_index_bits"] == 16
        out = {
            "vals": np.zeros(L, vdt), "cols": np.zeros(L, np.uint16 if col16 else np.int32),
            "chunk_rows": np.zeros(nc + 1, np.int32), "seg": np.zeros(segcap, np.int32),
            "seg_count": np.zeros(max(nc, 1), np.int32), "split_n": np.zeros(max(nc, 1), np.int32),
            "split_tab": np.zeros(2 * SPLIT_CAP * max(nc, 1), np.int32), "fix": np.zeros(2 * max(nc, 1), np.int32),
            "col_base": np.zeros(max(nc, 1), np.int32), "end_mask": np.zeros(max(L // 64, 1), np.uint64),
        }
        _check(lib().cvr_hip_export_layout(self._h, *[_ptr(out[k]) for k in
                                                      ["vals", "cols", "chunk_rows", "seg", "seg_count", "split_n",
                                                       "split_tab", "fix", "col_base", "end_mask"]]))
        out["info"] = inf
        return out


def conduct_cvr_hip(A: CSR, x: np.ndarray, name: str = "matrix", iterations: int = 10) -> np.ndarray:
    """Drop-in conduct_cvr: converts, verifies into y, times, prints the report line."""
    y = np.empty(A.numRows)
    x = np.ascontiguousarray(x, np.float64)
    _check(lib().conduct_cvr_hip(A.numRows, A.numCols, A.nnz, _ptr(A.rowOffsets), _ptr(A.cols), _ptr(A.vals),
                                 _ptr(x), _ptr(y), 1, name.encode(), iterations))
    return y


def version() -> str:
    return lib().pava_hip_version().decode()


class Batch:
    """Batched multi-RHS SpMM over several CvrHip matrices (persistent launch)."""

    def __init__(self, mats: list, k: int):
        arr = (C.c_void_p * len(mats))(*[m.handle for m in mats])
        self._b = C.c_void_p()
        self.mats = mats  # keep the handles alive
        self.k = k
        _check(lib().cvr_hip_batch_create(len(mats), arr, k, C.byref(self._b)))

    def set_operands(self, X_ptrs: list, Y_ptrs: list) -> None:
        n = len(self.mats)
        xs = (C.c_void_p * n)(*X_ptrs)
        ys = (C.c_void_p * n)(*Y_ptrs)
        _check(lib().cvr_hip_batch_set_operands(self._b, xs, ys))

    def spmm(self, stream: int = 0) -> None:
        _check(lib().cvr_hip_batch_spmm(self._b, stream))

    def capture(self, stream: int) -> None:
        _check(lib().cvr_hip_batch_graph(self._b, stream))

    def replay(self, stream: int) -> None:
        _check(lib().cvr_hip_batch_replay(self._b, stream))

    def set_engine(self, engine: int) -> None:
        """SPMM_FMA (the default) or SPMM_MFMA for the batch's launch."""
        _check(lib().cvr_hip_batch_set_engine(self._b, engine))

    def set_x_window(self, wgs_per_cu: int) -> None:
        """LDS window of X (k <= 8) sized for wgs_per_cu workgroups per CU; 0 = off."""
        lib().cvr_hip_batch_set_x_window.argtypes = [C.c_void_p, C.c_int]
        _check(lib().cvr_hip_batch_set_x_window(self._b, wgs_per_cu))

    def info(self) -> dict:
        """items: queue work items (chunks, or groups of 4 with the X window);
        chunks: all chunks of the batch; grid: persistent workgroups."""
        t, g, n = C.c_int(), C.c_int(), C.c_longlong()
        _check(lib().cvr_hip_batch_info(self._b, C.byref(t), C.byref(g), C.byref(n)))
        return {"items": t.value, "chunks": sum(m.info()["nchunks"] for m in self.mats), "grid": g.value,
                "nnz": n.value}

    def queue_counts(self) -> tuple:
        """(pops per queue range in the last completed launch, number of ranges)."""
        counts = (C.c_int * 8)()
        n = C.c_int()
        _check(lib().cvr_hip_batch_queue_counts(self._b, counts, C.byref(n)))
        return list(counts), n.value

    def close(self) -> None:
        if self._b:
            lib().cvr_hip_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Csr5Hip:
    """The csr5_hip format: CSR5 tiles of 64 x sigma nonzeros, FP64."""

    def __init__(self, A: CSR, sigma: int = 16, device: int = 0):
        L = lib()
        P = C.c_void_p
        L.csr5_hip_create.argtypes = [C.POINTER(_CSR), C.c_int, C.c_int, C.POINTER(P)]
        L.csr5_hip_spmv.argtypes = [P] * 4
        L.csr5_hip_spmv_host.argtypes = [P] * 3
        L.csr5_hip_get_info.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.csr5_hip_export_layout.argtypes = [P] * 9
        L.csr5_hip_destroy.argtypes = [P]
        self._h = P()
        s = A.c_struct()
        _check(L.csr5_hip_create(C.byref(s), sigma, device, C.byref(self._h)))
        self.shape = (A.numRows, A.numCols)
        self.n_nonempty = int(np.count_nonzero(np.diff(A.rowOffsets)))

    def info(self) -> dict:
        nt, sg, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().csr5_hip_get_info(self._h, C.byref(nt), C.byref(sg), C.byref(ms)))
        return {"ntiles": nt.value, "sigma": sg.value, "convert_ms": ms.value}

    def spmv(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        """y = A x on device pointers, asynchronous on `stream`."""
        _check(lib().csr5_hip_spmv(self._h, x_ptr, y_ptr, stream))

    def spmv_host(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty(self.shape[0])
        _check(lib().csr5_hip_spmv_host(self._h, _ptr(x), _ptr(y)))
        return y

    def export_layout(self) -> dict:
        inf = self.info()
        nt, sg = inf["ntiles"], inf["sigma"]
        L = nt * 64 * sg
        out = {"vals": np.zeros(max(L, 1)), "cols": np.zeros(max(L, 1), np.int32),
               "flags": np.zeros(max(nt * 64, 1), np.uint32), "yoff": np.zeros(max(nt * 64, 1), np.int32),
               "tile_base": np.zeros(max(nt, 1), np.int32), "nerows": np.zeros(max(self.n_nonempty, 1), np.int32),
               "fix": np.zeros(max(2 * nt, 2), np.int32), "cont": np.zeros(max(nt, 1), np.uint8)}
        _check(lib().csr5_hip_export_layout(self._h, *[_ptr(out[k]) for k in ["vals", "cols", "flags", "yoff",
                                                                               "tile_base", "nerows", "fix",
                                                                               "cont"]]))
        out.update(inf)
        return out

    def close(self) -> None:
        if self._h:
            lib().csr5_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def conduct_csr5_hip(A: CSR, x: np.ndarray, name: str = "matrix", iterations: int = 10) -> np.ndarray:
    """Drop-in conduct_csr5: converts, verifies into y, times, prints the report line."""
    L = lib()
    L.conduct_csr5_hip.argtypes = [C.c_int] * 3 + [C.c_void_p] * 5 + [C.c_double, C.c_char_p, C.c_int]
    y = np.empty(A.numRows)
    x = np.ascontiguousarray(x, np.float64)
    _check(L.conduct_csr5_hip(A.numRows, A.numCols, A.nnz, _ptr(A.rowOffsets), _ptr(A.cols), _ptr(A.vals),
                              _ptr(x), _ptr(y), 1.0, name.encode(), iterations))
    return y


def conduct_coo_hip(A: CSR, x: np.ndarray, name: str = "matrix", iterations: int = 10) -> np.ndarray:
    """The COO_HIP driver block: converts, verifies into y, times, prints the report line."""
    L = lib()
    L.conduct_coo_hip.argtypes = [C.c_int] * 3 + [C.c_void_p] * 5 + [C.c_char_p, C.c_int]
    y = np.empty(A.numRows)
    x = np.ascontiguousarray(x, np.float64)
    _check(L.conduct_coo_hip(A.numRows, A.numCols, A.nnz, _ptr(A.rowOffsets), _ptr(A.cols), _ptr(A.vals),
                             _ptr(x), _ptr(y), name.encode(), iterations))
    return y


class CooHip:
    """The coo_hip format: COO SpMV (row index per nonzero), FP64, deterministic."""

    def __init__(self, A: CSR, device: int = 0):
        L = lib()
        L.coo_hip_create.argtypes = [C.POINTER(_CSR), C.c_int, C.POINTER(C.c_void_p)]
        L.coo_hip_spmv.argtypes = [C.c_void_p] * 4
        L.coo_hip_get_info.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.coo_hip_destroy.argtypes = [C.c_void_p]
        self._h = C.c_void_p()
        s = A.c_struct()
        _check(L.coo_hip_create(C.byref(s), device, C.byref(self._h)))
        self.shape = (A.numRows, A.numCols)

    def info(self) -> dict:
        nt, ms = C.c_int(), C.c_double()
        _check(lib().coo_hip_get_info(self._h, C.byref(nt), C.byref(ms)))
        return {"ntiles": nt.value, "convert_ms": ms.value}

    def spmv(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        _check(lib().coo_hip_spmv(self._h, x_ptr, y_ptr, stream))

    def close(self) -> None:
        if self._h:
            lib().coo_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class CsrHip:
    """The csr_hip format: CSR SpMV with `lanes_per_row` lanes per row."""

    def __init__(self, A: CSR, lanes_per_row: int = 8, precision: int = PREC_F64, device: int = 0):
        L = lib()
        L.csr_hip_create.argtypes = [C.POINTER(_CSR), C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.csr_hip_spmv.argtypes = [C.c_void_p] * 4
        L.csr_hip_destroy.argtypes = [C.c_void_p]
        self._h = C.c_void_p()
        s = A.c_struct()
        _check(L.csr_hip_create(C.byref(s), lanes_per_row, precision, device, C.byref(self._h)))
        self.shape = (A.numRows, A.numCols)
        self.precision = precision

    def spmv(self, x_ptr: int, y_ptr: int, stream: int = 0) -> None:
        _check(lib().csr_hip_spmv(self._h, x_ptr, y_ptr, stream))

    def close(self) -> None:
        if self._h:
            lib().csr_hip_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

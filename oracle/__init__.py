"""CPU oracle for the CVR SpMV path -- TEST INFRASTRUCTURE ONLY.

Python loader for ``oracle/liboracle.so`` (built from ``oracle/oracle.c`` by
``oracle/build.py``). Used by tests/, ``__graft_entry__.smoke()`` and the
cpu_baseline leg of bench.py as the checker; the product never imports it.
Each function's reference citation and what pins it are in oracle.c's header.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liboracle.so"
SPLIT_CAP = 72
_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build as _b
            _b.build()
        L = C.CDLL(str(LIB_PATH))
        P, i = C.c_void_p, C.c_int
        for f in ["orc_csr_spmv_fma", "orc_csr_spmv_muladd", "orc_csr_spmv_dd", "orc_row_abs"]:
            getattr(L, f).argtypes = [i, P, P, P, P, P]
            getattr(L, f).restype = None
        L.orc_csr_spmv_omp.argtypes = [i, P, P, P, P, P, i]
        L.orc_csr_spmv_omp.restype = None
        L.orc_check_results.argtypes = [i, P, P]
        L.orc_cvr64_sizes.argtypes = [i, i, i, C.POINTER(i), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        L.orc_cvr64_sizes.restype = None
        L.orc_cvr64_build.argtypes = [i, i, i, P, P, P, i, i] + [P] * 9
        L.orc_cvr64_spmv.argtypes = [i, i, i, i] + [P] * 10
        L.orc_cvr64_spmv.restype = None
        L.orc_cvr64_spmv_par.argtypes = [i, i, i, i] + [P] * 12 + [C.c_double, P, P]
        L.orc_cvr64_spmv_par.restype = None
        L.orc_csr5_ntiles.argtypes = [i, i]
        L.orc_csr5_build.argtypes = [i, i, P, P, P, i] + [P] * 8
        L.orc_csr5_spmv.argtypes = [i, i, i] + [P] * 10
        L.orc_csr5_spmv.restype = None
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def _csr(A):
    return (np.ascontiguousarray(A.rowOffsets, np.int32), np.ascontiguousarray(A.cols, np.int32),
            np.ascontiguousarray(A.vals, np.float64))


def _spmv(fn, A, x):
    rp, ci, va = _csr(A)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.numRows)
    fn(A.numRows, _p(rp), _p(ci), _p(va), _p(x), _p(y))
    return y


def spmv_fma(A, x):
    """referenceSpMV (utils/basicMatrix.cpp:538-556), sequential fma per row."""
    return _spmv(lib().orc_csr_spmv_fma, A, x)


def spmv_muladd(A, x):
    return _spmv(lib().orc_csr_spmv_muladd, A, x)


def spmv_dd(A, x):
    """Near-exact (double-double accumulated) y."""
    return _spmv(lib().orc_csr_spmv_dd, A, x)


def row_abs(A, x):
    return _spmv(lib().orc_row_abs, A, x)


def spmv_omp(A, x, threads: int):
    rp, ci, va = _csr(A)
    x = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.numRows)
    lib().orc_csr_spmv_omp(A.numRows, _p(rp), _p(ci), _p(va), _p(x), _p(y), threads)
    return y


def check_results(y, y_ref) -> int:
    y = np.ascontiguousarray(y, np.float64)
    y_ref = np.ascontiguousarray(y_ref, np.float64)
    return lib().orc_check_results(int(y.shape[0]), _p(y), _p(y_ref))


def cvr64_sizes(m: int, nnz: int, S: int):
    nc, L, sc = C.c_int(), C.c_longlong(), C.c_longlong()
    lib().orc_cvr64_sizes(m, nnz, S, C.byref(nc), C.byref(L), C.byref(sc))
    return nc.value, L.value, sc.value


def cvr64_build(A, S: int = 256, prec: int = 0) -> dict:
    """CPU emulation of the GPU CVR-64 layout (same arrays as CvrHip.export_layout)."""
    rp, ci, va = _csr(A)
    nc, L, sc = cvr64_sizes(A.numRows, A.nnz, S)
    vdt = {0: np.float64, 1: np.float32, 2: np.uint16}[prec]
    out = {"vals": np.zeros(L, vdt), "cols": np.zeros(L, np.int32), "chunk_rows": np.zeros(nc + 1, np.int32),
           "seg": np.zeros(sc, np.int32), "seg_count": np.zeros(max(nc, 1), np.int32),
           "split_n": np.zeros(max(nc, 1), np.int32), "split_tab": np.zeros(2 * SPLIT_CAP * max(nc, 1), np.int32),
           "fix": np.zeros(2 * max(nc, 1), np.int32), "info": np.zeros(3, np.int32)}
    rc = lib().orc_cvr64_build(A.numRows, A.numCols, A.nnz, _p(rp), _p(ci), _p(va), S, prec,
                               *[_p(out[k]) for k in ["vals", "cols", "chunk_rows", "seg", "seg_count", "split_n",
                                                      "split_tab", "fix", "info"]])
    if rc != 0:
        raise RuntimeError(f"orc_cvr64_build failed: {rc}")
    out.update(nchunks=nc, layout_elems=L, seg_capacity=sc, S=S, prec=prec)
    return out


PARITY_TOL = 1e-12  # the library's default parity_tol for FP64 handles


def cvr64_spmv(A, lay: dict, x, parity_tol: float = PARITY_TOL, details: bool = False):
    """y over an emulated layout, in the GPU kernel's summation order. For an
    FP64 layout, split rows whose running error bound cannot show agreement
    with the serial sum within parity_tol are recomputed serially, as the GPU
    does with its default parity_tol (0: the plain CVR order). details=True
    returns (y, bound, n_recomputed), bound[r] = u * A of split row r (0 on
    unsplit rows)."""
    rp, ci, va = _csr(A)
    prec = lay["prec"]
    if prec == 0:
        xx = np.ascontiguousarray(x, np.float64)
        y = np.empty(A.numRows, np.float64)
    else:
        xx = np.ascontiguousarray(x, np.float32)
        y = np.empty(A.numRows, np.float32)
        parity_tol = 0.0
    bound = np.zeros(A.numRows) if details else None
    nrc = C.c_int(0)
    lib().orc_cvr64_spmv_par(A.numRows, A.nnz, lay["S"], prec, *[_p(lay[k]) for k in
                                                                  ["vals", "cols", "chunk_rows", "seg", "split_n",
                                                                   "split_tab", "fix"]], _p(rp), _p(ci), _p(va),
                             _p(xx), _p(y), float(parity_tol), _p(bound) if details else None, C.byref(nrc))
    if details:
        return y, bound, nrc.value
    return y


def error_stats(y, y_ref, scale=None) -> dict:
    """Per-row relative error |y - y_ref| / |y_ref| (rows with y_ref == 0 use
    the absolute error), and, with scale = sum_j |a_ij x_j|, the normalised
    error |y - y_ref| / scale."""
    y = np.asarray(y, np.float64)
    y_ref = np.asarray(y_ref, np.float64)
    d = np.abs(y - y_ref)
    den = np.abs(y_ref)
    rel = np.where(den > 0, d / np.where(den > 0, den, 1.0), d)
    out = {"max_rel": float(rel.max()) if rel.size else 0.0,
           "median_rel": float(np.median(rel)) if rel.size else 0.0}
    if scale is not None:
        sc = np.asarray(scale, np.float64)
        nrm = np.where(sc > 0, d / np.where(sc > 0, sc, 1.0), d)
        out["max_scaled"] = float(nrm.max()) if nrm.size else 0.0
    return out


def csr5_build(A, sigma: int = 16) -> dict:
    """CPU emulation of the GPU CSR5-64 layout (same arrays as Csr5Hip.export_layout)."""
    rp, ci, va = _csr(A)
    nt = lib().orc_csr5_ntiles(A.nnz, sigma)
    L = nt * 64 * sigma
    out = {"vals": np.zeros(max(L, 1)), "cols": np.zeros(max(L, 1), np.int32),
           "flags": np.zeros(max(nt * 64, 1), np.uint32), "yoff": np.zeros(max(nt * 64, 1), np.int32),
           "tile_base": np.zeros(max(nt, 1), np.int32), "nerows": np.zeros(max(A.numRows, 1), np.int32),
           "fix": np.zeros(max(2 * nt, 2), np.int32), "cont": np.zeros(max(nt, 1), np.uint8)}
    rc = lib().orc_csr5_build(A.numRows, A.nnz, _p(rp), _p(ci), _p(va), sigma,
                              *[_p(out[k]) for k in ["vals", "cols", "flags", "yoff", "tile_base", "nerows", "fix",
                                                     "cont"]])
    if rc != 0:
        raise RuntimeError(f"orc_csr5_build failed: {rc}")
    out.update(ntiles=nt, sigma=sigma, n_nonempty=int(np.count_nonzero(np.diff(rp))))
    return out


def csr5_spmv(A, lay: dict, x):
    """y over an emulated CSR5-64 layout, in the GPU kernel's summation order."""
    xx = np.ascontiguousarray(x, np.float64)
    y = np.empty(A.numRows)
    lib().orc_csr5_spmv(A.numRows, A.nnz, lay["sigma"], *[_p(lay[k]) for k in
                                                          ["vals", "cols", "flags", "yoff", "tile_base", "nerows",
                                                           "fix", "cont"]], _p(xx), _p(y))
    return y

"""The C-ABI library loads, exports every symbol include/pava_hip.h declares,
and its host-side checks fail cleanly (no GPU needed)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    text = (ROOT / "include" / "pava_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?(?:long\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_symbols_exported(P):
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(P.LIB_PATH)], capture_output=True, text=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    assert sorted(P.EXPORTED) == names


def test_library_has_gfx950_code(P):
    # the embedded code objects carry the target id amdgcn-amd-amdhsa--gfx950
    data = P.LIB_PATH.read_bytes()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def test_version(P):
    assert "gfx950" in P.version()


@pytest.mark.parametrize("kw", [dict(steps_per_chunk=7), dict(steps_per_chunk=8192), dict(precision=5),
                                dict(col_index_bits=12), dict(x_hot_kb=-2)])
def test_invalid_options_rejected_before_device(P, kw):
    A = P.laplacian2d(10, 10)
    with pytest.raises(P.PavaError) as e:
        P.CvrHip(A, **kw)
    assert e.value.code == -1


def test_malformed_csr_rejected(P):
    bad = [
        P.CSR(3, 3, np.array([0, 2, 1, 3]), np.array([0, 1, 2]), np.ones(3)),  # offsets not monotone
        P.CSR(3, 3, np.array([0, 1, 2, 3]), np.array([0, 5, 2]), np.ones(3)),  # column out of range
        P.CSR(3, 3, np.array([1, 1, 2, 3]), np.array([0, 1, 2]), np.ones(3)),  # offsets[0] != 0
    ]
    for A in bad:
        with pytest.raises(P.PavaError):
            P.validate(A)
        with pytest.raises(P.PavaError):
            P.CvrHip(A)


def test_print_performance_line(P, capfd):
    P.lib().pava_print_performance(b"grp mat", b"CVR_HIP", 4, 0.5, 0.25)
    assert capfd.readouterr().out == " Performance grp mat CVR_HIP 4 0.500000 0.250000 \n"


def test_init_vectors_pattern(P):
    x, y, yr = P.init_vectors(5, 7)
    assert x.tolist() == [0, 1, 2, 0, 1, 2, 0] and not y.any() and not yr.any()


def test_sizes_beyond_gpu_limit_rejected_before_device(P):
    """Every GPU format rejects a dimension or nnz above PAVA_HIP_MAX_SIZE
    (2^31 - 2^16) with PAVA_HIP_ERANGE before it reads the arrays or touches a
    device; the limit keeps element positions plus padding inside int32."""
    text = (ROOT / "include" / "pava_hip.h").read_text()
    limit = int(re.search(r"#define PAVA_HIP_MAX_SIZE (\d+)", text).group(1))
    assert limit == 2 ** 31 - 2 ** 16
    L = P.lib()
    V = C.c_void_p
    L.csr5_hip_create.argtypes = [C.POINTER(P._CSR), C.c_int, C.c_int, C.POINTER(V)]
    L.coo_hip_create.argtypes = [C.POINTER(P._CSR), C.c_int, C.POINTER(V)]
    L.csr_hip_create.argtypes = [C.POINTER(P._CSR), C.c_int, C.c_int, C.c_int, C.POINTER(V)]
    rp = np.zeros(2, np.int32)
    one = np.zeros(1, np.int32)
    val = np.zeros(1)
    for M, N, nnz in [(limit + 1, 4, 0), (1, limit + 1, 0), (1, 4, limit + 1)]:
        s = P._CSR()
        s.numRows, s.numCols, s.nnz = M, N, nnz
        s.rowOffsets = rp.ctypes.data_as(C.POINTER(C.c_int))
        s.cols = one.ctypes.data_as(C.POINTER(C.c_int))
        s.vals = val.ctypes.data_as(C.POINTER(C.c_double))
        o = P._Opts()
        L.cvr_hip_default_opts(C.byref(o))
        h = V()
        calls = {
            "cvr_hip_create": lambda: L.cvr_hip_create(C.byref(s), C.byref(o), C.byref(h)),
            "cvr_hip_create_device": lambda: L.cvr_hip_create_device(M, N, nnz, rp.ctypes.data, one.ctypes.data,
                                                                     val.ctypes.data, C.byref(o), C.byref(h)),
            "csr5_hip_create": lambda: L.csr5_hip_create(C.byref(s), 16, 0, C.byref(h)),
            "coo_hip_create": lambda: L.coo_hip_create(C.byref(s), 0, C.byref(h)),
            "csr_hip_create": lambda: L.csr_hip_create(C.byref(s), 4, 0, 0, C.byref(h)),
        }
        for name, call in calls.items():
            assert call() == -6, (name, M, N, nnz)
            assert "limit" in L.pava_hip_last_error().decode(), name
            assert not h.value, name


def test_columns_beyond_2_30_rejected_for_32bit_layouts(P):
    """32-bit CVR layouts keep flags in bits 31 and 30 of a column: numCols
    above 2^30 is PAVA_HIP_ERANGE before anything is read or a device is used."""
    L = P.lib()
    V = C.c_void_p
    rp = np.zeros(2, np.int32)
    one = np.zeros(1, np.int32)
    val = np.zeros(1)
    N = 2 ** 30 + 1
    s = P._CSR()
    s.numRows, s.numCols, s.nnz = 1, N, 0
    s.rowOffsets = rp.ctypes.data_as(C.POINTER(C.c_int))
    s.cols = one.ctypes.data_as(C.POINTER(C.c_int))
    s.vals = val.ctypes.data_as(C.POINTER(C.c_double))
    o = P._Opts()
    L.cvr_hip_default_opts(C.byref(o))
    o.col_index_bits = 32
    h = V()
    assert L.cvr_hip_create(C.byref(s), C.byref(o), C.byref(h)) == -6
    assert "2^30" in L.pava_hip_last_error().decode()
    assert L.cvr_hip_create_device(1, N, 0, rp.ctypes.data, one.ctypes.data, val.ctypes.data, C.byref(o),
                                   C.byref(h)) == -6
    assert not h.value

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
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*([a-z_0-9]+)\s*\(", text, flags=re.M)
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
                                dict(col_index_bits=16), dict(col_index_bits=12)])
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

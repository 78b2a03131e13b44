"""Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).

Two harnesses under tests/sanitize/ are compiled with gcc/g++
``-fsanitize=address,undefined -fno-sanitize-recover=all`` and run:

* oracle_harness.c: the oracle's CVR-64 and CSR5-64 emulations (layout build
  and SpMV) on 60 random CSR matrices with empty, short and long rows, every
  buffer allocated at its exact size, y checked against the serial FMA sum.
* host_harness.cpp: the library's host side without any HIP code (Matrix
  Market reader, COO->CSR, binary cache, writer, generators, features) on
  valid files (duplicates, symmetric/pattern, CRLF, no final newline, empty
  matrix, a file above the parallel-parse threshold) and on malformed or
  truncated ones, each of which must be rejected with an error; and with the
  n-th host allocation failing (pava_hip_test_fail_alloc), where the reader
  returns PAVA_MM_ENOMEM (-5, as readCOOMatrix, utils/basicMatrix.cpp:84-132)
  and the converter PAVA_HIP_ENOMEM, with nothing leaked (LeakSanitizer).

A planted out-of-bounds read is compiled the same way first, to show that
the sanitizer is active in this toolchain (a harness that "passes" with the
instrumentation silently missing would prove nothing)."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None,
                                reason="gcc/g++ not on PATH")


def _build(tmp: Path, cc: str, srcs, out: str, extra=()):
    exe = tmp / out
    cmd = [cc, *SAN, *extra, *map(str, srcs), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    return exe


def _run(exe: Path, *args, timeout=300):
    env = {"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
           "PATH": "/usr/bin:/bin", "OMP_NUM_THREADS": "4"}
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=timeout, env=env)


def test_sanitizer_is_active(tmp_path):
    src = tmp_path / "oob.c"
    src.write_text("#include <stdlib.h>\nint main(int c, char **v) { int *a = malloc(4 * sizeof(int));"
                   " volatile int r = a[c + 3]; free(a); return r == 12345; }\n")
    r = _run(_build(tmp_path, "gcc", [src], "oob"))
    # UBSan's object-size check or ASan's shadow check, whichever reports first
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr or "insufficient space" in r.stderr, r.stderr[-2000:]


def test_oracle_emulations_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "gcc", [ROOT / "tests/sanitize/oracle_harness.c", ROOT / "oracle/oracle.c"], "oracle_h",
                 ["-I", str(ROOT / "oracle"), "-fopenmp", "-lm"])
    r = _run(exe)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "0 failures" in r.stdout


def test_host_io_under_sanitizers(tmp_path):
    cs = ROOT / "pava_amd/csrc"
    exe = _build(tmp_path, "g++", [ROOT / "tests/sanitize/host_harness.cpp", cs / "host_io.cpp", cs / "gen.cpp",
                                   cs / "features.cpp"], "host_h",
                 ["-std=c++17", "-fopenmp", "-I", str(ROOT / "include"), "-I", str(cs)])
    scratch = tmp_path / "files"
    scratch.mkdir()
    r = _run(exe, scratch)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "0 failures" in r.stdout

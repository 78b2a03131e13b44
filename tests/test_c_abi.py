"""The C-ABI header is plain C: a C99 program includes it, links against
libpava_hip.so and calls the host-side entry points (no GPU needed), as a C,
cgo or FFI caller would."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "pava_hip.h"

int main(void) {
    cvr_hip_opts o;
    cvr_hip_default_opts(&o);
    if (o.precision != CVR_HIP_PREC_F64 || o.steps_per_chunk != 0 || o.col_index_bits != 0 || o.x_hot_kb != -1)
        return 1;
    /* 3 x 3: [[1, 0, 2], [0, 0, 0], [0, 3, 0]] */
    int rp[4] = {0, 2, 2, 3}, ci[3] = {0, 2, 1};
    double va[3] = {1.0, 2.0, 3.0}, x[3] = {1.0, 10.0, 100.0}, y[3] = {-1.0, -1.0, -1.0};
    pava_csr_matrix A = {3, 3, 3, va, ci, rp};
    if (pava_csr_validate(&A) != 0) return 2;
    pava_reference_spmv(&A, x, y);
    if (y[0] != 201.0 || y[1] != 0.0 || y[2] != 30.0) return 3;
    /* beyond the GPU formats' size limit: refused before any array access */
    pava_csr_matrix big = {PAVA_HIP_MAX_SIZE + 1, 1, 4, va, ci, rp};
    cvr_hip_handle *h = NULL;
    if (cvr_hip_create(&big, &o, &h) != PAVA_HIP_ERANGE || h != NULL) return 4;
    if (strstr(pava_hip_last_error(), "limit") == NULL) return 5;
    printf("%s\n", pava_hip_version());
    return 0;
}
"""


def test_c_program_builds_links_and_runs(P, tmp_path):
    src = tmp_path / "abi.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "abi"
    lib_dir = P.LIB_PATH.parent
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", str(ROOT / "include"), str(src),
           "-o", str(exe), "-L", str(lib_dir), "-l:" + P.LIB_PATH.name, "-Wl,-rpath," + str(lib_dir)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert "gfx950" in run.stdout

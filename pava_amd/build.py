"""Builds the in-tree shared library ``pava_amd/libpava_hip.so`` for gfx950.

HIP sources are compiled with hipcc (``--offload-arch=gfx950``), host C++ with
g++ (OpenMP via libgomp, the runtime PyTorch also loads), and the objects are
linked with g++ against libamdhip64. Run ``python -m pava_amd.build``.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OBJDIR = HERE / "_build"
LIB = HERE / "libpava_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIP_SOURCES = ["cvr_convert.hip", "cvr_spmv.hip", "cvr_spmm.hip", "cvr_handle.hip", "csr_vector.hip", "csr5_hip.hip",
               "coo_hip.hip", "mlp.hip", "driver.hip"]
CXX_SOURCES = ["host_io.cpp", "gen.cpp", "features.cpp"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
HIP_FLAGS = COMMON + [
    "--offload-arch=gfx950",
    # packed FP32 (v_pk_*_f32) is kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
]
CXX_FLAGS = COMMON + ["-fopenmp", "-ffp-contract=off", "-march=x86-64-v2"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"compile failed: {cmd[-1]}")
    err = "\n".join(
        l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip()
    )
    if err:
        sys.stderr.write(err + "\n")


def build(verbose: bool = False) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = OBJDIR / (src + ".o")
        jobs.append([str(ROCM / "bin" / "hipcc"), *HIP_FLAGS, "-c", "-o", str(obj), str(CSRC / src)])
    for src in CXX_SOURCES:
        obj = OBJDIR / (src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", "-o", str(obj), str(CSRC / src)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        list(ex.map(_run, jobs))
    objs = [str(OBJDIR / (s + ".o")) for s in HIP_SOURCES + CXX_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-o", str(tmp), *objs, f"-L{ROCM / 'lib'}", "-lamdhip64", "-fopenmp",
          f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)

"""Builds the CPU oracle ``oracle/liboracle.so`` (test infrastructure).

gcc, -O2 with -ffp-contract=off so the only fused operations are the explicit
fma() calls. ``python -m oracle.build`` or ``build()``.
"""
from __future__ import annotations

import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           "-Wall", "-o", str(tmp), str(HERE / "oracle.c"), "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build())

"""Writes tests/golden/scipy_spmv.npz: seeded random CSR matrices, x, and y
computed by scipy.sparse (an independent SpMV implementation), used to
cross-check the oracle. Run: python tests/golden/make_fixtures.py"""
from pathlib import Path

import numpy as np
import scipy.sparse as sp

OUT = Path(__file__).resolve().parent / "scipy_spmv.npz"


def main():
    rng = np.random.default_rng(20261203)
    arrays = {}
    shapes = [(37, 41, 0.2), (200, 150, 0.05), (1000, 1000, 0.01), (64, 64, 0.9)]
    for k, (m, n, dens) in enumerate(shapes):
        A = sp.random(m, n, density=dens, format="csr", random_state=rng, dtype=np.float64)
        A.data = rng.uniform(-1, 1, A.nnz)
        A.sort_indices()
        x = rng.uniform(-1, 1, n)
        y = A @ x
        arrays[f"m{k}_shape"] = np.array([m, n])
        arrays[f"m{k}_rp"] = A.indptr.astype(np.int32)
        arrays[f"m{k}_ci"] = A.indices.astype(np.int32)
        arrays[f"m{k}_va"] = A.data
        arrays[f"m{k}_x"] = x
        arrays[f"m{k}_y"] = y
    np.savez_compressed(OUT, count=np.array(len(shapes)), **arrays)
    print(OUT)


if __name__ == "__main__":
    main()

"""Failure paths on the GPU (needs an MI355X): with the n-th device allocation
forced to fail (pava_hip_test_fail_alloc), cvr_hip_create and
cvr_hip_batch_create return PAVA_HIP_ENOMEM and release every device
allocation they had taken, for every n until creation succeeds. The
reference checks one allocation and prints (cvr/cvr.h:2021-2022); here each
allocation of the create path is forced to fail in turn."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ENOMEM = -2


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _sweep(P, make):
    L = P.lib()
    base = L.pava_hip_test_live_device_allocs()
    failures = 0
    for n in range(1, 200):
        L.pava_hip_test_fail_alloc(0, n)
        try:
            obj = make()
        except P.PavaError as e:
            assert e.code == ENOMEM, (n, str(e))
            assert L.pava_hip_test_live_device_allocs() == base, (n, str(e))
            failures += 1
            continue
        finally:
            L.pava_hip_test_fail_alloc(0, 0)
        obj.close()
        assert L.pava_hip_test_live_device_allocs() == base
        return failures
    raise AssertionError("creation never succeeded")


@pytest.mark.parametrize("case", ["rmat_fp64_parity_lds", "laplacian_col16", "rmat_fp32_recompute", "device_create"])
def test_device_allocation_failure_releases_everything(P, case):
    import torch
    A = P.rmat(13, 16, seed=5) if case != "laplacian_col16" else P.laplacian2d(200, 300)
    _say(f"alloc failure sweep: {case}")
    if case == "device_create":
        rp = torch.from_numpy(A.rowOffsets.astype(np.int32)).cuda()
        ci = torch.from_numpy(A.cols.astype(np.int32)).cuda()
        va = torch.from_numpy(A.vals).cuda()
        make = lambda: P.CvrHip.from_device(A.numRows, A.numCols, A.nnz, rp.data_ptr(), ci.data_ptr(),
                                            va.data_ptr(), steps_per_chunk=16)
    elif case == "rmat_fp32_recompute":
        make = lambda: P.CvrHip(A, precision=1, recompute=True, steps_per_chunk=16)
    else:
        make = lambda: P.CvrHip(A, steps_per_chunk=16)
    n = _sweep(P, make)
    _say(f"{case}: {n} allocation sites failed cleanly")
    # the handle's arrays come from one arena allocation; the CSR copy, the
    # temporaries of hot-x staging and the parity plan are allocated apart
    assert n >= 3
    # the library still works afterwards
    x = P.fill_uniform(A.numCols, 1)
    with P.CvrHip(A, steps_per_chunk=16) as h:
        assert np.allclose(h.spmv_host(x), P.reference_spmv(A, x), rtol=1e-12, atol=1e-300)


def test_batch_allocation_failure_releases_everything(P):
    mats = [P.synthetic(3000 + 500 * i, P.RL_POISSON, 12.0, col_struct=P.COL_MIXED, half_band=300, band_frac=0.7,
                        seed=40 + i) for i in range(3)]
    hs = [P.CvrHip(A, precision=P.PREC_F32, steps_per_chunk=16) for A in mats]
    try:
        n = _sweep(P, lambda: P.Batch(hs, 8))
        assert n >= 3
    finally:
        for h in hs:
            h.close()


def test_duplicate_handle_in_batch_rejected(P):
    A = P.laplacian2d(40, 40)
    with P.CvrHip(A, precision=P.PREC_F32, steps_per_chunk=16) as h:
        with pytest.raises(P.PavaError) as e:
            P.Batch([h, h], 8)
        assert e.value.code == -1 and "more than once" in str(e.value)

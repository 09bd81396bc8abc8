"""GPU parity of aclblasSmatinvBatched (getrf followed by getri, n <= 32).

Called through the C ABI on the reference's CSV cases (tests/golden/cases/
smatinv_batched.csv, from test/matinv_batched/arch35/matinv_batched_test.csv).
The bar: the inverse and info bit-exact against the oracle's composition
(oracle.matinv_batched: its getrf, then its getri), columns beyond n of each
lda_inv-strided output untouched, and for nonsingular members within the
reference's MIXED_TOLERANCE (test/frame/verify.h:276-333) of LAPACK's
inverse, the reference's golden (matinv_batched_golden.h:26-95).
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host

pytestmark = pytest.mark.gpu

PAD = 9.0


def _matinv(handle_value, A, ldi):
    import torch
    batch, n, lda = A.shape
    dA = to_dev(A)
    C = torch.full((batch, n, ldi), PAD, dtype=torch.float32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    pa, pc = B.ptr_array(dA), B.ptr_array(C)
    st = B.smatinv_batched(handle_value, n, pa.data_ptr(), lda, pc.data_ptr(), ldi, info.data_ptr(), batch)
    return st, to_host(C), to_host(info)


CASES = [r for r in load_cases("smatinv_batched.csv")]


@pytest.mark.parametrize("row", CASES, ids=[r["case_name"] for r in CASES])
def test_matinv_reference_cases(handle, row):
    n, lda, ldi, batch = int(row["n"]), int(row["lda"]), int(row["lda_inv"]), int(row["batch_size"])
    expect = status_code(row["expect_result"])
    kind = row["matrix_type"]
    if expect != 0 or n <= 0 or batch <= 0 or kind.startswith("NULLPTR"):
        pytest.skip("argument-check row: covered by tests/test_abi.py on the CPU")
    print("matinv case", row["case_name"], n, lda, ldi, batch, kind, flush=True)
    A = O.lapack_batch(n, lda, kind, int(row["random_seed"]), batch)
    st, C, info = _matinv(handle.value, A, ldi)
    assert st == expect
    gC, ginfo = O.matinv_batched(A, lda_inv=ldi)
    assert np.array_equal(info, ginfo), (info, ginfo)
    assert np.array_equal(C[:, :, :n], gC[:, :, :n], equal_nan=True)
    assert np.all(C[:, :, n:] == PAD)
    if float(row["mere_threshold"]) <= 0.0:
        return                      # the reference checks only info on these rows
    for b in range(batch):
        if ginfo[b] != 0:
            continue
        lu, piv, linfo = lapack.sgetrf(A[b, :, :n].T)
        ref, linfo2 = lapack.sgetri(lu, piv)
        assert linfo == 0 and linfo2 == 0
        ok, detail = O.mixed_tolerance(C[b, :, :n].T, ref)
        assert ok, detail


def test_matinv_singular_member_info_and_neighbours(handle):
    # one member with a zero column among nonsingular ones: its info is the
    # column index + 1, the others are untouched by it
    n, batch = 12, 8
    A = O.normal((batch, n, n), 3, n=n, batch=batch, shift=4.0)
    A[5, 4, :] = 0.0
    st, C, info = _matinv(handle.value, A, n)
    assert st == 0
    gC, ginfo = O.matinv_batched(A)
    assert np.array_equal(info, ginfo) and info[5] == 5 and np.count_nonzero(info) == 1
    assert np.array_equal(C, gC, equal_nan=True)


def test_matinv_full_size_batch(handle):
    # 4096 systems of order 32 (the LU workload's shape): bit-exact and A A^-1 = I
    n, batch = 32, 4096
    A = O.normal((batch, n, n), 11, n=n, batch=batch, shift=4.0)
    st, C, info = _matinv(handle.value, A, n)
    assert st == 0 and np.all(info == 0)
    gC, _ = O.matinv_batched(A)
    assert np.array_equal(C, gC)
    Ad, Cd = A.astype(np.float64), C.astype(np.float64)
    prod = np.einsum("bji,bkj->bki", Ad, Cd)          # (A Ainv) column-major
    err = np.abs(prod - np.eye(n)).max(axis=(1, 2))
    # normwise residual bound of an LU-based inverse: n u ||A||inf ||Ainv||inf
    norm_a = np.abs(Ad).sum(axis=1).max(axis=1)
    norm_c = np.abs(Cd).sum(axis=1).max(axis=1)
    assert np.all(err <= n * 2.0 ** -24 * norm_a * norm_c)


def test_matinv_workspace_too_small_is_execution_failed():
    import torch
    n, batch = 32, 64
    need = batch * 8 + n * n * batch * 4 + n * batch * 4
    A = O.normal((batch, n, n), 17, n=n, batch=batch, shift=4.0)
    small = torch.empty(need - 4, dtype=torch.uint8, device="cuda")
    exact = torch.empty(need, dtype=torch.uint8, device="cuda")
    with B.Handle() as h:
        assert h.set_stream(torch.cuda.current_stream().cuda_stream) == 0
        lib = B.load()
        assert lib.aclblasSetWorkspace(h.value, ctypes.c_void_p(small.data_ptr()), small.numel()) == 0
        st, _, _ = _matinv(h.value, A, n)
        assert st == status_code("EXECUTION_FAILED")
        assert lib.aclblasSetWorkspace(h.value, ctypes.c_void_p(exact.data_ptr()), exact.numel()) == 0
        st, C, info = _matinv(h.value, A, n)
        assert st == 0 and np.all(info == 0)
        torch.cuda.synchronize()
    gC, _ = O.matinv_batched(A)
    assert np.array_equal(C, gC)

"""GPU parity of aclblasSgetrfBatched / SgetrsBatched / SgetriBatched.

The GPU library is called through its C ABI (ctypes), on the same seeded inputs
as the oracle.  The bar: pivots and info bit-exact; LU factors, solutions and
inverses bit-exact against the oracle (the kernels perform the same roundings in
the same order), and within the reference's MIXED_TOLERANCE (test/frame/
verify.h:276-333) of LAPACK, the reference's own golden.  For 64 < n <= 256 the
LU is blocked (LAPACK sgetrf's algorithm, trailing updates summed blockwise):
there pivots and info are bit-exact on tie-free inputs and the factors are held
to MIXED_TOLERANCE against the oracle and LAPACK; on tie-heavy inputs whose LU
is exact in any summation order (integer matrices P L U, see
test_getrf_exact_ties_all_paths) everything is bit-exact on every path.  Case
lists are the reference's CSV tables (tests/golden/cases).
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg.lapack as lapack

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host

pytestmark = pytest.mark.gpu


def _batched_getrf(handle, A, pivot=True, lda=None):
    """A (batch, n, lda) numpy -> (LU, ipiv, info, status) from the GPU."""
    import torch
    batch, n, ld = A.shape
    dA = to_dev(A)
    ptrs = B.ptr_array(dA)
    piv = torch.zeros((batch, n), dtype=torch.int32, device="cuda") if pivot else None
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    st = B.sgetrf_batched(handle.value, n, ptrs.data_ptr(), ld, piv.data_ptr() if pivot else None,
                          info.data_ptr(), batch)
    return to_host(dA), (to_host(piv) if pivot else None), to_host(info), st


def _matrix_kind(row):
    return row["matrix_type"]


def _blocked(n):
    """Orders factored by the blocked kernel."""
    return 64 < n <= 256


def _check_factors(LU, gLU, n, rows):
    """Bit-exact against the oracle, or MIXED_TOLERANCE for the blocked orders."""
    if not _blocked(n):
        assert np.array_equal(LU[rows][:, :, :n], gLU[rows][:, :, :n])
        return
    for b in np.nonzero(rows)[0]:
        ok, detail = O.mixed_tolerance(LU[b, :, :n], gLU[b, :, :n])
        assert ok, "matrix %d: %s" % (b, detail)


# ------------------------------------------------------------------ getrf CSV
GETRF_CASES = [r for r in load_cases("sgetrf_batched.csv")]


@pytest.mark.parametrize("row", GETRF_CASES, ids=[r["case_name"] for r in GETRF_CASES])
def test_getrf_reference_cases(handle, row):
    n, lda, batch = int(row["n"]), int(row["lda"]), int(row["batch_size"])
    pivot = row["pivot_mode"] != "NO_PIVOT"
    expect = status_code(row["expect_result"])
    kind = _matrix_kind(row)
    seed = int(row["random_seed"])
    if expect != 0 or n <= 0 or batch <= 0 or kind.startswith("NULLPTR"):
        pytest.skip("argument-check row: covered by tests/test_abi.py on the CPU and test_status_rows_on_gpu")
    A = O.lapack_batch(n, lda, kind, seed, batch)
    LU, ipiv, info, st = _batched_getrf(handle, A, pivot)
    assert st == expect
    gLU, gpiv, ginfo = O.getrf_batched(A.copy(), pivot=pivot)
    assert np.array_equal(info, ginfo), (info, ginfo)
    if pivot:
        assert np.array_equal(ipiv, gpiv)
    ok_rows = info == 0
    # factors: bit-exact against the oracle (MIXED_TOLERANCE for the blocked orders)
    _check_factors(LU, gLU, n, ok_rows)
    # padding rows (lda > n) untouched
    assert np.array_equal(LU[:, :, n:], A[:, :, n:])
    # and the reference's own criterion against LAPACK (skipping singular batches)
    for b in np.nonzero(ok_rows)[0]:
        if pivot:
            lu, lpiv, linfo = lapack.sgetrf(A[b, :, :n].T)
            assert np.array_equal(lpiv + 1, ipiv[b])
            ok, detail = O.mixed_tolerance(LU[b, :, :n].T, lu)
            assert ok, detail


def test_getrf_c2_config(handle):
    """4096 x 32x32, N(0,1) + 4I (the solve benchmark's inputs): pivots bit-exact, x within 1e-5 of FP64."""
    import torch
    batch, n = 4096, 32
    A = O.normal((batch, n, n), 0, n=n, batch=batch, shift=4.0)
    rhs = O.normal((batch, 1, n), 1)
    LU, ipiv, info, st = _batched_getrf(handle, A)
    assert st == 0 and np.all(info == 0)
    gLU, gpiv, _ = O.getrf_batched(A.copy())
    assert np.array_equal(ipiv, gpiv)
    assert np.array_equal(LU, gLU)
    # solve on the GPU with the GPU factors
    dLU = to_dev(LU)
    dX = to_dev(rhs)
    dpiv = to_dev(ipiv)
    inf = ctypes.c_int(-1)
    pa, px = B.ptr_array(dLU), B.ptr_array(dX)
    st = B.sgetrs_batched(handle.value, B.OP_N, n, 1, pa.data_ptr(), n, dpiv.data_ptr(), px.data_ptr(), n, inf,
                          batch)
    assert st == 0 and inf.value == 0
    X = to_host(dX)
    assert np.array_equal(X, O.getrs_batched(0, gLU, gpiv, rhs))
    # FP64 oracle: normwise backward error of every solve at FP32 level, and
    # forward error within 1e-5 wherever the system allows it (cond_inf <= 1e3,
    # about 3/4 of these systems); elsewhere bounded by cond * u (FP32 LU is
    # backward stable, not exact).
    berr = O.backward_error(A, rhs, X)
    assert berr.max() <= 1e-6, berr.max()
    M = A.transpose(0, 2, 1).astype(np.float64)
    exact = np.linalg.solve(M, rhs.transpose(0, 2, 1).astype(np.float64))
    rel = np.abs(X.transpose(0, 2, 1) - exact).max(axis=(1, 2)) / np.abs(exact).max(axis=(1, 2))
    cond = np.linalg.cond(M, p=np.inf)
    well = cond <= 1e3
    assert well.sum() > batch // 2
    assert rel[well].max() <= 1e-5, rel[well].max()
    assert np.all(rel <= cond * 2.0 ** -24), (rel / cond).max()
    del torch


@pytest.mark.parametrize("n", [1, 7, 33, 64, 65, 100, 128, 129, 200, 256, 300])
def test_getrf_sizes_pivoting(handle, n):
    """Every kernel path (wave, block, global) against the oracle on pivoting-heavy inputs."""
    batch = 6 if n <= 256 else 2
    A = O.normal((batch, n, n), 100 + n)
    LU, ipiv, info, st = _batched_getrf(handle, A)
    assert st == 0
    gLU, gpiv, ginfo = O.getrf_batched(A.copy())
    assert np.array_equal(info, ginfo)
    assert np.array_equal(ipiv, gpiv)
    _check_factors(LU, gLU, n, np.ones(batch, bool))


def _exact_ties_batch(n, batch, seed):
    """A = P L U with L unit lower holding one +-1 below the diagonal in each row, U upper with small integer
    entries (diagonal +-1..+-4), P a random row permutation.  Partial pivoting then meets exact ties in |a| on
    most columns (the lowest current position must win) and every multiplier stays in {0, +-1}: all
    intermediates are small integers, exact in FP32 in any summation order (checked against FP64 on the host
    when the test was written: identical factors and pivots, thousands of tie-driven interchanges)."""
    rng = np.random.default_rng(seed)
    A = np.zeros((batch, n, n), np.float32)
    for b in range(batch):
        L = np.eye(n)
        for i in range(1, n):
            L[i, rng.integers(0, i)] = rng.choice([-1.0, 1.0])
        U = np.triu(rng.integers(-4, 5, size=(n, n)).astype(np.float64))
        U[np.arange(n), np.arange(n)] = rng.integers(1, 5, size=n) * rng.choice([-1, 1], size=n)
        M = (L @ U)[rng.permutation(n)]
        A[b] = M.T.astype(np.float32)               # A[b, j, i] = M[i, j]
    return A


@pytest.mark.parametrize("n", [7, 32, 64, 65, 100, 128, 200, 256, 300])
def test_getrf_exact_ties_all_paths(handle, n):
    """Tie-heavy inputs on every kernel path (n <= 64 wave kernel, 64 < n <= 256 blocked, n > 256 global):
    factors, pivots and info bit-exact against the oracle."""
    batch = 4 if n <= 256 else 2
    A = _exact_ties_batch(n, batch, 500 + n)
    LU, ipiv, info, st = _batched_getrf(handle, A)
    assert st == 0
    gLU, gpiv, ginfo = O.getrf_batched(A.copy())
    assert np.array_equal(info, ginfo)
    assert np.array_equal(ipiv, gpiv)
    assert np.array_equal(LU, gLU)


@pytest.mark.parametrize("n,zero_col", [(100, 70), (129, 0), (200, 191), (256, 255)])
def test_getrf_blocked_zero_pivot(handle, n, zero_col):
    """An exact zero pivot inside a panel of the blocked path: info = column + 1, the pivots up to that column
    as the oracle's, later pivots untouched, and the matrix left in the oracle's stopping state (to FP32
    accuracy)."""
    import torch
    batch = 3
    A = O.normal((batch, n, n), 900 + n, n=n, batch=batch, shift=2.0)
    A[1, zero_col, :] = 0.0                          # column zero_col of matrix 1 is zero
    dA = to_dev(A)
    piv = torch.full((batch, n), -7, dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    pa = B.ptr_array(dA)
    assert B.sgetrf_batched(handle.value, n, pa.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch) == 0
    LU, p, ih = to_host(dA), to_host(piv), to_host(info)
    gLU, gpiv, ginfo = O.getrf_batched(A.copy())
    assert np.array_equal(ih, ginfo) and ginfo[1] == zero_col + 1
    assert np.array_equal(p[1, :zero_col + 1], gpiv[1, :zero_col + 1])
    assert np.all(p[1, zero_col + 1:] == -7)
    for b in range(batch):
        ok, detail = O.mixed_tolerance(LU[b], gLU[b])
        assert ok, "matrix %d: %s" % (b, detail)


def test_getrf_blocked_no_pivot(handle):
    """Without a pivot array the blocked path factors in place without interchanges (diagonally dominant input)."""
    n, batch = 150, 4
    A = O.normal((batch, n, n), 77, n=n, batch=batch, shift=float(n))
    LU, _, info, st = _batched_getrf(handle, A, pivot=False)
    assert st == 0 and not info.any()
    gLU, _, ginfo = O.getrf_batched(A.copy(), pivot=False)
    for b in range(batch):
        ok, detail = O.mixed_tolerance(LU[b], gLU[b])
        assert ok, detail


def test_getrf_zero_pivot_leaves_later_pivots(handle):
    import torch
    n, batch = 40, 4
    A = O.lapack_batch(n, n, "SINGULAR_ZERO_COL", 42, batch)
    dA = to_dev(A)
    piv = torch.full((batch, n), -7, dtype=torch.int32, device="cuda")
    info = torch.zeros(batch, dtype=torch.int32, device="cuda")
    pa = B.ptr_array(dA)
    st = B.sgetrf_batched(handle.value, n, pa.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch)
    assert st == 0
    p = to_host(piv)
    assert to_host(info).tolist() == [1] * batch
    assert np.all(p[:, 0] == 1) and np.all(p[:, 1:] == -7)
    assert np.array_equal(to_host(dA), A)


# ------------------------------------------------------------------ getrs CSV
GETRS_CASES = [r for r in load_cases("getrs_batched.csv")]


def _factor_with_lapack(A, pivot):
    """LU factors as the reference's getrs/getri tests make them (LAPACK sgetrf on the CPU)."""
    batch, n, lda = A.shape
    LU = A.copy()
    ipiv = np.zeros((batch, n), np.int32) if pivot else None
    for b in range(batch):
        if pivot:
            lu, p, _ = lapack.sgetrf(A[b, :, :n].T)
            LU[b, :, :n] = lu.T
            ipiv[b] = p + 1
        else:
            tmp = A[b:b + 1].copy()
            O.getrf_batched(tmp, pivot=False)
            LU[b] = tmp[0]
    return LU, ipiv


@pytest.mark.parametrize("row", GETRS_CASES, ids=[r["case_name"] for r in GETRS_CASES])
def test_getrs_reference_cases(handle, row):
    n, nrhs = int(row["n"]), int(row["nrhs"])
    lda, ldb, batch = int(row["lda"]), int(row["ldb"]), int(row["batch_count"])
    expect = status_code(row["expect_result"])
    kind = row["matrix_type"]
    if expect != 0 or n <= 0 or nrhs <= 0 or batch <= 0 or kind.startswith("NULLPTR"):
        pytest.skip("argument-check row: covered by tests/test_abi.py")
    pivot = row["pivot_mode"] != "NO_PIVOT"
    trans = B.OPS[row["trans"].replace("ACLBLAS_OP_", "")]
    seed = int(row["random_seed"])
    A = O.lapack_batch(n, lda, kind, seed, batch)
    LU, ipiv = _factor_with_lapack(A, pivot)
    Bh = np.stack([O.getrs_rhs(n, nrhs, ldb, seed, b) for b in range(batch)])
    dLU, dB = to_dev(LU), to_dev(Bh)
    dpiv = to_dev(ipiv) if pivot else None
    info = ctypes.c_int(-1)
    use_info = row.get("info_mode", "NORMAL") != "NULLPTR"
    pa, pb = B.ptr_array(dLU), B.ptr_array(dB)
    st = B.sgetrs_batched(handle.value, trans, n, nrhs, pa.data_ptr(), lda,
                          dpiv.data_ptr() if pivot else None, pb.data_ptr(), ldb,
                          info if use_info else None, batch)
    assert st == expect
    if use_info:
        assert info.value == 0
    X = to_host(dB)
    gX = O.getrs_batched(0 if trans == B.OP_N else 1, LU, ipiv, Bh)
    assert np.array_equal(X[:, :, :n], gX[:, :, :n], equal_nan=True)
    assert np.array_equal(X[:, :, n:], Bh[:, :, n:])
    if n <= 1024:
        for b in range(batch):
            ref = lapack.sgetrs(LU[b, :, :n].T, (ipiv[b] - 1) if pivot else np.arange(n),
                                Bh[b, :, :n].T, trans=0 if trans == B.OP_N else 1)[0]
            ok, detail = O.mixed_tolerance(X[b, :, :n].T, ref)
            assert ok or kind.startswith("SINGULAR") or kind == "MIXED", detail


# ------------------------------------------------------------------ getri CSV
GETRI_CASES = [r for r in load_cases("sgetri_batched.csv")]


@pytest.mark.parametrize("row", GETRI_CASES, ids=[r["case_name"] for r in GETRI_CASES])
def test_getri_reference_cases(handle, row):
    import torch
    n, lda, ldc, batch = int(row["n"]), int(row["lda"]), int(row["ldc"]), int(row["batch_size"])
    expect = status_code(row["expect_result"])
    kind = row["matrix_type"]
    if expect != 0 or n <= 0 or batch <= 0 or kind.startswith("NULLPTR"):
        pytest.skip("argument-check row: covered by tests/test_abi.py")
    pivot = row["pivot_mode"] != "NO_PIVOT"
    seed = int(row["random_seed"])
    A = O.lapack_batch(n, lda, kind, seed, batch)
    LU, ipiv = _factor_with_lapack(A, pivot)
    dLU = to_dev(LU)
    C = torch.full((batch, n, ldc), 9.0, dtype=torch.float32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    dpiv = to_dev(ipiv) if pivot else None
    pa, pc = B.ptr_array(dLU), B.ptr_array(C)
    st = B.sgetri_batched(handle.value, n, pa.data_ptr(), lda, dpiv.data_ptr() if pivot else None,
                          pc.data_ptr(), ldc, info.data_ptr(), batch)
    assert st == expect
    Ch, ih = to_host(C), to_host(info)
    gC, ginfo = O.getri_batched(LU, ipiv, ldc=ldc)
    assert np.array_equal(ih, ginfo)
    assert np.array_equal(Ch[:, :, :n], gC[:, :, :n], equal_nan=True)
    assert np.all(Ch[:, :, n:] == 9.0)
    if n <= 512:
        for b in range(batch):
            if ginfo[b] != 0:
                continue
            ref, linfo = lapack.sgetri(LU[b, :, :n].T, (ipiv[b] - 1) if pivot else np.arange(n))
            assert linfo == 0
            ok, detail = O.mixed_tolerance(Ch[b, :, :n].T, ref)
            assert ok, detail


@pytest.mark.parametrize("n", [1, 7, 16, 17, 31, 32, 33, 63, 64, 65, 100])
@pytest.mark.parametrize("pivot", [True, False])
def test_getri_sizes_and_zero_diagonals(handle, n, pivot):
    # n <= 64 runs the register kernel, larger n the LDS kernel: both bit-exact
    # against the oracle; members 1 and 2 get zero diagonal entries (info is the
    # largest such j + 1 and the U solve is skipped)
    import torch
    batch, ldc = 6, n + 3
    A = O.normal((batch, n, n), 40 + n, n=n, batch=batch, shift=4.0)
    LU, ipiv, _ = O.getrf_batched(A.copy(), pivot=pivot)
    LU[1, 0, 0] = 0.0
    LU[2, n // 2, n // 2] = 0.0
    LU[2, n - 1, n - 1] = 0.0
    dLU = to_dev(LU)
    C = torch.full((batch, n, ldc), 9.0, dtype=torch.float32, device="cuda")
    info = torch.full((batch,), -1, dtype=torch.int32, device="cuda")
    dpiv = to_dev(ipiv) if pivot else None
    pa, pc = B.ptr_array(dLU), B.ptr_array(C)
    st = B.sgetri_batched(handle.value, n, pa.data_ptr(), n, dpiv.data_ptr() if pivot else None, pc.data_ptr(),
                          ldc, info.data_ptr(), batch)
    assert st == 0
    Ch, ih = to_host(C), to_host(info)
    gC, ginfo = O.getri_batched(LU, ipiv, ldc=ldc)
    assert ginfo[1] == 1 and ginfo[2] == n and ginfo[0] == 0
    assert np.array_equal(ih, ginfo)
    assert np.array_equal(Ch[:, :, :n], gC[:, :, :n], equal_nan=True)
    assert np.all(Ch[:, :, n:] == 9.0)

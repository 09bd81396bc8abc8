"""Pins the CPU oracle (CPU only).

Every oracle function is checked against something it does not restate:
the reference's known-answer examples, LAPACK (scipy's sgetrf/sgetrs/sgetri,
the library the reference's golden path calls), FP64 linear algebra, the
published MT19937 stream, and invariants (P A = L U, A A^-1 = I).  The checks
are chosen so that a dropped term, a wrong index or a transposed operand in the
oracle fails at least one of them.
"""
import numpy as np
import pytest
import scipy.linalg as sla
import scipy.linalg.lapack as lapack

import oracle as O


def _col(a):
    """(batch, n, ld) oracle layout of a single column-major matrix given as a 2-D row-major array."""
    return np.ascontiguousarray(np.asarray(a, np.float32).T)[None]


# ------------------------------------------------------------------ generators
def test_uniform_generator_matches_mt19937_stream():
    # std::mt19937(42) first outputs (published reference values) and the
    # libstdc++ uniform_real_distribution<float> mapping, rebuilt with numpy's
    # legacy MT19937 (same init_genrand seeding).
    rs = np.random.RandomState(42)
    raw = rs.randint(0, 2 ** 32, size=1000, dtype=np.uint64)
    assert raw[0] == 1608637542
    u = raw.astype(np.float32) / np.float32(2 ** 32)
    u = np.where(u >= 1, np.nextafter(np.float32(1), np.float32(0)), u).astype(np.float32)
    lo, hi = np.float32(-1), np.float32(1)
    expect = (u * (hi - lo) + lo).astype(np.float32)
    assert np.array_equal(O.fill_uniform(1000, 42, -1.0, 1.0), expect)
    # seed 0 maps to 42 in the reference's fill helpers (test/frame/fill.h:497)
    assert np.array_equal(O.fill_uniform(1000, 0, -1.0, 1.0), expect)


def test_lapack_matrix_kinds():
    n = 6
    ident = O.lapack_matrix(n, n, "IDENTITY", 42, 0)
    assert np.array_equal(ident, np.eye(n, dtype=np.float32))
    hil = O.lapack_matrix(4, 4, "ILL_CONDITIONED", 42, 0)
    assert np.allclose(hil, 1.0 / (np.arange(4)[:, None] + np.arange(4)[None, :] + 1))
    rnd = O.lapack_matrix(n, 8, "RANDOM_NONSINGULAR", 42, 3)
    # draws come from mt19937(42 + 3*1000) in column order, n added on the diagonal
    rs = np.random.RandomState(42 + 3000)
    raw = rs.randint(0, 2 ** 32, size=n * n, dtype=np.uint64).astype(np.float32) / np.float32(2 ** 32)
    expect = raw.reshape(n, n).astype(np.float32) + np.float32(n) * np.eye(n, dtype=np.float32)
    assert np.array_equal(rnd[:, :n], expect)
    assert np.all(rnd[:, n:] == 0)
    z = O.lapack_matrix(n, n, "SINGULAR_ZERO_COL", 42, 0)
    assert np.all(z[0] == 0)
    d = O.lapack_matrix(n, n, "SINGULAR_DEPENDENT_ROW", 42, 0)
    assert np.array_equal(d[:, 2], 2 * d[:, 0])
    mixed1 = O.lapack_matrix(n, n, "MIXED", 42, 1)
    assert np.all(mixed1[0] == 0)


# ------------------------------------------------------------------ getrf
def test_getrf_known_answer():
    # reference blas/getrf_batched/README.md:143-144: A = [[4,1],[2,3]] (column-major {4,2,1,3})
    A = np.array([[[4, 2], [1, 3]]], np.float32)
    LU, ipiv, info = O.getrf_batched(A.copy())
    assert info[0] == 0
    assert ipiv.tolist() == [[1, 2]]
    assert LU.reshape(-1).tolist() == [4.0, 0.5, 1.0, 2.5]


def _c2_batch(batch, n=32, seed=0):
    return O.normal((batch, n, n), seed, n=n, batch=batch, shift=4.0)


@pytest.mark.parametrize("n,batch", [(32, 64), (8, 32), (128, 4)])
def test_getrf_pivots_match_lapack(n, batch):
    A = _c2_batch(batch, n)
    LU, ipiv, info = O.getrf_batched(A.copy())
    assert np.all(info == 0)
    for b in range(batch):
        lu, piv, inf = lapack.sgetrf(A[b].T)  # scipy takes a 2-D row-major view of the matrix
        assert inf == 0
        assert np.array_equal(piv + 1, ipiv[b]), "pivot order differs from LAPACK"
        ok, detail = O.mixed_tolerance(LU[b].T, lu)
        assert ok, detail


@pytest.mark.parametrize("n", [5, 32, 100])
def test_getrf_reconstructs(n):
    A = _c2_batch(3, n, seed=7)
    LU, ipiv, info = O.getrf_batched(A.copy())
    for b in range(3):
        M = LU[b].T.astype(np.float64)
        L = np.tril(M, -1) + np.eye(n)
        U = np.triu(M)
        PA = A[b].T.astype(np.float64).copy()
        for k in range(n):
            p = ipiv[b, k] - 1
            PA[[k, p]] = PA[[p, k]]
        err = np.abs(L @ U - PA).max() / np.abs(A[b]).max()
        assert err < 8 * n * 2.0 ** -24
        assert np.abs(np.tril(M, -1)).max() <= 1.0


def test_getrf_zero_pivot_rules():
    n = 8
    A = np.stack([O.lapack_matrix(n, n, "SINGULAR_ZERO_COL", 42, b) for b in range(2)])
    ipiv = np.full((2, n), -7, np.int32)
    info = np.zeros(2, np.int32)
    LU = A.copy()
    O.lib().orc_getrf_batched(LU.reshape(-1), n * n, n, n, ipiv.ctypes.data_as(O.ctypes.c_void_p),
                              info.ctypes.data_as(O.ctypes.c_void_p), 2, 1)
    # reference sgetrf_batched_kernel.cpp:128-137: piv[col] written, info = col+1, stop
    assert info.tolist() == [1, 1]
    assert np.all(ipiv[:, 0] == 1) and np.all(ipiv[:, 1:] == -7)
    assert np.array_equal(LU, A)  # nothing else modified
    # LAPACK agrees on info and on the first pivot
    _, piv, inf = lapack.sgetrf(A[0].T)
    assert inf == 1 and piv[0] == 0
    # dependent row: exactly singular, last pivot zero
    D = np.stack([O.lapack_matrix(32, 32, "SINGULAR_DEPENDENT_ROW", 42, b) for b in range(4)])
    _, _, info = O.getrf_batched(D.copy())
    for b in range(4):
        assert info[b] == lapack.sgetrf(D[b].T)[2]


def test_getrf_no_pivot():
    A = np.stack([O.lapack_matrix(16, 16, "DIAGONALLY_DOMINANT", 42, b) for b in range(3)])
    LU, ipiv, info = O.getrf_batched(A.copy(), pivot=False)
    assert ipiv is None and np.all(info == 0)
    for b in range(3):
        M = LU[b].T.astype(np.float64)
        L = np.tril(M, -1) + np.eye(16)
        assert np.abs(L @ np.triu(M) - A[b].T).max() < 1e-4


def test_getrf_planted_mistake_is_caught():
    # A wrong multiplier index (a plausible slip) must break the reconstruction check.
    n = 16
    A = _c2_batch(1, n, seed=3)
    LU, ipiv, _ = O.getrf_batched(A.copy())
    bad = LU.copy()
    bad[0, 0, 5] = bad[0, 0, 6]  # l_50 replaced by l_60
    M = bad[0].T.astype(np.float64)
    PA = A[0].T.astype(np.float64).copy()
    for k in range(n):
        p = ipiv[0, k] - 1
        PA[[k, p]] = PA[[p, k]]
    L = np.tril(M, -1) + np.eye(n)
    assert np.abs(L @ np.triu(M) - PA).max() > 1e-3


# ------------------------------------------------------------------ getrs / getri
@pytest.mark.parametrize("trans", [0, 1])
def test_getrs_matches_lapack_and_fp64(trans):
    n, nrhs, batch = 48, 5, 6
    A = _c2_batch(batch, n, seed=11)
    B = O.normal((batch, nrhs, n), 12)
    LU, ipiv, _ = O.getrf_batched(A.copy())
    X = O.getrs_batched(trans, LU, ipiv, B)
    for b in range(batch):
        ref = lapack.sgetrs(LU[b].T, ipiv[b] - 1, B[b].T, trans=trans)[0]
        ok, detail = O.mixed_tolerance(X[b].T, ref)
        assert ok, detail
        M = A[b].T.astype(np.float64)
        exact = np.linalg.solve(M.T if trans else M, B[b].T.astype(np.float64))
        assert np.abs(X[b].T - exact).max() / np.abs(exact).max() < 1e-4
    assert O.backward_error(A, B, X).max() < 1e-6 if trans == 0 else True


def test_getrs_identity_and_invalid_pivots():
    A = np.eye(2, dtype=np.float32)[None]
    B = np.array([[[3, 7]]], np.float32)
    # ipiv (1, 2): no swap; (0, 1) is the reference README's example, whose 0
    # entry is out of range and skipped, and 1 swaps rows 1 and 0.
    assert O.getrs_batched(0, A, np.array([[1, 2]], np.int32), B).reshape(-1).tolist() == [3, 7]
    assert O.getrs_batched(0, A, np.array([[0, 1]], np.int32), B).reshape(-1).tolist() == [7, 3]
    assert O.getrs_batched(0, A, None, B).reshape(-1).tolist() == [3, 7]


def test_getri_matches_lapack():
    n, batch = 40, 3
    A = _c2_batch(batch, n, seed=21)
    LU, ipiv, _ = O.getrf_batched(A.copy())
    C, info = O.getri_batched(LU, ipiv)
    assert np.all(info == 0)
    for b in range(batch):
        ref = lapack.sgetri(LU[b].T, ipiv[b] - 1)[0]
        ok, detail = O.mixed_tolerance(C[b].T, ref)
        assert ok, detail
        assert np.abs(C[b].T.astype(np.float64) @ A[b].T - np.eye(n)).max() < 1e-4


def test_getri_info_is_largest_zero_diagonal():
    n = 10
    LU = np.eye(n, dtype=np.float32)[None].copy()
    LU[0, 3, 3] = 0.0
    LU[0, 7, 7] = 0.0
    C, info = O.getri_batched(LU, None)
    # reference getri_batched_kernel.cpp:142-151 scans j = n-1 down to 0
    assert info[0] == 8
    # with a zero diagonal the backward substitution is skipped: C = L^-1 P = I here
    assert np.array_equal(C[0], np.eye(n, dtype=np.float32))


# ------------------------------------------------------------------ matinv
def test_matinv_readme_example_closed_form():
    # reference blas/matinv_batched/README.md example: a_ii = 2, a_ij = 0.1 (n 16,
    # batch 4), i.e. A = 1.9 I + 0.1 J; Sherman-Morrison gives
    # A^-1 = I / 1.9 - 0.1 / (1.9 (1.9 + 1.6)) J
    n, batch = 16, 4
    A = np.full((batch, n, n), 0.1, np.float32)
    for b in range(batch):
        np.fill_diagonal(A[b], 2.0)
    C, info = O.matinv_batched(A)
    assert np.all(info == 0)
    exact = np.eye(n) / 1.9 - 0.1 / (1.9 * 3.5) * np.ones((n, n))
    bound = 4 * n * 2.0 ** -24 * np.abs(exact).max()   # cond(A) = 3.5 / 1.9 < 2
    for b in range(batch):
        assert np.abs(C[b].T - exact).max() < bound


def test_matinv_matches_fp64_inverse_and_keeps_padding_layout():
    rng = np.random.default_rng(5)
    for n, lda, ldi in ((1, 1, 1), (5, 9, 7), (32, 32, 40)):
        A = np.zeros((3, n, lda), np.float32)
        A[:, :, :n] = rng.standard_normal((3, n, n)) + 4 * np.eye(n)
        C, info = O.matinv_batched(A, lda_inv=ldi)
        assert C.shape == (3, n, ldi) and np.all(info == 0)
        for b in range(3):
            ref = np.linalg.inv(A[b, :, :n].T.astype(np.float64))
            ok, detail = O.mixed_tolerance(C[b, :, :n].T, ref)
            assert ok, detail
            # the composition is literally getrf then getri
            LU, piv, _ = O.getrf_batched(A[b:b + 1, :, :n].copy())
            C2, _ = O.getri_batched(LU, piv, ldc=ldi)
            assert np.array_equal(C[b], C2[0])


def test_matinv_info_is_the_inverse_rule():
    # column 1 is zero: the LU stops at k = 1 (its info would be 2, LAPACK's
    # too); the inverse then finds u_33 = 0 in the unfactored part and reports
    # the largest zero diagonal, 4 -- the reference's composition returns the
    # inverse's info (matinv_batched_host.cpp:157-186)
    A = np.zeros((1, 4, 4), np.float32)
    A[0, 0, 0] = 1.0
    A[0, 2, 2] = 1.0
    C, info = O.matinv_batched(A)
    assert info[0] == 4
    _, _, finfo = O.getrf_batched(A.copy())
    assert finfo[0] == 2
    # zero column only: both rules give its index
    A2 = (np.eye(6, dtype=np.float32) * 3 + 0.5)[None].copy()
    A2[0, 2, :] = 0.0
    _, info2 = O.matinv_batched(A2)
    assert info2[0] == 3 == lapack.sgetrf(A2[0].T)[2]


# ------------------------------------------------------------------ GEMM
def test_gemm_known_answer():
    # reference blas/gemm_ex/README.md:184-188
    A = np.array([[1, 2, 3], [4, 5, 6]], np.float32)       # 2x3
    B = np.array([[1, 0], [0, 1], [1, 1]], np.float32)    # 3x2
    Ac = np.ascontiguousarray(A.T)
    Bc = np.ascontiguousarray(B.T)
    C = np.zeros((2, 2), np.float32)
    O.gemm(0, 0, 2, 2, 3, 1.0, Ac, O.DT_F32, 2, Bc, O.DT_F32, 3, 0.0, C, O.DT_F32, 2)
    assert C.T.tolist() == [[4, 5], [10, 11]]


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_gemm_matches_fp64_matmul(ta, tb):
    m, n, k = 17, 19, 23
    rs = np.random.RandomState(5)
    opA = rs.uniform(-1, 1, (m, k)).astype(np.float32)
    opB = rs.uniform(-1, 1, (k, n)).astype(np.float32)
    C0 = rs.uniform(-1, 1, (m, n)).astype(np.float32)
    # column-major storage of A (m x k or, when transposed, k x m)
    Ast = np.ascontiguousarray((opA.T if ta else opA).T)
    Bst = np.ascontiguousarray((opB.T if tb else opB).T)
    lda = k if ta else m
    ldb = n if tb else k
    C = np.ascontiguousarray(C0.T).copy()
    O.gemm(ta, tb, m, n, k, 0.75, Ast, O.DT_F32, lda, Bst, O.DT_F32, ldb, -0.5, C, O.DT_F32, m)
    expect = 0.75 * (opA.astype(np.float64) @ opB) - 0.5 * C0
    assert np.abs(C.T - expect).max() < 1e-6


def test_dtype_conversions_match_torch():
    import torch
    x = np.concatenate([np.random.RandomState(1).normal(0, 3, 4000).astype(np.float32),
                        np.array([0.0, -0.0, 1e-8, -1e-8, 65504.0, 3.0e38], np.float32)])
    bf = O.to_bf16(x)
    tb = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(bf, tb)
    assert np.array_equal(O.from_bf16(bf), torch.from_numpy(x).to(torch.bfloat16).float().numpy())
    h = O.to_f16(x)
    assert np.array_equal(h, x.astype(np.float16).view(np.uint16))
    assert np.array_equal(O.from_f16(h), x.astype(np.float16).astype(np.float32))
    inr = np.clip(x, -440, 440)
    for e4m3, tdt in ((True, torch.float8_e4m3fn), (False, torch.float8_e5m2)):
        q = O.to_fp8(inr, e4m3)
        tq = torch.from_numpy(inr).to(tdt).view(torch.uint8).numpy()
        assert np.array_equal(q, tq)
        assert np.array_equal(O.from_fp8(q, e4m3), torch.from_numpy(inr).to(tdt).float().numpy())
    # saturation (reference test/frame/dtype_utils.h): beyond max -> max
    assert O.from_fp8(O.to_fp8(np.array([1e4, -1e4], np.float32), True), True).tolist() == [448.0, -448.0]


def test_backward_error_oracle():
    A = np.eye(3, dtype=np.float32)[None]
    B = np.array([[[1, 2, 3]]], np.float32)
    assert O.backward_error(A, B, B)[0] == 0.0
    X = B.copy()
    X[0, 0, 2] = 3.5   # residual 0.5, ||A||=1, ||x||=3.5, ||b||=3
    assert abs(O.backward_error(A, B, X)[0] - 0.5 / (3.5 + 3.0)) < 1e-12


# ------------------------------------------------------------------ triangular solve (aclblasStrsm golden)
def _tri(m, upper, seed, boost=True):
    rng = np.random.default_rng(seed)
    A = rng.uniform(-1, 1, (m, m)).astype(np.float32)
    A = np.triu(A) if upper else np.tril(A)
    if boost:
        d = np.diag(A).copy()
        A[np.arange(m), np.arange(m)] = d + np.where(d >= 0, 1, -1) * max(5.0, m)
    return A


@pytest.mark.parametrize("side", [0, 1])
@pytest.mark.parametrize("upper", [0, 1])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("unit", [0, 1])
def test_oracle_trsm_matches_fp64_and_blas(side, upper, trans, unit):
    """orc_trsm against an FP64 triangular solve (FP32-level agreement) and scipy's BLAS strsm (another FP32
    summation order), every side / uplo / trans / diag case, alpha != 1."""
    import scipy.linalg as sl
    import scipy.linalg.blas as sb
    m, n = 37, 29
    k = m if side == 0 else n
    A = _tri(k, upper, 10 + 8 * side + 4 * upper + 2 * trans + unit)
    lda = k + 3
    Ap = np.zeros((k, lda), np.float32)          # column-major with padding: Ap[j, i] = A(i, j)
    Ap[:, :k] = A.T
    rng = np.random.default_rng(99)
    Bm = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    ldb = m + 2
    Bp = np.zeros((n, ldb), np.float32)
    Bp[:, :m] = Bm.T
    alpha = 0.75
    out = Bp.copy().reshape(-1)
    O.trsm(side, upper, trans, unit, m, n, alpha, Ap.reshape(-1), lda, out, ldb)
    X = out.reshape(n, ldb)[:, :m].T.astype(np.float64)
    assert np.array_equal(out.reshape(n, ldb)[:, m:], Bp[:, m:])          # padding untouched
    A64 = A.astype(np.float64)
    if unit:
        A64[np.arange(k), np.arange(k)] = 1.0
    opA = A64.T if trans else A64
    ref = np.linalg.solve(opA, alpha * Bm.astype(np.float64)) if side == 0 else \
        np.linalg.solve(opA.T, alpha * Bm.astype(np.float64).T).T
    assert np.abs(X - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())
    blas = sb.strsm(alpha, A, Bm, side=side, lower=1 - upper, trans_a=trans, diag=unit)
    assert np.abs(X - blas).max() <= 1e-5 * max(1.0, np.abs(blas).max())
    # the check separates the cases: the solution with the operand transposed is far off
    X2 = np.linalg.solve(opA.T, alpha * Bm.astype(np.float64)) if side == 0 else \
        np.linalg.solve(opA, alpha * Bm.astype(np.float64).T).T
    assert np.abs(X - X2).max() > 1e-3


def test_oracle_trsm_known_answers():
    """Identity A: B scaled by alpha exactly; diagonal A (right side, N): column j divided through the reciprocal
    of a_jj as strsm.f does; alpha = 0: B zeroed, NaN included."""
    m, n = 5, 4
    B0 = np.arange(1, m * n + 1, dtype=np.float32)
    I = np.eye(m, dtype=np.float32).reshape(-1)
    for side, k in ((0, m), (1, n)):
        Ik = np.eye(k, dtype=np.float32).reshape(-1)
        for upper in (0, 1):
            for trans in (0, 1):
                out = B0.copy()
                O.trsm(side, upper, trans, 0, m, n, 2.0, Ik, k, out, m)
                assert np.array_equal(out, 2 * B0)
    d = np.array([2.0, 4.0, 8.0, 3.0], np.float32)
    out = B0.copy()
    O.trsm(1, 1, 0, 0, m, n, 1.0, np.diag(d).T.reshape(-1).copy(), n, out, m)
    expect = (B0.reshape(n, m) * (np.float32(1.0) / d)[:, None]).reshape(-1)
    assert np.array_equal(out, expect)
    out = B0.copy()
    out[3] = np.nan
    O.trsm(0, 0, 0, 0, m, n, 0.0, I, m, out, m)
    assert not out.any()

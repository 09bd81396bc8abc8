"""The adjoint of X = A^-1 B on the CPU: the oracle's composition
(transposed solve with the LU factors, then dA = -dB X^T) against FP64
central finite differences of L(A, B) = <dX, A^-1 B>, and a planted mistake
(the untransposed solve) that the check must reject.
"""
import numpy as np

import oracle as O


def _fd_grad(A, Bm, dX, i, j, h):
    """dL/dA_ij by central differences in FP64; A, Bm, dX are row-major (n x n, n x r, n x r)."""
    E = np.zeros_like(A)
    E[i, j] = h
    lp = np.sum(dX * np.linalg.solve(A + E, Bm))
    lm = np.sum(dX * np.linalg.solve(A - E, Bm))
    return (lp - lm) / (2 * h)


def _setup(n, nrhs, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n)) + 8 * np.eye(n)
    Bm = rng.standard_normal((n, nrhs))
    dX = rng.standard_normal((n, nrhs))
    A32, B32, dX32 = A.astype(np.float32), Bm.astype(np.float32), dX.astype(np.float32)
    LU, piv, info = O.getrf_batched(A32.T.copy()[None])         # column-major storage [b, j, i]
    assert info[0] == 0
    X = O.getrs_batched(0, LU, piv, B32.T.copy()[None])         # (1, nrhs, n)
    return A32.astype(np.float64), B32.astype(np.float64), dX32.astype(np.float64), LU, piv, X, dX32


def test_adjoint_matches_finite_differences():
    n, nrhs = 24, 3
    A, Bm, dX, LU, piv, X, dX32 = _setup(n, nrhs, 0)
    dB, dA = O.getrs_adjoint(LU, piv, X, dX32.T.copy()[None])
    dA_rm = dA[0].T.astype(np.float64)              # row-major dL/dA
    dB_rm = dB[0].T.astype(np.float64)              # row-major dL/dB (n x nrhs)
    # dL/dB = A^-T dX exactly (FP64 reference)
    ref_dB = np.linalg.solve(A.T, dX)
    assert np.abs(dB_rm - ref_dB).max() <= 1e-5 * np.abs(ref_dB).max()
    h = 1e-6 * np.abs(A).max()
    rng = np.random.default_rng(1)
    for _ in range(12):
        i, j = rng.integers(0, n, 2)
        fd = _fd_grad(A, Bm, dX, i, j, h)
        assert abs(dA_rm[i, j] - fd) <= 1e-4 * max(1.0, abs(fd)), (i, j, dA_rm[i, j], fd)


def test_adjoint_planted_mistake_is_caught():
    n, nrhs = 16, 2
    A, Bm, dX, LU, piv, X, dX32 = _setup(n, nrhs, 3)
    wrong = O.getrs_batched(0, LU, piv, dX32.T.copy()[None])[0].T.astype(np.float64)  # A^-1 dX, not A^-T dX
    ref_dB = np.linalg.solve(A.T, dX)
    assert np.abs(wrong - ref_dB).max() > 1e-3 * np.abs(ref_dB).max()

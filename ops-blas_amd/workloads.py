"""Synthetic inputs of the benchmark workloads (numpy, host memory).

Batches of column-major matrices follow the layout used throughout: a batch of
n x n matrices with leading dimension n is an array of shape (batch, n, n)
whose [b, j, i] element is entry (i, j) of matrix b.
"""
import numpy as np


def lu_batch(batch, n, nrhs, seed=0, shift=4.0):
    """Fixed-order systems: A ~ N(0,1) + shift*I, B ~ N(0,1).  Returns (A (batch, n, n), B (batch, nrhs, n))."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((batch, n, n), dtype=np.float32)
    idx = np.arange(n)
    A[:, idx, idx] += np.float32(shift)
    B = rng.standard_normal((batch, nrhs, n), dtype=np.float32)
    return A, B


def conditioned_systems(count, seed=0, nmin=16, nmax=256, rmin=1, rmax=16, log_cond=(2.0, 6.0)):
    """Mixed-order systems A = U diag(s) V^T and B ~ N(0,1): for every system its own U and V, the Q
    factors of the QR of two Gaussian n x n matrices (FP64); s_i = 10^(-c i / (n-1)) with
    c = log10(cond) uniform in log_cond; nrhs uniform in [rmin, rmax].

    Returns a list of dicts {n, nrhs, cond, A (n*n float32, column-major), B (n*nrhs float32)}
    in generation order.  The QR factorisations of one order run as one stacked numpy call.
    """
    rng = np.random.default_rng(seed)
    ns = rng.integers(nmin, nmax + 1, size=count)
    rs = rng.integers(rmin, rmax + 1, size=count)
    cs = rng.uniform(log_cond[0], log_cond[1], size=count)
    out = [None] * count
    for n in np.unique(ns):
        members = np.nonzero(ns == n)[0]
        n = int(n)
        k = len(members)
        U, _ = np.linalg.qr(rng.standard_normal((k, n, n)))
        V, _ = np.linalg.qr(rng.standard_normal((k, n, n)))
        expo = np.arange(n) / max(n - 1, 1)
        s = 10.0 ** (-cs[members][:, None] * expo[None, :])            # (k, n)
        A = (U * s[:, None, :]) @ np.swapaxes(V, 1, 2)                 # row-major (k, n, n)
        for t, idx in enumerate(members):
            r = int(rs[idx])
            Bm = rng.standard_normal((r, n)).astype(np.float32)          # column-major n x r
            out[idx] = {"n": n, "nrhs": r, "cond": float(10.0 ** cs[idx]),
                        "A": np.ascontiguousarray(A[t].T, dtype=np.float32).reshape(-1),
                        "B": Bm.reshape(-1)}
    return out


def normwise_backward_error(A_cm, n, B_cm, X_cm, nrhs):
    """max over right-hand sides of |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf), in FP64.

    A_cm, B_cm, X_cm are flat column-major float arrays (lda = ldb = ldx = n)."""
    A = np.asarray(A_cm, dtype=np.float64).reshape(n, n).T
    Bm = np.asarray(B_cm, dtype=np.float64).reshape(nrhs, n).T
    Xm = np.asarray(X_cm, dtype=np.float64).reshape(nrhs, n).T
    R = Bm - A @ Xm
    an = np.abs(A).sum(axis=1).max()
    num = np.abs(R).max(axis=0)
    den = an * np.abs(Xm).max(axis=0) + np.abs(Bm).max(axis=0)
    with np.errstate(divide="ignore", invalid="ignore"):
        be = np.where(den > 0, num / den, np.where(num > 0, np.inf, 0.0))
    return float(be.max()) if be.size else 0.0

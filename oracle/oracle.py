"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's CPU-baseline leg, always as the checker or the timed CPU baseline,
never as a substitute for the GPU library.

All arrays are numpy arrays; matrices are column-major.  A batch of column-major
n x n matrices with leading dimension lda is stored as a numpy array of shape
(batch, n, lda) whose [b, j, i] element is entry (i, j) of matrix b.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIBRARY selects another build of oracle.cpp (the sanitizer build in tests)
_LIB_PATH = os.environ.get("ORACLE_LIBRARY") or os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u16p = np.ctypeslib.ndpointer(dtype=np.uint16, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")

# LAPACK matrix kinds of the reference's makeBlasLapackMatrix (test/frame/fill.h:586-599)
MATRIX_TYPES = {
    "RANDOM_NONSINGULAR": 0, "IDENTITY": 1, "DIAGONALLY_DOMINANT": 2, "ILL_CONDITIONED": 3,
    "DIAGONAL": 4, "UPPER_TRIANGULAR": 5, "LOWER_TRIANGULAR": 6, "SINGULAR_ZERO_COL": 7,
    "SINGULAR_DEPENDENT_ROW": 8, "MIXED": 9,
}

DT_F32, DT_F16, DT_BF16, DT_E5M2, DT_E4M3 = 0, 1, 27, 35, 36


def build():
    subprocess.check_call(["make", "-s", "-B", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c = ctypes
        L.orc_fill_uniform.argtypes = [_f32p, c.c_size_t, c.c_uint32, c.c_float, c.c_float, c.c_int]
        L.orc_lapack_matrix.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_uint32, c.c_int]
        L.orc_getrs_rhs.argtypes = [_f32p, c.c_int, c.c_int, c.c_int, c.c_uint32, c.c_int]
        L.orc_fill_normal.argtypes = [_f32p, c.c_size_t, c.c_uint64, c.c_int, c.c_int, c.c_float]
        L.orc_getrf_batched.argtypes = [_f32p, c.c_longlong, c.c_int, c.c_int, c.c_void_p, c.c_void_p, c.c_int,
                                        c.c_int]
        L.orc_getrs_batched.argtypes = [c.c_int, c.c_int, c.c_int, _f32p, c.c_longlong, c.c_int, c.c_void_p,
                                        _f32p, c.c_longlong, c.c_int, c.c_int, c.c_int]
        L.orc_getri_batched.argtypes = [c.c_int, _f32p, c.c_longlong, c.c_int, c.c_void_p, _f32p, c.c_longlong,
                                        c.c_int, _i32p, c.c_int, c.c_int]
        L.orc_time_getrf_getrs.argtypes = [_f32p, _f32p, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int]
        L.orc_time_getrf_getrs.restype = c.c_double
        L.orc_trsm.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_float, _f32p, c.c_int,
                               _f32p, c.c_int]
        L.orc_time_getrf_getrs_mixed.argtypes = [_f32p, _i64p, _f32p, _i64p, _i32p, _i32p, c.c_int, c.c_int,
                                                 c.c_int]
        L.orc_time_getrf_getrs_mixed.restype = c.c_double
        L.orc_hardware_threads.restype = c.c_int
        L.orc_gemm.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_double, c.c_void_p, c.c_int,
                               c.c_int, c.c_void_p, c.c_int, c.c_int, c.c_double, c.c_void_p, c.c_int, c.c_int,
                               c.c_double, c.c_int]
        L.orc_lt_matmul.argtypes = [c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_int, c.c_float,
                                    c.c_void_p, c.c_longlong, c.c_void_p, c.c_longlong, c.c_void_p, c.c_void_p,
                                    c.c_float, c.c_void_p, c.c_int, _f32p, c.c_int, c.c_int]
        L.orc_f32_to_bf16.argtypes = [_f32p, _u16p, c.c_size_t]
        L.orc_bf16_to_f32.argtypes = [_u16p, _f32p, c.c_size_t]
        L.orc_f32_to_f16.argtypes = [_f32p, _u16p, c.c_size_t]
        L.orc_f16_to_f32.argtypes = [_u16p, _f32p, c.c_size_t]
        L.orc_f32_to_fp8.argtypes = [_f32p, _u8p, c.c_size_t, c.c_int]
        L.orc_fp8_to_f32.argtypes = [_u8p, _f32p, c.c_size_t, c.c_int]
        L.orc_backward_error.argtypes = [_f32p, _f32p, _f32p, c.c_int, c.c_int, c.c_int, _f64p, c.c_int]
        _lib = L
    return _lib


def threads():
    return max(1, int(lib().orc_hardware_threads()))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# ---------------------------------------------------------------- generators
def fill_uniform(count, seed, lo, hi, map_zero=True):
    out = np.empty(count, dtype=np.float32)
    lib().orc_fill_uniform(out, count, seed, lo, hi, int(map_zero))
    return out


def blas_matrix(m, n, ld, seed, lo=-1.0, hi=1.0):
    """makeBlasMatrix(m, n, ld, "RANDOM_v", seed): ld*n draws, column-major."""
    return fill_uniform(ld * n, seed, lo, hi).reshape(n, ld)


def lapack_matrix(n, lda, kind, seed, batch_idx):
    out = np.empty(lda * max(1, n), dtype=np.float32)
    t = MATRIX_TYPES[kind] if isinstance(kind, str) else int(kind)
    lib().orc_lapack_matrix(out, n, lda, t, seed, batch_idx)
    return out.reshape(max(1, n), lda)


def lapack_batch(n, lda, kind, seed, batch):
    return np.stack([lapack_matrix(n, lda, kind, seed, b) for b in range(batch)]) if batch > 0 else \
        np.zeros((0, max(1, n), lda), np.float32)


def getrs_rhs(n, nrhs, ldb, seed, batch_idx):
    out = np.empty(ldb * max(1, nrhs), dtype=np.float32)
    lib().orc_getrs_rhs(out, n, nrhs, ldb, seed, batch_idx)
    return out.reshape(max(1, nrhs), ldb)


def normal(shape, seed, n=0, batch=0, shift=0.0):
    out = np.empty(int(np.prod(shape)), dtype=np.float32)
    lib().orc_fill_normal(out, out.size, seed, n, batch, shift)
    return out.reshape(shape)


# ---------------------------------------------------------------- LAPACK
def getrf_batched(A, pivot=True, nthreads=None):
    """In-place batched LU of A (batch, n, lda).  Returns (ipiv (batch, n) or None, info (batch,))."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    batch, n, lda = A.shape
    ipiv = np.zeros((batch, n), dtype=np.int32) if pivot else None
    info = np.zeros(batch, dtype=np.int32)
    lib().orc_getrf_batched(A.reshape(-1), n * lda, n, lda, _ptr(ipiv), _ptr(info), batch,
                            nthreads or threads())
    return A, ipiv, info


def getrs_batched(trans, LU, ipiv, B, nthreads=None):
    """Solve op(A) X = B in place; LU (batch, n, lda), B (batch, nrhs, ldb)."""
    LU = np.ascontiguousarray(LU, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32).copy()
    batch, n, lda = LU.shape
    _, nrhs, ldb = B.shape
    ipiv = None if ipiv is None else np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_getrs_batched(int(trans), n, nrhs, LU.reshape(-1), n * lda, lda, _ptr(ipiv), B.reshape(-1),
                            nrhs * ldb, ldb, batch, nthreads or threads())
    return B


def getrs_adjoint(LU, ipiv, X, dX, nthreads=None):
    """Reverse-mode derivative of X = A^-1 B from the LU factors of A.

    dB = A^-T dX by the transposed solve (getrs_batched(1, ...)), then
    dA = -dB X^T accumulated in FP64 and rounded once to FP32.  Shapes as in
    getrs_batched: LU (batch, n, lda), X and dX (batch, nrhs, ld).  Returns
    (dB like dX, dA (batch, n, n) column-major).
    """
    dB = getrs_batched(1, LU, ipiv, dX, nthreads)
    n = LU.shape[1]
    Xm = np.asarray(X, dtype=np.float64)[:, :, :n]      # (batch, nrhs, n): [b, c, i] = X_ic
    dBm = dB.astype(np.float64)[:, :, :n]
    # dA_ij = -sum_c dB_ic X_jc; column-major storage [b, j, i]
    dA = -np.einsum("bcj,bci->bji", Xm, dBm)
    return dB, dA.astype(np.float32)


def getri_batched(LU, ipiv, ldc=None, nthreads=None):
    LU = np.ascontiguousarray(LU, dtype=np.float32)
    batch, n, lda = LU.shape
    ldc = ldc or n
    C = np.zeros((batch, n, ldc), dtype=np.float32)
    info = np.zeros(batch, dtype=np.int32)
    ipiv = None if ipiv is None else np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_getri_batched(n, LU.reshape(-1), n * lda, lda, _ptr(ipiv), C.reshape(-1), n * ldc, ldc, info,
                            batch, nthreads or threads())
    return C, info


def matinv_batched(A, lda_inv=None, nthreads=None):
    """Ainv = A^-1 as the reference composes it (blas/matinv_batched/arch35/
    matinv_batched_host.cpp:165-195): pivoted getrf on a copy of A, then getri
    from those factors; the returned info is getri's, which replaces getrf's.
    Pivots after a zero pivot stay 0 (no interchange).  A (batch, n, lda);
    returns (Ainv (batch, n, lda_inv), info (batch,))."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    n = A.shape[1]
    LU, ipiv, _ = getrf_batched(np.ascontiguousarray(A[:, :, :n]).copy(), pivot=True, nthreads=nthreads)
    return getri_batched(LU, ipiv, ldc=lda_inv or n, nthreads=nthreads)


def time_getrf_getrs(A, B, nthreads, reps=3):
    """Median seconds of one getrf+getrs pass over A (batch, n, n), B (batch, nrhs, n)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    batch, n, _ = A.shape
    nrhs = B.shape[1]
    return float(lib().orc_time_getrf_getrs(A.reshape(-1), B.reshape(-1), n, nrhs, batch, nthreads, reps))


def time_getrf_getrs_mixed(A, offA, B, offB, n, nrhs, nthreads, reps=3):
    """Median seconds of one getrf+getrs pass over mixed-order systems packed in the flat float32 arrays A and
    B (system i at offA[i] / offB[i], column-major, leading dimension n[i])."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    offA = np.ascontiguousarray(offA, dtype=np.int64)
    offB = np.ascontiguousarray(offB, dtype=np.int64)
    n = np.ascontiguousarray(n, dtype=np.int32)
    nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
    return float(lib().orc_time_getrf_getrs_mixed(A, offA, B, offB, n, nrhs, len(n), nthreads, reps))


def trsm(side, upper, trans, unit, m, n, alpha, A, lda, B, ldb):
    """B (numpy float32, modified in place) := alpha op(A)^-1 B (side 0) or alpha B op(A)^-1 (side 1), netlib
    strsm loop order in FP32."""
    lib().orc_trsm(int(side), int(upper), int(trans), int(unit), m, n, float(alpha),
                   np.ascontiguousarray(A, dtype=np.float32), lda, B, ldb)
    return B


# ---------------------------------------------------------------- GEMM
def gemm(transA, transB, m, n, k, alpha, A, Atype, lda, B, Btype, ldb, beta, C, Ctype, ldc, scale_ab=1.0,
         nthreads=None):
    """C (numpy, modified in place) = alpha*op(A)op(B)*scale_ab + beta*C with FP64 accumulation."""
    lib().orc_gemm(int(transA), int(transB), m, n, k, float(alpha), _ptr(A), Atype, lda, _ptr(B), Btype, ldb,
                   float(beta), _ptr(C), Ctype, ldc, float(scale_ab), nthreads or threads())
    return C


LT_FP32, LT_MXFP8, LT_MXFP4 = 0, 36, 40


def lt_matmul(type_, bf16_out, transA, transB, m, n, k, alpha, A, lda, B, ldb, scaleA, scaleB, beta, C, ldc, ldd,
              nthreads=None):
    """Row-major BLASLt golden: returns D (m x ldd float32; BF16 results truncated to BF16, as the reference's
    golden).  A, B: numpy arrays of the stored operands (float32 for FP32, uint8 bytes for MX); scaleA/B: uint8 E8M0
    bytes (MX) or None; C: float32 (m x ldc) or None."""
    D = np.zeros((m, ldd), dtype=np.float32)
    lib().orc_lt_matmul(int(type_), int(bool(bf16_out)), int(transA), int(transB), m, n, k, float(alpha),
                        _ptr(np.ascontiguousarray(A)), lda, _ptr(np.ascontiguousarray(B)), ldb,
                        _ptr(None if scaleA is None else np.ascontiguousarray(scaleA, dtype=np.uint8)),
                        _ptr(None if scaleB is None else np.ascontiguousarray(scaleB, dtype=np.uint8)),
                        float(beta), _ptr(None if C is None else np.ascontiguousarray(C, dtype=np.float32)),
                        ldc, D.reshape(-1), ldd, nthreads or threads())
    return D


def to_bf16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_bf16(x.reshape(-1), out.reshape(-1), x.size)
    return out


def from_bf16(h):
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_bf16_to_f32(h.reshape(-1), out.reshape(-1), h.size)
    return out


def to_f16(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint16)
    lib().orc_f32_to_f16(x.reshape(-1), out.reshape(-1), x.size)
    return out


def from_f16(h):
    h = np.ascontiguousarray(h, dtype=np.uint16)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_f16_to_f32(h.reshape(-1), out.reshape(-1), h.size)
    return out


def to_fp8(x, e4m3=True):
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty(x.shape, dtype=np.uint8)
    lib().orc_f32_to_fp8(x.reshape(-1), out.reshape(-1), x.size, int(e4m3))
    return out


def from_fp8(h, e4m3=True):
    h = np.ascontiguousarray(h, dtype=np.uint8)
    out = np.empty(h.shape, dtype=np.float32)
    lib().orc_fp8_to_f32(h.reshape(-1), out.reshape(-1), h.size, int(e4m3))
    return out


# ---------------------------------------------------------------- FP64 residual
def backward_error(A, B, X, nthreads=None):
    """Per-system normwise backward error; A (batch, n, n), B and X (batch, nrhs, n)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    B = np.ascontiguousarray(B, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    batch, n, _ = A.shape
    nrhs = B.shape[1]
    out = np.zeros(batch, dtype=np.float64)
    lib().orc_backward_error(A.reshape(-1), B.reshape(-1), X.reshape(-1), n, nrhs, batch, out,
                             nthreads or threads())
    return out


# ---------------------------------------------------------------- tolerances
def mixed_tolerance(out, gold, dtype="fp32"):
    """MIXED_TOLERANCE of reference test/frame/verify.h:234-333.  Returns (pass, detail)."""
    rtol, atol, fixed, mbits, emin = {
        "fp32": (0.0009765625, 1.52587890625e-5, 1e-2, 23, -126),
        "fp16": (0.001953125, 0.001953125, 1e-1, 10, -14),
        "bf16": (0.015625, 0.015625, 1.0, 7, -126),
    }[dtype]
    o = np.asarray(out, dtype=np.float32).reshape(-1).astype(np.float64)
    g = np.asarray(gold, dtype=np.float32).reshape(-1).astype(np.float64)
    skip = (o == g) | (np.isnan(o) & np.isnan(g))
    o, g = o[~skip], g[~skip]
    count = skip.size
    special = ~np.isfinite(o) | ~np.isfinite(g)
    diff = np.abs(o - g)
    gabs = np.abs(g)
    with np.errstate(divide="ignore", invalid="ignore"):
        e = np.where(gabs > 0, np.floor(np.log2(np.where(gabs > 0, gabs, 1.0))), emin)
    ulp = np.where(gabs > 0, np.ldexp(1.0, (np.maximum(e, emin) - mbits).astype(np.int64)),
                   2.0 ** (emin - mbits))
    fail = (diff > atol + rtol * gabs) | special
    limit_fail = bool(np.any(special) or np.any(diff > np.maximum(fixed, 32.0 * ulp)))
    matched = 1.0 - (np.count_nonzero(fail) / count if count else 0.0)
    ok = matched >= 0.99 and not limit_fail
    return ok, "matched=%.5f max_abs=%.3e special=%d" % (matched, float(diff.max()) if diff.size else 0.0,
                                                           int(special.sum()))


def mere_mare(out, gold, threshold, multiplier=10.0):
    """MERE_MARE of reference test/frame/verify.h:153-208.  Returns (pass, detail)."""
    o = np.asarray(out, dtype=np.float32).reshape(-1).astype(np.float64)
    g = np.asarray(gold, dtype=np.float32).reshape(-1).astype(np.float64)
    skip = (o == g) | (np.isnan(o) & np.isnan(g))
    o, g = o[~skip], g[~skip]
    inf = np.isinf(o) | np.isinf(g)
    mism = int(inf.sum())
    o, g = o[~inf], g[~inf]
    rel = np.abs(o - g) / (np.abs(g) + 2.0 ** -14)
    rel = np.where(np.isnan(rel), np.inf, rel)
    mere = float(rel.mean()) if rel.size else 0.0
    mare = float(rel.max()) if rel.size else 0.0
    ok = mism == 0 and mere < threshold and mare < multiplier * threshold
    detail = "MERE=%.3e MARE=%.3e mismatches=%d differing=%d" % (mere, mare, mism, int(rel.size))
    if rel.size:
        w = int(np.argmax(rel))
        detail += " worst out=%.9g gold=%.9g" % (o[w], g[w])
    return ok, detail

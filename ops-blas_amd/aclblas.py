"""ctypes binding of libops_blas.so, the aclblas C ABI built for MI355X.

The functions mirror the C prototypes of include/cann_ops_blas.h one to one:
same names, same argument order, pointers passed as integers (device addresses
from torch tensors' data_ptr(), host scalars as Python floats).  There is no
fallback: if the library is missing, load() raises.

Layout helpers: a batch of column-major n x n matrices with leading dimension
lda is a torch tensor of shape (batch, n, lda); element (i, j) of matrix b is
t[b, j, i].  ptr_array() builds the device array of per-matrix device pointers
that the batched LAPACK calls take.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# ACLBLAS_LIBRARY selects another build of the same ABI (the host-sanitizer build in tests)
LIB_PATH = os.environ.get("ACLBLAS_LIBRARY") or os.path.join(_HERE, "libops_blas.so")

# aclblasStatus_t (include/cann_ops_blas_common.h)
STATUS = {
    0: "SUCCESS", 1: "NOT_INITIALIZED", 2: "ALLOC_FAILED", 3: "INVALID_VALUE", 4: "MAPPING_ERROR",
    5: "EXECUTION_FAILED", 6: "INTERNAL_ERROR", 7: "NOT_SUPPORTED", 8: "ARCH_MISMATCH",
    9: "HANDLE_IS_NULLPTR", 10: "INVALID_ENUM", 11: "UNKNOWN",
}
STATUS_CODE = {v: k for k, v in STATUS.items()}
OP_N, OP_T, OP_C = 111, 112, 113
OPS = {"N": OP_N, "T": OP_T, "C": OP_C}
# aclDataType (include/acl/acl.h)
ACL_FLOAT, ACL_FLOAT16, ACL_BF16, ACL_FLOAT8_E5M2, ACL_FLOAT8_E4M3FN = 0, 1, 27, 35, 36
DTYPES = {"FP32": ACL_FLOAT, "FP16": ACL_FLOAT16, "BF16": ACL_BF16, "FP8_E5M2": ACL_FLOAT8_E5M2,
          "FP8_E4M3FN": ACL_FLOAT8_E4M3FN, "FP8_E4M3": ACL_FLOAT8_E4M3FN}
# aclblasComputeType_t
COMPUTE = {"COMPUTE_16F": 0, "COMPUTE_16F_PEDANTIC": 1, "COMPUTE_32F": 2, "COMPUTE_32F_PEDANTIC": 3,
           "COMPUTE_32F_FAST_16F": 4, "COMPUTE_32F_FAST_16BF": 5, "COMPUTE_32F_FAST_TF32": 6,
           "COMPUTE_64F": 7}
GEMM_DEFAULT = 0

_lib = None

_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.POINTER(ctypes.c_float)

# name -> argtypes; every entry is declared in include/cann_ops_blas.h
_SIGNATURES = {
    "aclblasCreate": [_P],
    "aclblasDestroy": [_P],
    "aclblasSetStream": [_P, _P],
    "aclblasGetStream": [_P, _P],
    "aclblasSetWorkspace": [_P, _P, ctypes.c_size_t],
    "aclblasGetVersion": [_P, _P],
    "aclblasLoggerConfigure": [_I, _I, _I, ctypes.c_char_p],
    "aclblasSetLoggerCallback": [_P],
    "aclblasGetLoggerCallback": [_P],
    "aclblasSgemm": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I],
    "aclblasSgemmEx": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _I],
    "aclblasGemmEx": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P, _I, _I, _P, _P, _I, _I, _I, _I],
    "aclblasGemmBatchedEx": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P, _I, _I, _P, _P, _I, _I, _I, _I, _I],
    "aclblasSgemmBatched": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _I],
    "aclblasSgetrfBatched": [_P, _I, _P, _I, _P, _P, _I],
    "aclblasSgetrsBatched": [_P, _I, _I, _I, _P, _I, _P, _P, _I, _P, _I],
    "aclblasSgetriBatched": [_P, _I, _P, _I, _P, _P, _I, _P, _I],
    "aclblasSmatinvBatched": [_P, _I, _P, _I, _P, _I, _P, _I],
    "aclblasGemmGroupedBatchedEx": [_P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _P, _I, _P, _P, _P, _I, _P, _I, _P, _I],
    "aclblasSgetrsBatchedAdjoint": [_P, _I, _I, _P, _I, _P, _P, _I, _P, _I, _P, _I, _I],
    "aclblasSgemmGroupedBatched": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "aclblasSgesvBatchedMixed": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, _I, _P, _P, _P],
    "aclblasSetLaunchEvents": [_P, _P, _I],
    "aclblasStrsm": [_P, _I, _I, _I, _I, _I, _I, _P, _P, _I, _P, _I],
    "aclblasSgesvBatchedMixedEx": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_double, _I, _I, _P, _P, _P, _P],
    "aclblasGemmExScaled": [_P, _I, _I, _I, _I, _I, _P, _P, _I, _I, _P, _P, _I, _I, _P, _P, _P, _I, _I],
    "aclblasQuantizeFp8": [_P, ctypes.c_longlong, _P, _I, _P, _P],
    "aclblasQuantizeFp8NoAlloc": [_P, ctypes.c_longlong, _P, _I, _P, _P],
    "aclblasSetTuningOverride": [_P, _I, _I, _I],
    "aclblasGetTuningChoice": [_P, _I, _I, _P, _P],
    "aclblasTuningBucket": [_P, _I, _I, ctypes.c_longlong, _I, _P],
    # include/cann_ops_blasLt.h
    "aclblasLtGetVersion": [_P],
    "aclblasLtGetProperty": [_I, _P],
    "aclblasLtCreate": [_P],
    "aclblasLtDestroy": [_P],
    "aclblasLtMatrixLayoutCreate": [_P, _I, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int64],
    "aclblasLtMatrixLayoutDestroy": [_P],
    "aclblasLtMatrixLayoutSetAttribute": [_P, _I, _P, ctypes.c_size_t],
    "aclblasLtMatrixLayoutGetAttribute": [_P, _I, _P, ctypes.c_size_t, _P],
    "aclblasLtMatmulDescCreate": [_P, _I, _I],
    "aclblasLtMatmulDescDestroy": [_P],
    "aclblasLtMatmulDescInit": [_P, _I, _I],
    "aclblasLtMatmulDescSetAttribute": [_P, _I, _P, ctypes.c_size_t],
    "aclblasLtMatmulDescGetAttribute": [_P, _I, _P, ctypes.c_size_t, _P],
    "aclblasLtMatmulPreferenceCreate": [_P],
    "aclblasLtMatmulPreferenceDestroy": [_P],
    "aclblasLtMatmulPreferenceSetAttribute": [_P, _I, _P, ctypes.c_size_t],
    "aclblasLtMatmulPreferenceGetAttribute": [_P, _I, _P, ctypes.c_size_t, _P],
    "aclblasLtMatmulAlgoInit": [_P, _I, _I, _I, _I, _I, _I, _I, _P],
    "aclblasLtMatmulAlgoGetIds": [_P, _I, _I, _I, _I, _I, _I, _P, _I, _P],
    "aclblasLtMatmulAlgoGetHeuristic": [_P, _P, _P, _P, _P, _P, _P, _I, _P, _P],
    "aclblasLtMatmul": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P],
}

# BLASLt (include/cann_ops_blasLt.h)
ACL_FLOAT4_E2M1 = 40
LT_ORDER_COL, LT_ORDER_ROW = 0, 1
LT_LAYOUT_BATCH_COUNT, LT_LAYOUT_OFFSET, LT_LAYOUT_TYPE, LT_LAYOUT_ORDER, LT_LAYOUT_ROWS, LT_LAYOUT_COLS, \
    LT_LAYOUT_LD = range(7)
LT_DESC_TRANSA, LT_DESC_TRANSB, LT_DESC_EPILOGUE, LT_DESC_BIAS_POINTER, LT_DESC_BIAS_DATA_TYPE, \
    LT_DESC_A_SCALE_POINTER, LT_DESC_B_SCALE_POINTER = range(7)
LT_DESC_A_SCALE_MODE, LT_DESC_B_SCALE_MODE = 31, 32
LT_EPILOGUE_DEFAULT, LT_EPILOGUE_RELU = 1, 2
LT_PREF_SEARCH_MODE, LT_PREF_MAX_WORKSPACE_BYTES = 0, 1


class LtAlgo(ctypes.Structure):
    _fields_ = [("data", ctypes.c_uint8 * 16), ("max_workspace_bytes", ctypes.c_size_t)]


class LtHeuristicResult(ctypes.Structure):
    _fields_ = [("algo", LtAlgo), ("workspaceSize", ctypes.c_size_t), ("state", ctypes.c_int),
                ("wavesCount", ctypes.c_float), ("reserved", ctypes.c_int * 4)]

# tuning operations (aclblasSetTuningOverride)
TUNE_GEMM, TUNE_GROUPED, TUNE_MIXED, TUNE_LU = 0, 1, 2, 3
# GEMM bucket element classes
CLASS_F32, CLASS_16, CLASS_F8, CLASS_F8_SCALED = 0, 1, 2, 3


def load(path=None):
    """Loads libops_blas.so (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError("libops_blas.so not found at %s; build it with `make -C ops-blas_amd`" % p)
        L = ctypes.CDLL(p)
        for name, args in _SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def exported_symbols():
    return sorted(_SIGNATURES)


def _fp(x):
    """Host float scalar (or None) -> pointer argument."""
    if x is None:
        return None
    return ctypes.cast(ctypes.pointer(ctypes.c_float(float(x))), _P)


class Handle:
    """Owns one aclblasHandle_t."""

    def __init__(self):
        self.h = _P()
        st = load().aclblasCreate(ctypes.byref(self.h))
        if st != 0:
            raise RuntimeError("aclblasCreate failed: %s" % STATUS.get(st, st))

    @property
    def value(self):
        return self.h.value

    def set_stream(self, stream_ptr):
        return load().aclblasSetStream(self.h, _P(stream_ptr))

    def destroy(self):
        if self.h.value:
            st = load().aclblasDestroy(self.h)
            self.h = _P()
            return st
        return 0

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.destroy()


# ---------------------------------------------------------------- raw calls
def sgemm(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    return load().aclblasSgemm(h, ta, tb, m, n, k, _fp(alpha), A, lda, B, ldb, _fp(beta), C, ldc)


def sgemm_ex(h, ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, algo=GEMM_DEFAULT):
    return load().aclblasSgemmEx(h, ta, tb, m, n, k, _fp(alpha), A, lda, B, ldb, _fp(beta), C, ldc, algo)


def gemm_ex(h, ta, tb, m, n, k, alpha, A, atype, lda, B, btype, ldb, beta, C, ctype, ldc, compute,
            algo=GEMM_DEFAULT):
    return load().aclblasGemmEx(h, ta, tb, m, n, k, _fp(alpha), A, atype, lda, B, btype, ldb, _fp(beta), C,
                                ctype, ldc, compute, algo)


def gemm_batched_ex(h, ta, tb, m, n, k, alpha, Aarr, atype, lda, Barr, btype, ldb, beta, Carr, ctype, ldc, batch,
                    compute, algo=GEMM_DEFAULT):
    return load().aclblasGemmBatchedEx(h, ta, tb, m, n, k, _fp(alpha), Aarr, atype, lda, Barr, btype, ldb, _fp(beta),
                                       Carr, ctype, ldc, batch, compute, algo)


def sgemm_batched(h, ta, tb, m, n, k, alpha, Aarr, lda, Barr, ldb, beta, Carr, ldc, batch):
    return load().aclblasSgemmBatched(h, ta, tb, m, n, k, _fp(alpha), Aarr, lda, Barr, ldb, _fp(beta), Carr, ldc,
                                      batch)


def sgetrf_batched(h, n, Aarray, lda, piv, info, batch):
    return load().aclblasSgetrfBatched(h, n, Aarray, lda, piv, info, batch)


def sgetrs_batched(h, trans, n, nrhs, Aarray, lda, piv, Barray, ldb, info_host, batch):
    """info_host: a ctypes.c_int (or None); it receives the LAPACK-style argument info."""
    ip = ctypes.byref(info_host) if info_host is not None else None
    return load().aclblasSgetrsBatched(h, trans, n, nrhs, Aarray, lda, piv, Barray, ldb, ip, batch)


def sgetri_batched(h, n, Aarray, lda, piv, Carray, ldc, info, batch):
    return load().aclblasSgetriBatched(h, n, Aarray, lda, piv, Carray, ldc, info, batch)


def smatinv_batched(h, n, Aarray, lda, Ainv, lda_inv, info, batch):
    return load().aclblasSmatinvBatched(h, n, Aarray, lda, Ainv, lda_inv, info, batch)


def _host_array(ctype, values):
    """Python list -> ctypes host array; a ctypes array (host_ints / host_ptrs) passes through."""
    if values is None or isinstance(values, ctypes.Array):
        return values
    arr = (ctype * max(1, len(values)))(*values)
    return arr


def host_ints(values):
    """Host int array for the calls that take host arrays (built once, reusable across calls)."""
    return (ctypes.c_int * max(1, len(values)))(*values)


def host_ptrs(values):
    """Host array of device addresses for the calls that take host pointer arrays."""
    return (ctypes.c_void_p * max(1, len(values)))(*values)


def sgemm_grouped_batched(h, transa, transb, m, n, k, alpha, A, lda, Bp, ldb, beta, C, ldc, group_size):
    """All arguments are Python lists (host arrays) or None; A/Bp/C are lists of device addresses."""
    args = [_host_array(ctypes.c_int, transa), _host_array(ctypes.c_int, transb), _host_array(ctypes.c_int, m),
            _host_array(ctypes.c_int, n), _host_array(ctypes.c_int, k), _host_array(ctypes.c_float, alpha),
            _host_array(ctypes.c_void_p, A), _host_array(ctypes.c_int, lda), _host_array(ctypes.c_void_p, Bp),
            _host_array(ctypes.c_int, ldb), _host_array(ctypes.c_float, beta), _host_array(ctypes.c_void_p, C),
            _host_array(ctypes.c_int, ldc), _host_array(ctypes.c_int, group_size)]
    count = len(group_size) if group_size is not None else (len(m) if m is not None else 0)
    return sgemm_grouped_batched_raw(h, count, *args)


def sgemm_grouped_batched_raw(h, group_count, *arrays):
    ptrs = [ctypes.cast(a, ctypes.c_void_p) if a is not None else None for a in arrays]
    return load().aclblasSgemmGroupedBatched(h, group_count, *ptrs)


def gemm_grouped_batched_ex(h, transa, transb, m, n, k, alpha, Aarray, atype, lda, Barray, btype, ldb, beta,
                            Carray, ctype, ldc, group_size, compute, group_count=None):
    """transa .. ldc, group_size: Python lists (host arrays) or None; alpha / beta: lists of floats (COMPUTE_32F)
    or ctypes arrays of FP16 bit patterns (COMPUTE_16F); Aarray / Barray / Carray: device addresses of device
    pointer arrays holding every problem of every group, group after group."""
    if group_count is None:
        group_count = len(group_size) if group_size is not None else 0
    ints = lambda v: _host_array(ctypes.c_int, v)
    arrs = [ints(transa), ints(transb), ints(m), ints(n), ints(k), _host_array(ctypes.c_float, alpha)]
    ptrs = [ctypes.cast(a, ctypes.c_void_p) if a is not None else None for a in arrs]
    lda_a, ldb_a, beta_a, ldc_a, gs_a = ints(lda), ints(ldb), _host_array(ctypes.c_float, beta), ints(ldc), \
        ints(group_size)
    cast = lambda a: ctypes.cast(a, ctypes.c_void_p) if a is not None else None
    return load().aclblasGemmGroupedBatchedEx(h, *ptrs, Aarray, atype, cast(lda_a), Barray, btype, cast(ldb_a),
                                              cast(beta_a), Carray, ctype, cast(ldc_a), int(group_count),
                                              cast(gs_a), compute)


def sgesv_batched_mixed(h, n, nrhs, A, lda, Bp, ldb, X, ldx, tol, max_iter, iters=None, berr=None, info=None):
    """n .. ldx: Python lists (host arrays); A/Bp/X: lists of device addresses; iters/berr/info: device addresses."""
    count = len(n) if n is not None else 0
    arrs = [_host_array(ctypes.c_int, n), _host_array(ctypes.c_int, nrhs), _host_array(ctypes.c_void_p, A),
            _host_array(ctypes.c_int, lda), _host_array(ctypes.c_void_p, Bp), _host_array(ctypes.c_int, ldb),
            _host_array(ctypes.c_void_p, X), _host_array(ctypes.c_int, ldx)]
    ptrs = [ctypes.cast(a, ctypes.c_void_p) if a is not None else None for a in arrs]
    return load().aclblasSgesvBatchedMixed(h, count, *ptrs, float(tol), int(max_iter), iters, berr, info)


SIDE = {"LEFT": 141, "RIGHT": 142}
FILL = {"UPPER": 121, "LOWER": 122}
DIAG = {"NON_UNIT": 131, "UNIT": 132}


def strsm(h, side, uplo, trans, diag, m, n, alpha, A, lda, Bp, ldb):
    """alpha: Python float (or None for a NULL pointer); A, Bp: device addresses."""
    return load().aclblasStrsm(h, side, uplo, trans, diag, m, n, _fp(alpha), A, lda, Bp, ldb)


def set_launch_events(h, events, count):
    """events: a ctypes array of hipEvent_t handles (or None with count 0)."""
    return load().aclblasSetLaunchEvents(h, events, count)


MIXED_UPDATE = {"FP32": 0, "BF16": 1, "BF16X2": 2, "BF16X3": 3}


def sgesv_batched_mixed_ex(h, n, nrhs, A, lda, Bp, ldb, X, ldx, tol, max_iter, update, iters=None, berr=None,
                           info=None, fallback=None):
    """As sgesv_batched_mixed, with the trailing-update precision (MIXED_UPDATE value) and the fallback array."""
    count = len(n) if n is not None else 0
    arrs = [_host_array(ctypes.c_int, n), _host_array(ctypes.c_int, nrhs), _host_array(ctypes.c_void_p, A),
            _host_array(ctypes.c_int, lda), _host_array(ctypes.c_void_p, Bp), _host_array(ctypes.c_int, ldb),
            _host_array(ctypes.c_void_p, X), _host_array(ctypes.c_int, ldx)]
    ptrs = [ctypes.cast(a, ctypes.c_void_p) if a is not None else None for a in arrs]
    return load().aclblasSgesvBatchedMixedEx(h, count, *ptrs, float(tol), int(max_iter), int(update), iters, berr,
                                             info, fallback)


def gemm_ex_scaled(h, ta, tb, m, n, k, alpha, A, atype, lda, scale_a, B, btype, ldb, scale_b, beta, C, ctype, ldc):
    """scale_a / scale_b: device addresses of one float each (or None)."""
    return load().aclblasGemmExScaled(h, ta, tb, m, n, k, _fp(alpha), A, atype, lda, scale_a, B, btype, ldb, scale_b,
                                      _fp(beta), C, ctype, ldc)


def quantize_fp8_no_alloc(h, count, x, qtype, q, scale):
    """aclblasQuantizeFp8NoAlloc: as quantize_fp8, with the library's own scratch array (allocates nothing)."""
    return load().aclblasQuantizeFp8NoAlloc(h, count, x, qtype, q, scale)


def quantize_fp8(h, count, x, qtype, q, scale):
    return load().aclblasQuantizeFp8(h, count, x, qtype, q, scale)


def sgetrs_batched_adjoint(h, n, nrhs, LUarray, lda, piv, Xarray, ldx, dXarray, lddx, dAarray, ldda, batch):
    """dX <- A^-T dX (in place, the gradient of B) and dA <- -(A^-T dX) X^T; dAarray may be None."""
    return load().aclblasSgetrsBatchedAdjoint(h, n, nrhs, LUarray, lda, piv, Xarray, ldx, dXarray, lddx, dAarray,
                                              ldda, batch)


def set_tuning_override(h, op, a, b=0):
    """Forces launch choice (a, b) for every bucket of `op` on handle h; a < 0 removes it."""
    return load().aclblasSetTuningOverride(h, op, a, b)


def get_tuning_choice(h, op, bucket):
    """(status, a, b) of the choice handle h uses for (op, bucket)."""
    a, b = ctypes.c_int(0), ctypes.c_int(0)
    st = load().aclblasGetTuningChoice(h, op, bucket, ctypes.byref(a), ctypes.byref(b))
    return st, a.value, b.value


def tuning_bucket(h, op, cls, units, k=0):
    """(status, bucket) of a problem (see aclblasTuningBucket)."""
    out = ctypes.c_int(-1)
    st = load().aclblasTuningBucket(h, op, cls, units, k, ctypes.byref(out))
    return st, out.value


# ---------------------------------------------------------------- torch helpers
def ptr_array(t):
    """Device int64 tensor holding the address of each t[b] (t: (batch, ...) contiguous CUDA tensor).

    The pointer array is kept alive as an attribute of `t`: a temporary such as
    `ptr_array(t).data_ptr()` would otherwise be freed (and its memory handed to
    the next allocation) before the asynchronous kernel that reads it has run.
    """
    import torch
    batch = t.shape[0]
    stride = t.stride(0) * t.element_size()
    base = t.data_ptr()
    cached = getattr(t, "_aclblas_ptrs", None)
    if cached is not None and cached[0] == (base, stride, batch):
        return cached[1]
    ptrs = torch.tensor([base + b * stride for b in range(batch)], dtype=torch.int64, device=t.device)
    t._aclblas_ptrs = ((base, stride, batch), ptrs)
    return ptrs


# ---------------------------------------------------------------- BLASLt helpers
def lt_create():
    h = _P()
    st = load().aclblasLtCreate(ctypes.byref(h))
    return st, h


def lt_layout(dtype, rows, cols, ld, order=LT_ORDER_ROW):
    """A matrix layout with ORDER set (the reference's callers set ROW); returns (status, layout)."""
    lay = _P()
    st = load().aclblasLtMatrixLayoutCreate(ctypes.byref(lay), int(dtype), int(rows), int(cols), int(ld))
    if st == 0 and order is not None:
        o = ctypes.c_int32(order)
        st = load().aclblasLtMatrixLayoutSetAttribute(lay, LT_LAYOUT_ORDER, ctypes.byref(o), 4)
    return st, lay


def lt_desc(compute=2, scale_type=ACL_FLOAT, transA=OP_N, transB=OP_N, scaleA=None, scaleB=None):
    d = _P()
    st = load().aclblasLtMatmulDescCreate(ctypes.byref(d), compute, scale_type)
    for attr, v in ((LT_DESC_TRANSA, transA), (LT_DESC_TRANSB, transB)):
        if st == 0:
            x = ctypes.c_int32(v)
            st = load().aclblasLtMatmulDescSetAttribute(d, attr, ctypes.byref(x), 4)
    for attr, v in ((LT_DESC_A_SCALE_POINTER, scaleA), (LT_DESC_B_SCALE_POINTER, scaleB)):
        if st == 0 and v is not None:
            x = ctypes.c_void_p(v)
            st = load().aclblasLtMatmulDescSetAttribute(d, attr, ctypes.byref(x), ctypes.sizeof(ctypes.c_void_p))
    return st, d


def lt_matmul(h, desc, alpha, A, Adesc, B, Bdesc, beta, C, Cdesc, D, Ddesc, algo=None, workspace=None, ws_size=0,
              stream=None):
    ap = ctypes.byref(algo) if algo is not None else None
    return load().aclblasLtMatmul(h, desc, _fp(alpha), A, Adesc, B, Bdesc, _fp(beta), C, Cdesc, D, Ddesc, ap,
                                  workspace, ws_size, stream)

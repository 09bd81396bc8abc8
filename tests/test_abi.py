"""The C ABI on the CPU: exports, handle semantics and every argument-check row.

Runs without a GPU.  libops_blas.so is loaded with ctypes; on a host with no
visible GPU aclblasCreate returns a handle without workspace, so every status
code that the reference decides before touching device memory can be checked
here, row by row from the reference's CSV tables.  Pointers passed in these
rows are never dereferenced (the calls return during argument checking).
"""
import ctypes
import os
import re

import pytest

import aclblas as B
from common import load_cases, status_code

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", "cann_ops_blas.h"), os.path.join(ROOT, "include", "cann_ops_blasLt.h")]


def _declared():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"aclblasStatus_t\s+(aclblas\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = B.load()
    declared = _declared()
    assert declared, "no declarations parsed"
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(B.exported_symbols()) == declared


@pytest.fixture(scope="module")
def cpu_handle():
    try:
        import torch
        if torch.cuda.is_available():
            # with a GPU, the handle owns device memory: still fine for argument checks
            torch.cuda.init()
    except ImportError:
        pass
    h = B.Handle()
    yield h.value
    h.destroy()


_buf = (ctypes.c_float * 64)()
DUMMY = ctypes.addressof(_buf)


def test_handle_semantics():
    lib = B.load()
    assert lib.aclblasCreate(None) == status_code("HANDLE_IS_NULLPTR")
    h = ctypes.c_void_p(DUMMY)
    assert lib.aclblasCreate(ctypes.byref(h)) == status_code("INVALID_VALUE")
    assert lib.aclblasDestroy(None) == status_code("HANDLE_IS_NULLPTR")
    assert lib.aclblasSetStream(None, None) == status_code("HANDLE_IS_NULLPTR")
    assert lib.aclblasGetStream(None, None) == status_code("HANDLE_IS_NULLPTR")
    v = ctypes.c_int(0)
    assert lib.aclblasGetVersion(None, ctypes.byref(v)) == 0 and v.value == 10000
    assert lib.aclblasGetVersion(None, None) == status_code("INVALID_VALUE")
    h = B.Handle()
    s = ctypes.c_void_p(1)
    assert lib.aclblasGetStream(h.h, ctypes.byref(s)) == 0 and not s.value
    assert lib.aclblasGetStream(h.h, None) == status_code("INVALID_VALUE")
    assert lib.aclblasSetWorkspace(h.h, None, 16) == status_code("INVALID_VALUE")
    assert lib.aclblasSetWorkspace(h.h, DUMMY, 0) == status_code("INVALID_VALUE")
    assert h.destroy() == 0


def test_logger_entry_points():
    lib = B.load()
    assert lib.aclblasLoggerConfigure(0, 0, 0, None) == 0
    CB = ctypes.CFUNCTYPE(None, ctypes.c_char_p)
    seen = []
    cb = CB(lambda msg: seen.append(msg))
    assert lib.aclblasSetLoggerCallback(ctypes.cast(cb, ctypes.c_void_p)) == 0
    out = ctypes.c_void_p()
    assert lib.aclblasGetLoggerCallback(ctypes.byref(out)) == 0
    assert out.value == ctypes.cast(cb, ctypes.c_void_p).value
    assert lib.aclblasGetLoggerCallback(None) == status_code("INVALID_VALUE")
    # an argument error is reported through the callback
    assert B.sgetrf_batched(None, 4, DUMMY, 4, None, None, 1) == status_code("HANDLE_IS_NULLPTR")
    lib.aclblasSetLoggerCallback(None)


def _bool(v):
    return str(v).strip().lower() == "true"


def _op(s):
    s = s.replace("ACLBLAS_OP_", "")
    return B.OPS.get(s, 999)


# ------------------------------------------------------------------ LAPACK rows
def _lapack_rows(name, n_key, batch_key):
    rows = []
    for r in load_cases(name):
        n, batch = int(r[n_key]), int(r[batch_key])
        if status_code(r["expect_result"]) != 0 or n <= 0 or batch <= 0 or \
                ("nrhs" in r and int(r["nrhs"]) <= 0):
            rows.append(r)
    return rows


GETRF_ROWS = _lapack_rows("sgetrf_batched.csv", "n", "batch_size")


@pytest.mark.parametrize("row", GETRF_ROWS, ids=[r["case_name"] for r in GETRF_ROWS])
def test_getrf_argument_rows(cpu_handle, row):
    kind = row["matrix_type"]
    h = None if kind == "NULLPTR_HANDLE" else cpu_handle
    A = None if kind == "NULLPTR_AARRAY" else DUMMY
    piv = DUMMY if row["pivot_mode"] != "NO_PIVOT" else None
    info = None if kind == "NULLPTR_INFOARRAY" else DUMMY
    if status_code(row["expect_result"]) == 0:
        piv = None  # the reference's no-op path passes no pivot array
    st = B.sgetrf_batched(h, int(row["n"]), A, int(row["lda"]), piv, info, int(row["batch_size"]))
    assert st == status_code(row["expect_result"])


GETRS_ROWS = _lapack_rows("getrs_batched.csv", "n", "batch_count")


def _expected_getrs_info(r):
    """reference test/getrs_batched/arch35/getrs_batched_test.cpp:76-99"""
    n, nrhs, batch = int(r["n"]), int(r["nrhs"]), int(r["batch_count"])
    lda, ldb = int(r["lda"]), int(r["ldb"])
    kind = r["matrix_type"]
    if kind == "NULLPTR_HANDLE":
        return None
    if _op(r["trans"]) == 999:
        return -2
    if n < 0:
        return -3
    if nrhs < 0 or nrhs > 256:
        return -4
    if n > 0 and nrhs > 0 and batch > 0 and kind == "NULLPTR_AARRAY":
        return -5
    if lda < max(1, n):
        return -6
    if n > 0 and nrhs > 0 and batch > 0 and kind == "NULLPTR_BARRAY":
        return -8
    if ldb < max(1, n):
        return -9
    if batch < 0:
        return -11
    return 0


@pytest.mark.parametrize("row", GETRS_ROWS, ids=[r["case_name"] for r in GETRS_ROWS])
def test_getrs_argument_rows(cpu_handle, row):
    kind = row["matrix_type"]
    h = None if kind == "NULLPTR_HANDLE" else cpu_handle
    A = None if kind == "NULLPTR_AARRAY" else DUMMY
    Bp = None if kind == "NULLPTR_BARRAY" else DUMMY
    info = ctypes.c_int(12345)
    use_info = row.get("info_mode", "NORMAL") != "NULLPTR"
    st = B.sgetrs_batched(h, _op(row["trans"]), int(row["n"]), int(row["nrhs"]), A, int(row["lda"]), None, Bp,
                          int(row["ldb"]), info if use_info else None, int(row["batch_count"]))
    assert st == status_code(row["expect_result"])
    if use_info:
        exp = _expected_getrs_info(row)
        assert info.value == (12345 if exp is None else exp)


GETRI_ROWS = _lapack_rows("sgetri_batched.csv", "n", "batch_size")


@pytest.mark.parametrize("row", GETRI_ROWS, ids=[r["case_name"] for r in GETRI_ROWS])
def test_getri_argument_rows(cpu_handle, row):
    kind = row["matrix_type"]
    h = None if kind == "NULLPTR_HANDLE" else cpu_handle
    A = None if kind == "NULLPTR_AARRAY" else DUMMY
    C = None if kind == "NULLPTR_CARRAY" else DUMMY
    info = None if kind == "NULLPTR_INFOARRAY" else DUMMY
    st = B.sgetri_batched(h, int(row["n"]), A, int(row["lda"]), None, C, int(row["ldc"]), info,
                          int(row["batch_size"]))
    assert st == status_code(row["expect_result"])


def test_getri_null_handle():
    assert B.sgetri_batched(None, 4, DUMMY, 4, None, DUMMY, 4, DUMMY, 1) == status_code("HANDLE_IS_NULLPTR")


MATINV_ROWS = _lapack_rows("smatinv_batched.csv", "n", "batch_size")


@pytest.mark.parametrize("row", MATINV_ROWS, ids=[r["case_name"] for r in MATINV_ROWS])
def test_matinv_argument_rows(cpu_handle, row):
    kind = row["matrix_type"]
    n, batch = int(row["n"]), int(row["batch_size"])
    if status_code(row["expect_result"]) == 0:
        # no-op rows: the reference passes null arrays (matinv_batched_test.cpp:96-103)
        A = Ainv = info = None
    else:
        A = None if kind in ("NULLPTR_A", "NULLPTR_ALL") else DUMMY
        Ainv = None if kind in ("NULLPTR_AINV", "NULLPTR_ALL") else DUMMY
        info = None if kind in ("NULLPTR_INFO", "NULLPTR_ALL") else DUMMY
    st = B.smatinv_batched(cpu_handle, n, A, int(row["lda"]), Ainv, int(row["lda_inv"]), info, batch)
    assert st == status_code(row["expect_result"])


def test_matinv_null_handle():
    assert B.smatinv_batched(None, 4, DUMMY, 4, DUMMY, 4, DUMMY, 1) == status_code("HANDLE_IS_NULLPTR")


GROUPED_EX_ROWS = [r for r in load_cases("gemm_grouped_batched_ex.csv")
                   if status_code(r["expect_result"]) != 0 or int(r["group_count"]) == 0]


@pytest.mark.parametrize("row", GROUPED_EX_ROWS, ids=[r["case_name"] for r in GROUPED_EX_ROWS])
def test_grouped_ex_status_rows(cpu_handle, row):
    gc = int(row["group_count"])
    one = [1] * max(gc, 1)
    st = B.gemm_grouped_batched_ex(cpu_handle, [B.OP_N] * len(one), [B.OP_N] * len(one), [16] * len(one),
                                   [16] * len(one), [32] * len(one), [1.0] * len(one), DUMMY,
                                   B.DTYPES[row["Atype"]], [16] * len(one), DUMMY, B.DTYPES[row["Btype"]],
                                   [32] * len(one), [0.0] * len(one), DUMMY, B.DTYPES[row["Ctype"]],
                                   [16] * len(one), one, B.COMPUTE[row["computeType"]], group_count=gc)
    assert st == status_code(row["expect_result"])


def test_grouped_ex_argument_checks(cpu_handle):
    inv, ns = status_code("INVALID_VALUE"), status_code("NOT_SUPPORTED")
    F16, C32 = B.ACL_FLOAT16, B.COMPUTE["COMPUTE_32F"]

    def call(h=cpu_handle, ta=(B.OP_N,), m=(16,), k=(16,), lda=(16,), ldb=(16,), ldc=(16,), size=(1,), gc=1,
             A=DUMMY, at=F16, ct=F16, comp=C32, alpha=(1.0,)):
        return B.gemm_grouped_batched_ex(h, list(ta), [B.OP_N], list(m), [16], list(k), list(alpha), A, at,
                                         list(lda), DUMMY, at, list(ldb), [0.0], DUMMY, ct, list(ldc), list(size),
                                         comp, group_count=gc)

    assert call(h=None) == status_code("HANDLE_IS_NULLPTR")
    assert call(ct=B.ACL_BF16) == ns                       # FP16 x FP16 -> BF16
    assert call(at=B.ACL_FLOAT) == ns                      # FP32 is not a grouped-Ex type
    assert call(at=B.ACL_BF16, ct=B.ACL_BF16, comp=B.COMPUTE["COMPUTE_16F"]) == ns
    assert call(gc=-1) == inv
    assert call(ta=(7,)) == inv
    assert call(m=(-1,)) == inv
    assert call(size=(-2,)) == inv
    assert call(lda=(15,)) == inv
    assert call(ldb=(15,)) == inv
    assert call(ldc=(15,)) == inv
    assert call(A=None) == inv
    assert call(A=None, size=(0,)) == 0                    # no problems: arrays may be NULL, nothing runs
    assert call(gc=0) == 0


# ------------------------------------------------------------------ GEMM rows
def _gemm_degenerate(r):
    m, n = int(r["m"]), int(r["n"])
    return m == 0 or n == 0


SGEMM_ROWS = [r for r in load_cases("sgemm_ex.csv")
              if status_code(r["expect_result"]) != 0 or _gemm_degenerate(r)]


@pytest.mark.parametrize("row", SGEMM_ROWS, ids=[r["case_name"] for r in SGEMM_ROWS])
def test_sgemm_ex_argument_rows(cpu_handle, row):
    exp = status_code(row["expect_result"])
    h = None if exp == status_code("HANDLE_IS_NULLPTR") else cpu_handle
    alpha = None if _bool(row["alpha_null"]) else float(row["alpha"])
    beta = None if _bool(row["beta_null"]) else float(row["beta"])
    A = None if _bool(row["a_null"]) else DUMMY
    Bp = None if _bool(row["b_null"]) else DUMMY
    C = None if _bool(row["c_null"]) else DUMMY
    algo = 0 if row["algo"] == "DEFAULT" else 1 + int(row["algo"].replace("ALGO", ""))
    args = (h, _op(row["transA"]), _op(row["transB"]), int(row["m"]), int(row["n"]), int(row["k"]), alpha, A,
            int(row["lda"]), Bp, int(row["ldb"]), beta, C, int(row["ldc"]))
    assert B.sgemm_ex(*args, algo=algo) == exp
    if algo == 0:
        assert B.sgemm(*args) == exp


GEMMEX_ROWS = [r for r in load_cases("gemm_ex.csv")
               if status_code(r["expect_result"]) != 0 or _gemm_degenerate(r)]


def gemm_ex_lds(r):
    """Default leading dimensions of the reference's GemmEx table (blank = minimum)."""
    m, n, k = int(r["m"]), int(r["n"]), int(r["k"])
    ta, tb = r["transA"] != "N", r["transB"] != "N"
    lda = int(r["lda"]) if r["lda"] else max(1, k if ta else m)
    ldb = int(r["ldb"]) if r["ldb"] else max(1, n if tb else k)
    ldc = int(r["ldc"]) if r["ldc"] else max(1, m)
    return lda, ldb, ldc


@pytest.mark.parametrize("row", GEMMEX_ROWS, ids=[r["case_name"] for r in GEMMEX_ROWS])
def test_gemm_ex_argument_rows(cpu_handle, row):
    exp = status_code(row["expect_result"])
    h = None if exp == status_code("HANDLE_IS_NULLPTR") else cpu_handle
    alpha = None if _bool(row["alpha_null"]) else float(row["alpha"])
    beta = None if _bool(row["beta_null"]) else float(row["beta"])
    A = None if _bool(row["a_null"]) else DUMMY
    Bp = None if _bool(row["b_null"]) else DUMMY
    C = None if _bool(row["c_null"]) else DUMMY
    lda, ldb, ldc = gemm_ex_lds(row)
    st = B.gemm_ex(h, _op(row["transA"]), _op(row["transB"]), int(row["m"]), int(row["n"]), int(row["k"]), alpha,
                   A, B.DTYPES[row["Atype"]], lda, Bp, B.DTYPES[row["Btype"]], ldb, beta, C,
                   B.DTYPES[row["Ctype"]], ldc, B.COMPUTE[row["computeType"]])
    assert st == exp


def test_gemm_ex_extensions_are_accepted():
    """BF16 x BF16 -> FP32 and FP16 x FP16 -> FP32 (COMPUTE_32F) pass argument checks."""
    # m = 0 returns after the dtype check, so a null handle is not needed
    h = B.Handle()
    for at, ct in ((B.ACL_BF16, B.ACL_FLOAT), (B.ACL_FLOAT16, B.ACL_FLOAT)):
        st = B.gemm_ex(h.value, B.OP_N, B.OP_N, 0, 8, 8, 1.0, DUMMY, at, 1, DUMMY, at, 8, 0.0, DUMMY, ct, 1,
                       B.COMPUTE["COMPUTE_32F"])
        assert st == 0
    st = B.gemm_ex(h.value, B.OP_N, B.OP_N, 0, 8, 8, 1.0, DUMMY, B.ACL_BF16, 1, DUMMY, B.ACL_BF16, 8, 0.0, DUMMY,
                   B.ACL_FLOAT16, 1, B.COMPUTE["COMPUTE_32F"])
    assert st == status_code("NOT_SUPPORTED")
    h.destroy()


BATCHED_ROWS = [r for r in load_cases("gemm_batched_ex.csv")
                if status_code(r["expect_result"]) != 0 or _gemm_degenerate(r) or int(r["batchCount"]) == 0]


@pytest.mark.parametrize("row", BATCHED_ROWS, ids=[r["case_name"] for r in BATCHED_ROWS])
def test_gemm_batched_ex_argument_rows(cpu_handle, row):
    exp = status_code(row["expect_result"])
    h = None if exp == status_code("HANDLE_IS_NULLPTR") else cpu_handle
    alpha = None if _bool(row["alpha_null"]) else float(row["alpha"])
    beta = None if _bool(row["beta_null"]) else float(row["beta"])
    A = None if _bool(row["aarray_null"]) else DUMMY
    Bp = None if _bool(row["barray_null"]) else DUMMY
    C = None if _bool(row["carray_null"]) else DUMMY
    lda, ldb, ldc = gemm_ex_lds(row)
    algo = 0 if row["algo"] == "DEFAULT" else 1 + int(row["algo"].replace("ALGO", ""))
    st = B.gemm_batched_ex(h, _op(row["transA"]), _op(row["transB"]), int(row["m"]), int(row["n"]), int(row["k"]),
                           alpha, A, B.DTYPES[row["Atype"]], lda, Bp, B.DTYPES[row["Btype"]], ldb, beta, C,
                           B.DTYPES[row["Ctype"]], ldc, int(row["batchCount"]), B.COMPUTE[row["computeType"]], algo)
    assert st == exp


SGEMM_BATCHED_ROWS = [r for r in load_cases("sgemm_batched.csv")
                      if status_code(r["expect_result"]) != 0 or _gemm_degenerate(r) or int(r["batchCount"]) == 0]


@pytest.mark.parametrize("row", SGEMM_BATCHED_ROWS, ids=[r["case_name"] for r in SGEMM_BATCHED_ROWS])
def test_sgemm_batched_argument_rows(cpu_handle, row):
    """Error and no-work rows of test/gemm_batched/sgemm_batched/arch35/sgemm_batched_test.csv."""
    exp = status_code(row["expect_result"])
    h = None if exp == status_code("HANDLE_IS_NULLPTR") else cpu_handle
    alpha = None if _bool(row["alpha_null"]) else float(row["alpha"])
    beta = None if _bool(row["beta_null"]) else float(row["beta"])
    A = None if _bool(row["aarray_null"]) else DUMMY
    Bp = None if _bool(row["barray_null"]) else DUMMY
    C = None if _bool(row["carray_null"]) else DUMMY
    lda, ldb, ldc = gemm_ex_lds(row)
    st = B.sgemm_batched(h, _op(row["transA"]), _op(row["transB"]), int(row["m"]), int(row["n"]), int(row["k"]),
                         alpha, A, lda, Bp, ldb, beta, C, ldc, int(row["batchCount"]))
    assert st == exp


def test_sgemm_batched_argument_checks(cpu_handle):
    ok = status_code("SUCCESS")
    inv = status_code("INVALID_VALUE")
    f = B.sgemm_batched
    assert f(None, B.OP_N, B.OP_N, 4, 4, 4, 1.0, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == status_code("HANDLE_IS_NULLPTR")
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, None, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, -1, 4, 4, 1.0, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, -1) == inv
    assert f(cpu_handle, 999, B.OP_N, 4, 4, 4, 1.0, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, DUMMY, 3, DUMMY, 4, 0.0, DUMMY, 4, 1) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, None, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 0, 4, 4, 1.0, DUMMY, 4, DUMMY, 4, 0.0, DUMMY, 4, 1) == ok
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, None, 4, None, 4, 0.0, None, 4, 0) == ok


def test_getrs_adjoint_argument_checks(cpu_handle):
    ok = status_code("SUCCESS")
    inv = status_code("INVALID_VALUE")
    f = B.sgetrs_batched_adjoint
    D = DUMMY
    assert f(None, 4, 2, D, 4, D, D, 4, D, 4, D, 4, 1) == status_code("HANDLE_IS_NULLPTR")
    assert f(cpu_handle, -1, 2, D, 4, D, D, 4, D, 4, D, 4, 1) == inv
    assert f(cpu_handle, 4, -1, D, 4, D, D, 4, D, 4, D, 4, 1) == inv
    assert f(cpu_handle, 4, 257, D, 4, D, D, 4, D, 4, D, 4, 1) == inv
    assert f(cpu_handle, 4, 2, D, 4, D, D, 4, D, 4, D, 4, -1) == inv
    assert f(cpu_handle, 4, 2, D, 3, D, D, 4, D, 4, D, 4, 1) == inv      # lda
    assert f(cpu_handle, 4, 2, D, 4, D, D, 3, D, 4, D, 4, 1) == inv      # ldx
    assert f(cpu_handle, 4, 2, D, 4, D, D, 4, D, 3, D, 4, 1) == inv      # lddx
    assert f(cpu_handle, 4, 2, D, 4, D, D, 4, D, 4, D, 3, 1) == inv      # ldda
    assert f(cpu_handle, 4, 2, None, 4, D, D, 4, D, 4, D, 4, 1) == inv   # LU
    assert f(cpu_handle, 4, 2, D, 4, D, D, 4, None, 4, D, 4, 1) == inv   # dX
    assert f(cpu_handle, 4, 2, D, 4, D, None, 4, D, 4, D, 4, 1) == inv   # X needed for dA
    # quick returns before the device is needed
    assert f(cpu_handle, 0, 2, None, 1, None, None, 1, None, 1, None, 1, 1) == ok
    assert f(cpu_handle, 4, 2, None, 4, None, None, 4, None, 4, None, 4, 0) == ok


def test_gesv_mixed_argument_checks(cpu_handle):
    inv = status_code("INVALID_VALUE")
    ns = status_code("NOT_SUPPORTED")
    f = B.sgesv_batched_mixed
    D = DUMMY
    assert f(None, [4], [1], [D], [4], [D], [4], [D], [4], 1e-6, 3) == status_code("HANDLE_IS_NULLPTR")
    assert f(cpu_handle, [4], [1], [D], [4], [D], [4], [D], [4], -1.0, 3) == inv       # tol
    assert f(cpu_handle, [4], [1], [D], [4], [D], [4], [D], [4], 1e-6, -1) == inv      # maxIter
    assert f(cpu_handle, [-1], [1], [D], [1], [D], [1], [D], [1], 1e-6, 3) == inv      # n
    assert f(cpu_handle, [4], [-1], [D], [4], [D], [4], [D], [4], 1e-6, 3) == inv      # nrhs
    assert f(cpu_handle, [4], [1], [D], [3], [D], [4], [D], [4], 1e-6, 3) == inv       # lda
    assert f(cpu_handle, [4], [1], [D], [4], [D], [3], [D], [4], 1e-6, 3) == inv       # ldb
    assert f(cpu_handle, [4], [1], [D], [4], [D], [4], [D], [3], 1e-6, 3) == inv       # ldx
    assert f(cpu_handle, [4], [1], [None], [4], [D], [4], [D], [4], 1e-6, 3) == inv    # A
    assert f(cpu_handle, [257], [1], [D], [257], [D], [257], [D], [257], 1e-6, 3) == ns
    assert f(cpu_handle, [4], [17], [D], [4], [D], [4], [D], [4], 1e-6, 3) == ns
    assert f(cpu_handle, [], [], [], [], [], [], [], [], 1e-6, 3) == status_code("SUCCESS")
    g = B.sgesv_batched_mixed_ex
    for mode in B.MIXED_UPDATE.values():
        assert g(cpu_handle, [4], [1], [D], [3], [D], [4], [D], [4], 1e-6, 3, mode) == inv      # lda
    assert g(cpu_handle, [4], [1], [D], [4], [D], [4], [D], [4], 1e-6, 3, 4) == status_code("INVALID_ENUM")
    assert g(cpu_handle, [4], [1], [D], [4], [D], [4], [D], [4], 1e-6, 3, -1) == status_code("INVALID_ENUM")
    assert g(None, [4], [1], [D], [4], [D], [4], [D], [4], 1e-6, 3, 2) == status_code("HANDLE_IS_NULLPTR")
    assert g(cpu_handle, [], [], [], [], [], [], [], [], 1e-6, 3, 0) == status_code("SUCCESS")


def test_set_launch_events_argument_checks(cpu_handle):
    import ctypes
    assert B.set_launch_events(None, None, 0) == status_code("HANDLE_IS_NULLPTR")
    assert B.set_launch_events(cpu_handle, None, -1) == status_code("INVALID_VALUE")
    assert B.set_launch_events(cpu_handle, None, 2) == status_code("INVALID_VALUE")
    arr = (ctypes.c_void_p * 2)(None, None)
    assert B.set_launch_events(cpu_handle, arr, 2) == status_code("SUCCESS")
    assert B.set_launch_events(cpu_handle, None, 0) == status_code("SUCCESS")


def test_gemm_ex_scaled_and_quantizer_argument_checks(cpu_handle):
    inv = status_code("INVALID_VALUE")
    ns = status_code("NOT_SUPPORTED")
    e4, e5 = B.DTYPES["FP8_E4M3FN"], B.DTYPES["FP8_E5M2"]
    f = B.gemm_ex_scaled
    D = DUMMY
    assert f(None, B.OP_N, B.OP_N, 4, 4, 4, 1.0, D, e4, 4, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == \
        status_code("HANDLE_IS_NULLPTR")
    assert f(cpu_handle, 999, B.OP_N, 4, 4, 4, 1.0, D, e4, 4, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, -1, 4, 4, 1.0, D, e4, 4, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, D, e4, 3, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, None, D, e4, 4, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == inv
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, D, B.ACL_BF16, 4, D, D, e4, 4, D, 0.0, D, B.ACL_FLOAT16,
             4) == ns
    assert f(cpu_handle, B.OP_N, B.OP_N, 4, 4, 4, 1.0, D, e4, 4, D, D, e5, 4, D, 0.0, D, 2, 4) == ns
    assert f(cpu_handle, B.OP_N, B.OP_N, 0, 4, 4, 1.0, D, e4, 4, D, D, e5, 4, D, 0.0, D, B.ACL_FLOAT16, 4) == \
        status_code("SUCCESS")
    q = B.quantize_fp8
    assert q(None, 4, D, e4, D, D) == status_code("HANDLE_IS_NULLPTR")
    assert q(cpu_handle, -1, D, e4, D, D) == inv
    assert q(cpu_handle, 4, D, B.ACL_BF16, D, D) == inv
    assert q(cpu_handle, 4, D, e4, D, None) == inv


STRSM_ROWS = [r for r in load_cases("strsm.csv") if status_code(r["expect_result"]) != 0]


@pytest.mark.parametrize("row", STRSM_ROWS, ids=[r["case_name"] for r in STRSM_ROWS])
def test_strsm_argument_rows(cpu_handle, row):
    """The reference table's error rows (test/trsm/strsm/arch35/strsm_test.csv) on a handle without a device."""
    side = B.SIDE.get(row["side"], None) or int(row["side"])
    uplo = B.FILL.get(row["uplo"], None) or int(row["uplo"])
    trans = {"N": B.OP_N, "T": B.OP_T, "C": B.OP_C}.get(row["trans"], None) or int(row["trans"])
    diag = B.DIAG.get(row["diag"], None) or int(row["diag"])
    st = B.strsm(cpu_handle, side, uplo, trans, diag, int(row["m"]), int(row["n"]), float(row["alpha"]), DUMMY,
                 int(row["lda"]), DUMMY, int(row["ldb"]))
    assert st == status_code(row["expect_result"])


def test_strsm_argument_order(cpu_handle):
    """Check order of reference blas/trsm/arch35/strsm_host.cpp:406-452."""
    f = B.strsm
    L, R = B.SIDE["LEFT"], B.SIDE["RIGHT"]
    lo, up = B.FILL["LOWER"], B.FILL["UPPER"]
    nu = B.DIAG["NON_UNIT"]
    inv = status_code("INVALID_VALUE")
    D = DUMMY
    assert f(None, 999, lo, B.OP_N, nu, 4, 4, 1.0, D, 4, D, 4) == status_code("HANDLE_IS_NULLPTR")
    assert f(cpu_handle, 999, lo, B.OP_N, nu, -1, 4, 1.0, D, 4, D, 4) == inv          # side before m
    assert f(cpu_handle, L, 999, B.OP_N, nu, 0, 4, 1.0, D, 4, D, 4) == inv            # uplo before quick return
    assert f(cpu_handle, L, lo, B.OP_N, nu, 0, 4, None, D, 4, D, 4) == inv            # alpha before quick return
    assert f(cpu_handle, L, lo, B.OP_N, nu, 0, 4, 1.0, None, 0, None, 0) == status_code("SUCCESS")
    assert f(cpu_handle, L, lo, B.OP_N, nu, 4, 4, 1.0, D, 4, None, 4) == inv          # B
    assert f(cpu_handle, L, lo, B.OP_N, nu, 4, 4, 1.0, None, 4, D, 4) == inv          # A when alpha != 0
    assert f(cpu_handle, R, up, B.OP_T, nu, 8, 4, 1.0, D, 3, D, 8) == inv             # right: lda >= n
    assert f(cpu_handle, R, up, B.OP_T, nu, 8, 4, 1.0, D, 4, D, 8) == status_code("EXECUTION_FAILED")
    assert f(cpu_handle, L, lo, B.OP_C, B.DIAG["UNIT"], 4, 4, 0.0, None, 4, D, 4) == \
        status_code("EXECUTION_FAILED")   # alpha = 0 needs no A; no device here


def test_quantize_no_alloc_argument_checks(cpu_handle):
    """aclblasQuantizeFp8NoAlloc checks its arguments like aclblasQuantizeFp8 (no device here: valid arguments end
    at EXECUTION_FAILED before any launch)."""
    q = B.quantize_fp8_no_alloc
    D = DUMMY
    e4 = B.DTYPES["FP8_E4M3FN"]
    assert q(None, 4, D, e4, D, D) == status_code("HANDLE_IS_NULLPTR")
    assert q(cpu_handle, -1, D, e4, D, D) == status_code("INVALID_VALUE")
    assert q(cpu_handle, 4, D, B.ACL_BF16, D, D) == status_code("INVALID_VALUE")
    assert q(cpu_handle, 4, D, e4, D, None) == status_code("INVALID_VALUE")
    assert q(cpu_handle, 4, None, e4, D, D) == status_code("INVALID_VALUE")
    assert q(cpu_handle, 4, D, e4, D, D) == status_code("EXECUTION_FAILED")

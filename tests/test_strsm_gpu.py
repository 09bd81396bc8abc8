"""GPU parity of aclblasStrsm against the oracle's netlib-order FP32 strsm (the reference's golden,
test/trsm/strsm/strsm_golden.h:19-36) on every success row of the reference's table
(test/trsm/strsm/arch35/strsm_test.csv), with the reference's inputs and bar
(test/trsm/strsm/arch35/strsm_test.cpp:52-105):
  A: uniform [-1, 1] from mt19937(seed ? seed : 42), max(dim, lda) * dim values, the other triangle zeroed and
     the diagonal pushed away from zero by max(5, dim) (not for the ill-conditioned fill);
  B: uniform [-1, 1] (or all Inf / all NaN) from seed + 1, max(m, ldb) * n values;
  MERE / MARE over the whole ldb x n buffer with the row's threshold and multiplier.
Plus: alpha = 0 (zeros, A never read), alpha scaling, both sides with row padding, and the argument rows on the
GPU handle.
"""
import numpy as np
import pytest
import torch

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host

pytestmark = pytest.mark.gpu

OPS = {"N": B.OP_N, "T": B.OP_T, "C": B.OP_C}


def _enum(table, v):
    return table[v] if v in table else int(v)


def _inputs(row):
    side, uplo = row["side"], row["uplo"]
    m, n, lda, ldb = int(row["m"]), int(row["n"]), int(row["lda"]), int(row["ldb"])
    seed = int(row["random_seed"])
    dim = m if side == "LEFT" else n
    a = O.fill_uniform(max(dim, lda) * dim, seed, -1.0, 1.0) if dim > 0 else np.zeros(1, np.float32)
    upper = uplo == "UPPER"
    for j in range(dim):
        for i in range(dim):
            if (i > j) if upper else (i < j):
                a[i + j * lda] = 0.0
    if row["a"] != "RANDOM_ILLCOND":
        boost = np.float32(max(5.0, float(dim)))
        for i in range(dim):
            d = a[i + i * lda]
            a[i + i * lda] = d + boost if d >= 0 else d - boost
    count = max(m, ldb) * n
    if row["b"] == "VALUE_INF":
        b = np.full(count, np.inf, np.float32)
    elif row["b"] == "VALUE_NAN":
        b = np.full(count, np.nan, np.float32)
    else:
        b = O.fill_uniform(count, seed + 1, -1.0, 1.0)
    return a, b


def _call(handle, row, a, b, alpha=None):
    m, n, lda, ldb = int(row["m"]), int(row["n"]), int(row["lda"]), int(row["ldb"])
    alpha = float(row["alpha"]) if alpha is None else alpha
    dA, dB = to_dev(a), to_dev(b)
    st = B.strsm(handle.value, _enum(B.SIDE, row["side"]), _enum(B.FILL, row["uplo"]),
                 OPS.get(row["trans"], None) if row["trans"] in OPS else int(row["trans"]),
                 _enum(B.DIAG, row["diag"]), m, n, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb)
    return st, to_host(dB)


def _gold(row, a, b, alpha=None):
    m, n, lda, ldb = int(row["m"]), int(row["n"]), int(row["lda"]), int(row["ldb"])
    alpha = float(row["alpha"]) if alpha is None else alpha
    g = b.copy()
    O.trsm(0 if row["side"] == "LEFT" else 1, row["uplo"] == "UPPER", row["trans"] != "N",
           row["diag"] == "UNIT", m, n, alpha, a, lda, g, ldb)
    return g


ROWS = [r for r in load_cases("strsm.csv")]
SUCCESS = [r for r in ROWS if status_code(r["expect_result"]) == 0]
ERRORS = [r for r in ROWS if status_code(r["expect_result"]) != 0]


@pytest.mark.parametrize("row", SUCCESS, ids=[r["case_name"] for r in SUCCESS])
def test_strsm_reference_cases(handle, row):
    a, b = _inputs(row)
    st, out = _call(handle, row, a, b)
    assert st == 0
    m, n, ldb = int(row["m"]), int(row["n"]), int(row["ldb"])
    if m == 0 or n == 0:
        assert np.array_equal(out, b, equal_nan=True)
        return
    gold = _gold(row, a, b)
    count = ldb * n
    ok, detail = O.mere_mare(out[:count], gold[:count], float(row["mere_threshold"]), float(row["mare_multiplier"]))
    assert ok, detail
    # rows between m and ldb untouched
    assert np.array_equal(out[:count].reshape(n, ldb)[:, m:], b[:count].reshape(n, ldb)[:, m:], equal_nan=True)


@pytest.mark.parametrize("row", ERRORS, ids=[r["case_name"] for r in ERRORS])
def test_strsm_argument_rows_on_gpu(handle, row):
    m, n = int(row["m"]), int(row["n"])
    a = np.zeros(max(1, 64 * 64), np.float32)
    b = np.zeros(max(1, 64 * 64), np.float32)
    st, out = _call(handle, row, a, b)
    assert st == status_code(row["expect_result"])
    assert not out.any()
    del m, n


@pytest.mark.parametrize("side", ["LEFT", "RIGHT"])
@pytest.mark.parametrize("uplo", ["LOWER", "UPPER"])
@pytest.mark.parametrize("trans", ["N", "T"])
@pytest.mark.parametrize("diag", ["NON_UNIT", "UNIT"])
def test_strsm_blocked_sizes_and_alpha(handle, side, uplo, trans, diag):
    """Orders across three 128-blocks with a ragged last block (GEMM updates between the blocks), padded
    leading dimensions, alpha = -0.5.  With a unit diagonal the off-diagonal entries are scaled by 1/dim so the
    triangle stays well conditioned (a unit triangle of uniform [-1, 1] entries amplifies rounding
    differences exponentially in its order, which no summation order survives)."""
    m, n = (300, 70) if side == "LEFT" else (70, 300)
    dim = m if side == "LEFT" else n
    lda = dim + 5
    row = {"side": side, "uplo": uplo, "trans": trans, "diag": diag, "m": str(m), "n": str(n),
           "lda": str(lda), "ldb": str(m + 3), "alpha": "-0.5", "a": "RANDOM_NORM", "b": "RANDOM_NORM",
           "random_seed": "77"}
    a, b = _inputs(row)
    if diag == "UNIT":
        for j in range(dim):
            col = a[j * lda:j * lda + dim].copy()
            d = col[j]
            a[j * lda:j * lda + dim] = col / np.float32(dim)
            a[j * lda + j] = d
    st, out = _call(handle, row, a, b)
    assert st == 0
    gold = _gold(row, a, b)
    ok, detail = O.mere_mare(out, gold, 2.0 ** -13, 10.0)
    assert ok, detail


def test_strsm_alpha_zero_reads_no_a(handle):
    """alpha = 0: B becomes zero (NaN included) and A may be NULL."""
    m, n, ldb = 40, 30, 43
    b = np.full(ldb * n, np.nan, np.float32)
    dB = to_dev(b)
    st = B.strsm(handle.value, B.SIDE["LEFT"], B.FILL["LOWER"], B.OP_N, B.DIAG["NON_UNIT"], m, n, 0.0, None, m,
                 dB.data_ptr(), ldb)
    assert st == 0
    out = to_host(dB).reshape(n, ldb)
    assert not out[:, :m].any()
    assert np.isnan(out[:, m:]).all()


def test_strsm_graph_capture(handle):
    """The blocked solve (diagonal kernels and GEMM updates) captures into a HIP graph; replays equal eager calls."""
    row = {"side": "LEFT", "uplo": "UPPER", "trans": "N", "diag": "NON_UNIT", "m": "300", "n": "90",
           "lda": "300", "ldb": "300", "alpha": "1.5", "a": "RANDOM_NORM", "b": "RANDOM_NORM", "random_seed": "5"}
    a, b = _inputs(row)
    dA, dB0 = to_dev(a), to_dev(b)
    dB, dB2 = dB0.clone(), dB0.clone()
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        def call(x):
            return B.strsm(h.value, B.SIDE["LEFT"], B.FILL["UPPER"], B.OP_N, B.DIAG["NON_UNIT"], 300, 90, 1.5,
                           dA.data_ptr(), 300, x.data_ptr(), 300)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream, capture_error_mode="relaxed"):
            assert call(dB) == 0
        for rep in range(3):
            with torch.cuda.stream(stream):
                dB.copy_(dB0 * (rep + 1))
                dB2.copy_(dB0 * (rep + 1))
                g.replay()
                assert call(dB2) == 0
            stream.synchronize()
            assert torch.equal(dB, dB2), rep
    finally:
        torch.cuda.synchronize()
        h.destroy()

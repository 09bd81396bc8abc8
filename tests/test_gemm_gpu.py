"""GPU parity of aclblasSgemm / SgemmEx / GemmEx against the oracle.

Inputs follow the reference's tests (makeBlasMatrix fills, seeds A = s, B = s+1,
C = s+2; reference test/sgemm_ex/arch35/sgemm_ex_test.cpp:59-66) and are
quantised to the operand type first, so GPU and oracle see identical values.
The oracle accumulates in FP64 and rounds once to Ctype; the bar is the
reference's MERE/MARE criterion (test/frame/verify.h:153-208, 492-516):
mean relative error < 2^-13 (FP32), 2^-10 (FP16), 2^-7 (BF16) and maximum
relative error < 10x that.
"""
import numpy as np
import pytest

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host

pytestmark = pytest.mark.gpu

MERE = {"FP32": 2.0 ** -13, "FP16": 2.0 ** -10, "BF16": 2.0 ** -7}


def parse_fill(s, count, seed):
    """BlasFillMode strings of the reference (test/frame/fill.h:44-180) used by the GEMM tables."""
    parts = s.split("_")
    method = parts[0]
    nums = [p for p in parts[1:] if p not in ("NORM",)]
    if method == "NULLPTR":
        return None
    if method == "VALUE":
        v = nums[0] if nums else "0"
        val = {"NAN": np.nan, "INF": np.inf, "-INF": -np.inf}.get(v.upper(), None)
        val = float(v) if val is None else val
        return np.full(count, val, np.float32)
    if method == "RANDOM":
        lo = float(nums[0]) if nums else 1.0
        hi = float(nums[1]) if len(nums) > 1 else lo
        return O.fill_uniform(count, seed, -lo, hi)
    raise ValueError(s)


def quantize(x, dtype):
    """float32 host values -> (raw array of the operand type, float32 values it represents)."""
    if dtype == "FP32":
        return x.astype(np.float32), x.astype(np.float32)
    if dtype == "BF16":
        q = O.to_bf16(x)
        return q, O.from_bf16(q)
    if dtype == "FP16":
        q = O.to_f16(x)
        return q, O.from_f16(q)
    e4 = dtype == "FP8_E4M3FN"
    q = O.to_fp8(x, e4)
    return q, O.from_fp8(q, e4)


def run_gemm(handle, row, api):
    m, n, k = int(row["m"]), int(row["n"]), int(row["k"])
    ta, tb = row["transA"], row["transB"]
    rowsA, colsA = (m, k) if ta == "N" else (k, m)
    rowsB, colsB = (k, n) if tb == "N" else (n, k)
    if api in ("sgemm_ex", "sgemm"):
        at = bt = ct = "FP32"
        lda, ldb, ldc = int(row["lda"]), int(row["ldb"]), int(row["ldc"])
        seed = int(row["random_seed"])
    else:
        at, bt, ct = row["Atype"], row["Btype"], row["Ctype"]
        lda = int(row["lda"]) if row["lda"] else max(1, rowsA)
        ldb = int(row["ldb"]) if row["ldb"] else max(1, rowsB)
        ldc = int(row["ldc"]) if row["ldc"] else max(1, m)
        seed = int(row.get("random_seed") or 0)
    alpha, beta = float(row["alpha"]), float(row["beta"])
    a32 = parse_fill(row["a_fill"], lda * colsA, seed)
    b32 = parse_fill(row["b_fill"], ldb * colsB, seed + 1)
    c32 = parse_fill(row["c_fill"], ldc * n, seed + 2)
    aq, av = quantize(a32, at)
    bq, bv = quantize(b32, bt)
    cq, cv = quantize(c32, ct)
    dA, dB, dC = to_dev(aq), to_dev(bq), to_dev(cq)
    top = {"N": B.OP_N, "T": B.OP_T, "C": B.OP_C}
    if api == "sgemm":
        st = B.sgemm(handle.value, top[ta], top[tb], m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb,
                     beta, dC.data_ptr(), ldc)
    elif api == "sgemm_ex":
        st = B.sgemm_ex(handle.value, top[ta], top[tb], m, n, k, alpha, dA.data_ptr(), lda, dB.data_ptr(), ldb,
                        beta, dC.data_ptr(), ldc)
    else:
        st = B.gemm_ex(handle.value, top[ta], top[tb], m, n, k, alpha, dA.data_ptr(), B.DTYPES[at], lda,
                       dB.data_ptr(), B.DTYPES[bt], ldb, beta, dC.data_ptr(), B.DTYPES[ct], ldc,
                       B.COMPUTE[row["computeType"]])
    assert st == status_code(row["expect_result"])
    out = to_host(dC)
    gold = cq.copy()
    dt = {"FP32": O.DT_F32, "FP16": O.DT_F16, "BF16": O.DT_BF16}[ct]
    O.gemm(ta != "N", tb != "N", m, n, k, alpha, av, O.DT_F32, lda, bv, O.DT_F32, ldb, beta, gold, dt, ldc)
    if at.startswith("FP8"):
        # reference gemm_ex_golden.h QuantizeGoldenOutput: non-finite FP8 results are NaN
        gf = _as_f32(gold, ct)
        bad = ~np.isfinite(_fp64_result(av, bv, cv, m, n, k, lda, ldb, ldc, ta, tb, alpha, beta))
        gf = gf.reshape(n, ldc)
        gf[:, :m][bad] = np.nan
        gold = _from_f32(gf.reshape(-1), ct)
    return _as_f32(out, ct).reshape(n, ldc)[:, :m], _as_f32(gold, ct).reshape(n, ldc)[:, :m], ct, k


def _fp64_result(av, bv, cv, m, n, k, lda, ldb, ldc, ta, tb, alpha, beta):
    A = av.reshape(-1, lda)[:, : (m if ta == "N" else k)].astype(np.float64).T
    Bm = bv.reshape(-1, ldb)[:, : (k if tb == "N" else n)].astype(np.float64).T
    opA = A if ta == "N" else A.T
    opB = Bm if tb == "N" else Bm.T
    with np.errstate(all="ignore"):
        r = alpha * (opA @ opB).astype(np.float32)
        if beta != 0:
            r = r + beta * cv.reshape(n, ldc)[:, :m].T
    return r.T


def _as_f32(raw, ct):
    if ct == "FP32":
        return raw.astype(np.float32)
    if ct == "BF16":
        return O.from_bf16(raw)
    return O.from_f16(raw)


def _from_f32(x, ct):
    if ct == "FP32":
        return x
    if ct == "BF16":
        return O.to_bf16(x)
    return O.to_f16(x)


SGEMM = [r for r in load_cases("sgemm_ex.csv") if status_code(r["expect_result"]) == 0
         and int(r["m"]) > 0 and int(r["n"]) > 0]


@pytest.mark.parametrize("row", SGEMM, ids=[r["case_name"] for r in SGEMM])
def test_sgemm_ex_reference_cases(handle, row):
    out, gold, ct, k = run_gemm(handle, row, "sgemm_ex")
    ok, detail = O.mere_mare(out, gold, MERE[ct])
    assert ok, detail


GEMMEX = [r for r in load_cases("gemm_ex.csv") if status_code(r["expect_result"]) == 0
          and int(r["m"]) > 0 and int(r["n"]) > 0]


@pytest.mark.parametrize("row", GEMMEX, ids=[r["case_name"] for r in GEMMEX])
def test_gemm_ex_reference_cases(handle, row):
    out, gold, ct, k = run_gemm(handle, row, "gemm_ex")
    mult = 10.0
    if ct == "FP16" and row["computeType"] == "COMPUTE_16F" and k > 160:
        mult = max(10.0, k / 3.0)
    ok, detail = O.mere_mare(out, gold, MERE[ct], mult)
    assert ok, detail


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("dt", ["FP32", "BF16_F32", "FP16_F32", "E4M3"])
@pytest.mark.parametrize("shape", [(300, 257, 129), (64, 64, 512), (1024, 768, 640)])
def test_gemm_layouts_and_types(handle, ta, tb, dt, shape):
    """Every transpose combination and operand type, ragged and full tiles, beta != 0."""
    m, n, k = shape
    at, ct = {"FP32": ("FP32", "FP32"), "BF16_F32": ("BF16", "FP32"), "FP16_F32": ("FP16", "FP32"),
              "E4M3": ("FP8_E4M3FN", "FP16")}[dt]
    row = {"m": m, "n": n, "k": k, "transA": ta, "transB": tb, "Atype": at, "Btype": at, "Ctype": ct,
           "lda": "", "ldb": "", "ldc": str(m + 3), "alpha": "0.75", "beta": "-0.5", "a_fill": "RANDOM_1",
           "b_fill": "RANDOM_1", "c_fill": "RANDOM_1", "computeType": "COMPUTE_32F",
           "expect_result": "ACLBLAS_STATUS_SUCCESS", "random_seed": "7"}
    out, gold, ct, _ = run_gemm(handle, row, "gemm_ex")
    # relative to the output scale (MERE/MARE is ill-posed near cancellation)
    scale = np.abs(gold).max()
    tol = {"FP32": 2e-6, "FP16": 2e-3, "BF16": 1e-2}[ct] * max(1.0, np.sqrt(k / 64))
    err = np.abs(out.astype(np.float64) - gold).max() / scale
    assert err < tol, err


def test_sgemm_plain_entry_c1_shape(handle):
    """aclblasSgemm 512^3 NN, alpha = 1, beta = 0.5, U[-1,1] seeds 0/1/2 (the FP32 GEMM workload)."""
    row = {"m": "512", "n": "512", "k": "512", "transA": "N", "transB": "N", "lda": "512", "ldb": "512",
           "ldc": "512", "alpha": "1.0", "beta": "0.5", "a_fill": "RANDOM_1", "b_fill": "RANDOM_1",
           "c_fill": "RANDOM_1", "random_seed": "0", "expect_result": "ACLBLAS_STATUS_SUCCESS"}
    out, gold, ct, k = run_gemm(handle, row, "sgemm")
    err = np.abs(out.astype(np.float64) - gold).max() / np.abs(gold).max()
    assert err < 1e-5, err

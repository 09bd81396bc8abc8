"""GPU parity of aclblasLtMatmul (FP32, MXFP8, MXFP4; row-major) through the C ABI.

Every success row of the reference's table (test/blasLtMatmul/arch35/blasLtMatmul_test.csv) with work, inputs
as its test builds them (test/blasLtMatmul/arch35/blasLtMatmul_test.cpp:24-52, 200-265): FP32 operands uniform
in +-1/sqrt(K) from the reference generator with the case's seed for A, B and C alike, or A = 1/K, B = 1, C = 0
when K >= 128; MXFP8 bytes uniform in [0x38, 0x4F], MXFP4 nibbles in [3, 8] (two per byte, low nibble first),
E8M0 scales in [127, 130] or all zero for a VALUE_NORM_0 fill - drawn here from numpy with the same ranges (the
reference's std::uniform_int_distribution stream is not reproduced).  FP4 layouts get the bytes per row, as
the reference's wrapper sets them (blasLtMatmul_npu_wrapper.h:34-47).  The golden is the oracle's restatement
of the reference's (blasLtMatmul_golden.h), and the bar is the reference's: MERE / MARE with 2^-13 for FP32
and 2^-7 for MX, multiplier 10 (blasLtMatmul_test.cpp:358-383), BF16 results compared after widening.

Beyond the table: every transpose pair of both MX types at 200 x 136 x 512 with alpha, beta and C, with
unaligned leading dimensions (the kernel's byte-wise copies) and with zero scale blocks; the FP32 path with
C and D distinct and different leading dimensions.
"""
import numpy as np
import pytest

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host

pytestmark = pytest.mark.gpu

TYPES = {"FP32": (B.ACL_FLOAT, O.LT_FP32), "BF16": (B.ACL_BF16, None),
         "MXFP8_E4M3FN": (B.ACL_FLOAT8_E4M3FN, O.LT_MXFP8), "MXFP4_E2M1": (B.ACL_FLOAT4_E2M1, O.LT_MXFP4)}

ROWS = [r for r in load_cases("blaslt_matmul.csv")
        if status_code(r["expect_result"]) == 0 and int(r["M"]) > 0 and int(r["N"]) > 0]


@pytest.fixture(scope="module")
def lt():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.init()   # torch first: its runtime owns the device before the library's handle is made
    st, h = B.lt_create()
    assert st == 0
    yield h
    B.load().aclblasLtDestroy(h)


def _scale_bytes(rows, K, kpairs):
    return (((K + 63) // 64) * rows * 2) if kpairs else rows * 2 * ((K + 63) // 64)


def _mx_operand(rng, rows, cols, fp4):
    """Stored operand (rows x cols elements, row-major) and its bytes per row."""
    if fp4:
        nib = rng.integers(3, 9, size=(rows, cols + (cols & 1))).astype(np.uint8)
        packed = (nib[:, 0::2] | (nib[:, 1::2] << 4)).astype(np.uint8)
        return np.ascontiguousarray(packed), packed.shape[1]
    b = rng.integers(0x38, 0x50, size=(rows, cols)).astype(np.uint8)
    return b, cols


def _run(lt, typ, dt_out, ta, tb, M, N, K, alpha, beta, A, lda, Bm, ldb, sA, sB, C, ldc, ldd):
    """aclblasLtMatmul on the device; returns (status, D as float32 m x ldd)."""
    import torch
    at = {O.LT_FP32: B.ACL_FLOAT, O.LT_MXFP8: B.ACL_FLOAT8_E4M3FN, O.LT_MXFP4: B.ACL_FLOAT4_E2M1}[typ]
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    dA, dB = to_dev(A), to_dev(Bm)
    dsa = to_dev(sA) if sA is not None else None
    dsb = to_dev(sB) if sB is not None else None
    dC = to_dev(C) if C is not None else None
    bf16 = dt_out == B.ACL_BF16
    dD = torch.zeros((M, ldd), dtype=torch.bfloat16 if bf16 else torch.float32, device="cuda")
    L = B.load()
    lays = [B.lt_layout(at, ra, ca, lda)[1], B.lt_layout(at, rb, cb, ldb)[1],
            B.lt_layout(B.ACL_FLOAT, M, N, ldc)[1], B.lt_layout(dt_out, M, N, ldd)[1]]
    st, d = B.lt_desc(transA=B.OPS[ta], transB=B.OPS[tb], scaleA=dsa.data_ptr() if dsa is not None else None,
                      scaleB=dsb.data_ptr() if dsb is not None else None)
    assert st == 0
    try:
        stream = torch.cuda.current_stream().cuda_stream
        st = B.lt_matmul(lt, d, alpha, dA.data_ptr(), lays[0], dB.data_ptr(), lays[1], beta,
                         dC.data_ptr() if dC is not None else None, lays[2], dD.data_ptr(), lays[3], stream=stream)
        torch.cuda.synchronize()
        out = dD.float().cpu().numpy()
    finally:
        for x in lays:
            L.aclblasLtMatrixLayoutDestroy(x)
        L.aclblasLtMatmulDescDestroy(d)
    return st, out


def _check(out, gold, M, N, typ):
    thr = 2.0 ** -13 if typ == O.LT_FP32 else 2.0 ** -7
    ok, detail = O.mere_mare(out[:, :N], gold[:, :N], thr)
    assert ok, detail


@pytest.mark.parametrize("row", ROWS, ids=[r["case_name"] for r in ROWS])
def test_reference_cases(lt, row):
    at, typ = TYPES[row["dtypeA"]]
    dt_out = TYPES[row["dtypeD"]][0]
    M, N, K = int(row["M"]), int(row["N"]), int(row["K"])
    ta, tb = row["transA"], row["transB"]
    lda, ldb, ldc, ldd = int(row["lda"]), int(row["ldb"]), int(row["ldc"]), int(row["ldd"])
    alpha, beta = float(row["alpha"]), float(row["beta"])
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    seed = int(row.get("random_seed") or 0)
    sA = sB = None
    if typ == O.LT_FP32:
        if K >= 128:
            A = np.full(ra * lda, 1.0 / K, np.float32)
            Bm = np.ones(rb * ldb, np.float32)
            C = np.zeros(M * ldc, np.float32)
        else:
            bound = 1.0 / np.sqrt(np.float32(max(K, 1)))
            A = O.fill_uniform(ra * lda, seed, -bound, bound)
            Bm = O.fill_uniform(rb * ldb, seed, -bound, bound)
            C = O.fill_uniform(M * ldc, seed, -bound, bound)
    else:
        rng = np.random.default_rng(int(row["case_name"].rsplit("_", 1)[-1]))
        fp4 = typ == O.LT_MXFP4
        A, lda_b = _mx_operand(rng, ra, lda, fp4)
        Bm, ldb_b = _mx_operand(rng, rb, ldb, fp4)
        lda, ldb = lda_b, ldb_b
        C = rng.uniform(-1, 1, size=M * ldc).astype(np.float32)
        na = _scale_bytes(M, K, ta != "N")
        nb = _scale_bytes(N, K, tb == "N")
        zero_a = row["scaleA_fill"].startswith("VALUE_NORM_0")
        zero_b = row["scaleB_fill"].startswith("VALUE_NORM_0")
        sA = np.zeros(na, np.uint8) if zero_a else rng.integers(127, 131, size=na).astype(np.uint8)
        sB = np.zeros(nb, np.uint8) if zero_b else rng.integers(127, 131, size=nb).astype(np.uint8)
    Cuse = None if row["C_is_null"] == "true" else C
    st, out = _run(lt, typ, dt_out, ta, tb, M, N, K, alpha, beta, A, lda, Bm, ldb, sA, sB, Cuse, ldc, ldd)
    assert st == 0, st
    gold = O.lt_matmul(typ, dt_out == B.ACL_BF16, ta != "N", tb != "N", M, N, K, alpha, A, lda, Bm, ldb, sA, sB,
                       beta, Cuse, ldc, ldd)
    _check(out, gold, M, N, typ)


@pytest.mark.parametrize("fp4", [False, True], ids=["mxfp8", "mxfp4"])
@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("out", ["FP32", "BF16"])
def test_mx_larger(lt, fp4, trans, out):
    rng = np.random.default_rng(100 * int(fp4) + 10 * ["NN", "NT", "TN", "TT"].index(trans) + (out == "BF16"))
    M, N, K = 200, 136, 512
    ta, tb = trans[0], trans[1]
    typ = O.LT_MXFP4 if fp4 else O.LT_MXFP8
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    A, lda = _mx_operand(rng, ra, ca, fp4)
    Bm, ldb = _mx_operand(rng, rb, cb, fp4)
    sA = rng.integers(120, 135, size=_scale_bytes(M, K, ta != "N")).astype(np.uint8)
    sB = rng.integers(120, 135, size=_scale_bytes(N, K, tb == "N")).astype(np.uint8)
    ldc, ldd = N + 3, N + 5
    C = rng.uniform(-1, 1, size=M * ldc).astype(np.float32)
    dt_out = B.ACL_BF16 if out == "BF16" else B.ACL_FLOAT
    st, res = _run(lt, typ, dt_out, ta, tb, M, N, K, 0.75, 0.5, A, lda, Bm, ldb, sA, sB, C, ldc, ldd)
    assert st == 0
    gold = O.lt_matmul(typ, out == "BF16", ta != "N", tb != "N", M, N, K, 0.75, A, lda, Bm, ldb, sA, sB, 0.5, C,
                       ldc, ldd)
    _check(res, gold, M, N, typ)
    if out == "FP32":
        # the instruction sums exact FP8 / FP4 products: the error stays at FP32 rounding of a 512-term sum
        np.testing.assert_allclose(res[:, :N], gold[:, :N], rtol=2e-5, atol=1e-3)


@pytest.mark.parametrize("fp4", [False, True], ids=["mxfp8", "mxfp4"])
def test_mx_unaligned_rows_and_zero_blocks(lt, fp4):
    """Leading dimensions that are not multiples of 16 bytes (byte-wise copies), odd m and n, and scale bytes of
    zero in a few blocks (those blocks contribute exactly zero)."""
    rng = np.random.default_rng(77 + fp4)
    M, N, K = 67, 45, 160
    typ = O.LT_MXFP4 if fp4 else O.LT_MXFP8
    A, lda = _mx_operand(rng, M, K + 6, fp4)       # rows padded by 6 elements: unaligned pitch
    Bm, ldb = _mx_operand(rng, N, K + 2, fp4)      # B stored n x k (transB T)
    sA = rng.integers(126, 130, size=_scale_bytes(M, K, False)).astype(np.uint8)
    sB = rng.integers(126, 130, size=_scale_bytes(N, K, False)).astype(np.uint8)
    sA[::7] = 0
    sB[3::11] = 0
    st, res = _run(lt, typ, B.ACL_FLOAT, "N", "T", M, N, K, 1.0, 0.0, A, lda, Bm, ldb, sA, sB, None, N, N)
    assert st == 0
    gold = O.lt_matmul(typ, False, 0, 1, M, N, K, 1.0, A, lda, Bm, ldb, sA, sB, 0.0, None, N, N)
    _check(res, gold, M, N, typ)
    np.testing.assert_allclose(res, gold, rtol=2e-5, atol=1e-3)


@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
def test_fp32_distinct_c_and_d(lt, trans):
    rng = np.random.default_rng(5)
    M, N, K = 300, 200, 150
    ta, tb = trans[0], trans[1]
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    lda, ldb, ldc, ldd = ca + 1, cb + 2, N + 7, N + 3
    # non-negative data: no cancellation, so the relative bar measures the FP32 summation alone
    A = rng.uniform(0, 1, size=ra * lda).astype(np.float32)
    Bm = rng.uniform(0, 1, size=rb * ldb).astype(np.float32)
    C = rng.uniform(0, 1, size=M * ldc).astype(np.float32)
    st, res = _run(lt, O.LT_FP32, B.ACL_FLOAT, ta, tb, M, N, K, 1.5, 0.5, A, lda, Bm, ldb, None, None, C, ldc,
                   ldd)
    assert st == 0
    gold = O.lt_matmul(O.LT_FP32, False, ta != "N", tb != "N", M, N, K, 1.5, A, lda, Bm, ldb, None, None, 0.5, C,
                       ldc, ldd)
    _check(res, gold, M, N, O.LT_FP32)
    np.testing.assert_allclose(res[:, :N], gold[:, :N], rtol=1e-5, atol=1e-4)

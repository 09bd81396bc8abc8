"""GPU parity of aclblasSgemmGroupedBatched.

Reference case table test/gemm_grouped_batched/arch35/gemm_grouped_batched_test.csv
(copied as data under tests/golden/cases/): problem b of the flattened batch
is filled with makeBlasMatrix(ld, cols, ld, fill, seed + 3*b) for A and seed +
3*b + 1, + 2 for B and C (gemm_grouped_batched_test.cpp:98-117); the golden is
gemm_grouped_batched_golden.h (alpha == 0: C = beta * C, else sgemm), computed
here in FP64.  Bar: the case's MERE/MARE thresholds, with the FP32 dot-product
error bound as the fallback (see test_gemm_gpu.check_gemm).  Also: the grouped
workload of the benchmark (64 groups, m, n, k in {64, 80, ..., 2048}, seed 0)
and bit-identity with aclblasSgemm on the same problems.
"""
import numpy as np
import pytest
import torch

import aclblas as B
from common import load_cases, status_code, to_dev, to_host
from test_gemm_gpu import check_gemm, parse_fill

pytestmark = pytest.mark.gpu

OPS = {"ACLBLAS_OP_N": B.OP_N, "ACLBLAS_OP_T": B.OP_T, "ACLBLAS_OP_C": B.OP_C}


def _list(s, conv):
    return [conv(x) for x in str(s).split(";")]


def _gold(a, b, c, m, n, k, lda, ldb, ldc, ta, tb, alpha, beta):
    """FP64 reference of one problem; arrays are column-major flat float32."""
    out = c.astype(np.float64).copy().reshape(n, ldc)
    if alpha == 0.0 or k == 0:
        # alpha == 0: beta * C (golden :applyAlphaZeroGolden); k == 0: sgemm with an empty sum
        out[:, :m] = beta * out[:, :m] if (alpha == 0.0 or beta != 0.0) else 0.0
        return out.reshape(-1).astype(np.float32)
    A = a.reshape(-1, lda).astype(np.float64)[:, : (m if ta == B.OP_N else k)].T
    Bm = b.reshape(-1, ldb).astype(np.float64)[:, : (k if tb == B.OP_N else n)].T
    opA = A if ta == B.OP_N else A.T
    opB = Bm if tb == B.OP_N else Bm.T
    res = alpha * (opA @ opB)
    if beta != 0.0:
        res = res + beta * out[:, :m].T
    out[:, :m] = res.T
    return out.reshape(-1).astype(np.float32)


def run_grouped(handle, groups, fills, seed):
    """groups: list of dicts (ta, tb, m, n, k, alpha, beta, lda, ldb, ldc, size).  Returns per-problem records."""
    probs = []
    A_ptrs, B_ptrs, C_ptrs = [], [], []
    b = 0
    for g in groups:
        acols = g["k"] if g["ta"] == B.OP_N else g["m"]
        bcols = g["n"] if g["tb"] == B.OP_N else g["k"]
        for _ in range(g["size"]):
            a = parse_fill(fills[0], max(1, g["lda"] * acols), seed + 3 * b)
            bb = parse_fill(fills[1], max(1, g["ldb"] * bcols), seed + 3 * b + 1)
            c = parse_fill(fills[2], max(1, g["ldc"] * g["n"]), seed + 3 * b + 2)
            dA, dB, dC = to_dev(a), to_dev(bb), to_dev(c)
            probs.append(dict(g=g, a=a, b=bb, c=c, dA=dA, dB=dB, dC=dC))
            A_ptrs.append(dA.data_ptr())
            B_ptrs.append(dB.data_ptr())
            C_ptrs.append(dC.data_ptr())
            b += 1
    st = B.sgemm_grouped_batched(handle.value, [g["ta"] for g in groups], [g["tb"] for g in groups],
                                 [g["m"] for g in groups], [g["n"] for g in groups], [g["k"] for g in groups],
                                 [g["alpha"] for g in groups], A_ptrs, [g["lda"] for g in groups], B_ptrs,
                                 [g["ldb"] for g in groups], [g["beta"] for g in groups], C_ptrs,
                                 [g["ldc"] for g in groups], [g["size"] for g in groups])
    torch.cuda.synchronize()
    return st, probs


ROWS = [r for r in load_cases("gemm_grouped_batched.csv") if status_code(r["expect_result"]) == 0
        and int(r["groupCount"]) > 0]


@pytest.mark.parametrize("row", ROWS, ids=[r["case_name"] for r in ROWS])
def test_grouped_reference_cases(handle, row):
    gc = int(row["groupCount"])
    cols = dict(ta=_list(row["transaArray"], OPS.get), tb=_list(row["transbArray"], OPS.get),
                m=_list(row["mArray"], int), n=_list(row["nArray"], int), k=_list(row["kArray"], int),
                alpha=_list(row["alphaArray"], float), beta=_list(row["betaArray"], float),
                lda=_list(row["ldaArray"], int), ldb=_list(row["ldbArray"], int), ldc=_list(row["ldcArray"], int),
                size=_list(row["groupSizeArray"], int))
    groups = [{key: cols[key][g] for key in cols} for g in range(gc)]
    st, probs = run_grouped(handle, groups, (row["A_fill"], row["B_fill"], row["C_fill"]), int(row["random_seed"]))
    assert st == 0
    thr = float(row["mere_threshold"] or 2.0 ** -13)
    mult = float(row["mare_multiplier"] or 10)
    for idx, p in enumerate(probs):
        g = p["g"]
        m, n, k = g["m"], g["n"], g["k"]
        out = to_host(p["dC"])
        if m == 0 or n == 0:
            assert np.array_equal(out, p["c"])
            continue
        gold = _gold(p["a"], p["b"], p["c"], m, n, k, g["lda"], g["ldb"], g["ldc"], g["ta"], g["tb"], g["alpha"],
                     g["beta"])
        o = out.reshape(n, g["ldc"])[:, :m]
        gg = gold.reshape(n, g["ldc"])[:, :m]
        ref = (p["a"], p["b"], p["c"], m, n, k, g["lda"], g["ldb"], g["ldc"],
               "N" if g["ta"] == B.OP_N else "T", "N" if g["tb"] == B.OP_N else "T", g["alpha"], g["beta"])
        if k == 0 or g["alpha"] == 0.0:
            assert np.array_equal(o, gg)
        else:
            check_gemm(o, gg, "FP32", ref, thr, mult, case="gemm_grouped_batched:%s#%d" % (row["case_name"], idx))
        # rows between m and ldc are untouched
        assert np.array_equal(out.reshape(n, g["ldc"])[:, m:], p["c"].reshape(n, g["ldc"])[:, m:])


def _bench_groups(count=64, seed=0):
    rng = np.random.default_rng(seed)
    mnk = rng.integers(0, 125, size=(count, 3)) * 16 + 64
    return [dict(ta=B.OP_N, tb=B.OP_N, m=int(m), n=int(n), k=int(k), alpha=1.0, beta=0.5, lda=int(m), ldb=int(k),
                 ldc=int(m), size=1) for m, n, k in mnk]


def test_grouped_benchmark_workload_matches_sgemm(handle):
    """64 groups of one problem each; every C bit-identical to aclblasSgemm on the same inputs and within the
    FP32 bound of an FP64 product."""
    groups = _bench_groups()
    st, probs = run_grouped(handle, groups, ("RANDOM_1_1", "RANDOM_1_1", "RANDOM_1_1"), 7)
    assert st == 0
    for idx, p in enumerate(probs):
        g = p["g"]
        m, n, k = g["m"], g["n"], g["k"]
        dC2 = to_dev(p["c"])
        assert B.sgemm(handle.value, B.OP_N, B.OP_N, m, n, k, 1.0, p["dA"].data_ptr(), m, p["dB"].data_ptr(), k, 0.5,
                       dC2.data_ptr(), m) == 0
        torch.cuda.synchronize()
        out = to_host(p["dC"])
        assert np.array_equal(out, to_host(dC2)), "group %d differs from aclblasSgemm" % idx
        if idx % 8 == 0:
            gold = _gold(p["a"], p["b"], p["c"], m, n, k, m, k, m, B.OP_N, B.OP_N, 1.0, 0.5)
            ref = (p["a"], p["b"], p["c"], m, n, k, m, k, m, "N", "N", 1.0, 0.5)
            check_gemm(out.reshape(n, m), gold.reshape(n, m), "FP32", ref, 2.0 ** -13)


def test_grouped_mixed_transposes_and_inactive_groups(handle):
    groups = [dict(ta=B.OP_T, tb=B.OP_N, m=300, n=129, k=77, alpha=-1.25, beta=0.0, lda=80, ldb=77, ldc=301, size=3),
              dict(ta=B.OP_N, tb=B.OP_T, m=0, n=64, k=64, alpha=1.0, beta=1.0, lda=1, ldb=64, ldc=1, size=2),
              dict(ta=B.OP_T, tb=B.OP_T, m=65, n=31, k=513, alpha=0.5, beta=2.0, lda=513, ldb=31, ldc=65, size=2),
              dict(ta=B.OP_N, tb=B.OP_N, m=16, n=16, k=16, alpha=0.0, beta=-3.0, lda=16, ldb=16, ldc=16, size=1),
              dict(ta=B.OP_N, tb=B.OP_N, m=40, n=40, k=0, alpha=2.0, beta=0.25, lda=40, ldb=1, ldc=40, size=1)]
    st, probs = run_grouped(handle, groups, ("RANDOM_1_1", "RANDOM_1_1", "RANDOM_1_1"), 3)
    assert st == 0
    for p in probs:
        g = p["g"]
        m, n, k = g["m"], g["n"], g["k"]
        out = to_host(p["dC"])
        if m == 0:
            assert np.array_equal(out, p["c"])
            continue
        gold = _gold(p["a"], p["b"], p["c"], m, n, k, g["lda"], g["ldb"], g["ldc"], g["ta"], g["tb"], g["alpha"],
                     g["beta"])
        ref = (p["a"], p["b"], p["c"], m, n, k, g["lda"], g["ldb"], g["ldc"],
               "N" if g["ta"] == B.OP_N else "T", "N" if g["tb"] == B.OP_N else "T", g["alpha"], g["beta"])
        o = out.reshape(n, g["ldc"])[:, :m]
        gg = gold.reshape(n, g["ldc"])[:, :m]
        if k == 0 or g["alpha"] == 0.0:
            assert np.array_equal(o, gg)
        else:
            check_gemm(o, gg, "FP32", ref, 2.0 ** -13)

"""GPU parity of aclblasSgemmBatched on every success row of the reference's table
(test/gemm_batched/sgemm_batched/arch35/sgemm_batched_test.csv) that has work.

Inputs as the reference test builds them (test/gemm_batched/sgemm_batched/arch35/sgemm_batched_test.cpp:32-60):
member b uses seeds s + 3b (A), s + 3b + 1 (B), s + 3b + 2 (C), A holds lda x (k or m) values, B ldb x (n or k),
C ldc x n; blank leading dimensions are the minimum (sgemm_batched_param.h:66-72).  The golden is cblas_sgemm per
member, or beta * C when alpha == 0 (sgemm_batched_golden.h:18-50): here the oracle's GEMM with FP64 accumulation
rounded once to FP32, and beta * C in FP32 for alpha == 0.  The bar is the reference's: MIXED_TOLERANCE for FP32
over the whole ldc x n buffer of each member (sgemm_batched_test.cpp:84-106, test/frame/verify.h:234-333).
"""
import numpy as np
import pytest

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host
from test_gemm_gpu import parse_fill

pytestmark = pytest.mark.gpu

OPS = {"N": B.OP_N, "T": B.OP_T, "C": B.OP_C}

ROWS = [r for r in load_cases("sgemm_batched.csv")
        if status_code(r["expect_result"]) == 0 and int(r["m"]) > 0 and int(r["n"]) > 0
        and int(r["batchCount"]) > 0]


def _dims(row):
    m, n, k = int(row["m"]), int(row["n"]), int(row["k"])
    ta, tb = row["transA"], row["transB"]
    rowsA, colsA = (m, k) if ta == "N" else (k, m)
    rowsB, colsB = (k, n) if tb == "N" else (n, k)
    lda = int(row["lda"]) if row["lda"] else max(1, rowsA)
    ldb = int(row["ldb"]) if row["ldb"] else max(1, rowsB)
    ldc = int(row["ldc"]) if row["ldc"] else max(1, m)
    return m, n, k, ta, tb, lda, ldb, ldc, max(1, colsA), max(1, colsB)


@pytest.mark.parametrize("row", ROWS, ids=[r["case_name"] for r in ROWS])
def test_sgemm_batched_reference_cases(handle, row):
    import torch
    m, n, k, ta, tb, lda, ldb, ldc, colsA, colsB = _dims(row)
    batch = int(row["batchCount"])
    alpha, beta = float(row["alpha"]), float(row["beta"])
    seed0 = int(row.get("random_seed") or 0)
    As, Bs, Cs, golds = [], [], [], []
    for b in range(batch):
        seed = seed0 + 3 * b
        a = parse_fill(row["a_fill"], max(1, lda) * colsA, seed)
        bm = parse_fill(row["b_fill"], max(1, ldb) * colsB, seed + 1)
        c = parse_fill(row["c_fill"], max(1, ldc) * n, seed + 2)
        gold = c.copy()
        if alpha == 0.0:
            g2 = gold.reshape(n, ldc)
            g2[:, :m] = np.float32(beta) * g2[:, :m]
        else:
            O.gemm(ta != "N", tb != "N", m, n, k, alpha, a, O.DT_F32, lda, bm, O.DT_F32, ldb, beta, gold,
                   O.DT_F32, ldc)
        As.append(to_dev(a)); Bs.append(to_dev(bm)); Cs.append(to_dev(c)); golds.append(gold)
    pa = torch.tensor([x.data_ptr() for x in As], dtype=torch.int64, device="cuda")
    pb = torch.tensor([x.data_ptr() for x in Bs], dtype=torch.int64, device="cuda")
    pc = torch.tensor([x.data_ptr() for x in Cs], dtype=torch.int64, device="cuda")
    st = B.sgemm_batched(handle.value, OPS[ta], OPS[tb], m, n, k, alpha, pa.data_ptr(), lda, pb.data_ptr(), ldb,
                         beta, pc.data_ptr(), ldc, batch)
    assert st == 0
    for b in range(batch):
        out = to_host(Cs[b])
        ok, detail = O.mixed_tolerance(out, golds[b], "fp32")
        assert ok, "member %d: %s" % (b, detail)

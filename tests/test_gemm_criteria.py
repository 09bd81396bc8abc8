"""CPU checks of the GEMM acceptance criterion used by the GPU tests.

An FP32-accumulated product (computed here with float32 arithmetic in a
different summation order) must pass; a product that drops one term, flips a
sign or transposes an operand must fail.
"""
import numpy as np
import pytest

import oracle as O
from test_gemm_gpu import check_gemm, quantize


def _case(m, n, k, ct, seed=3, lo=1.0):
    a = O.fill_uniform(m * k, seed, -lo, lo)
    b = O.fill_uniform(k * n, seed + 1, -lo, lo)
    _, av = quantize(a, "FP16")
    _, bv = quantize(b, "FP16")
    cv = np.zeros(m * n, np.float32)
    gold = np.zeros(m * n, {"FP32": np.float32, "FP16": np.uint16, "BF16": np.uint16}[ct])
    dt = {"FP32": O.DT_F32, "FP16": O.DT_F16, "BF16": O.DT_BF16}[ct]
    O.gemm(False, False, m, n, k, 1.0, av, O.DT_F32, m, bv, O.DT_F32, k, 0.0, gold, dt, m)
    g = gold.astype(np.float32) if ct == "FP32" else (O.from_f16(gold) if ct == "FP16" else O.from_bf16(gold))
    A = av.reshape(k, m).T
    Bm = bv.reshape(n, k).T
    ref = (av, bv, cv, m, n, k, m, k, m, "N", "N", 1.0, 0.0)
    return A, Bm, g.reshape(n, m), ref


def _round(x, ct):
    if ct == "FP32":
        return x.astype(np.float32)
    if ct == "FP16":
        return O.from_f16(O.to_f16(x.astype(np.float32)))
    return O.from_bf16(O.to_bf16(x.astype(np.float32)))


def _fp32_blocked(A, Bm, blk=32):
    """FP32 accumulation in k-blocks of blk, block sums rounded to FP32 (MFMA-like order)."""
    acc = np.zeros((A.shape[0], Bm.shape[1]), np.float32)
    for k0 in range(0, A.shape[1], blk):
        part = (A[:, k0:k0 + blk].astype(np.float64) @ Bm[k0:k0 + blk].astype(np.float64)).astype(np.float32)
        acc = (acc + part).astype(np.float32)
    return acc


@pytest.mark.parametrize("ct", ["FP32", "FP16", "BF16"])
@pytest.mark.parametrize("shape", [(64, 64, 64), (96, 80, 300)])
def test_fp32_accumulation_passes(ct, shape):
    m, n, k = shape
    A, Bm, gold, ref = _case(m, n, k, ct)
    out = _round(_fp32_blocked(A, Bm), ct).T
    check_gemm(out, gold, ct, ref, {"FP32": 2.0 ** -13, "FP16": 0.0012, "BF16": 2.0 ** -7}[ct])


@pytest.mark.parametrize("ct", ["FP32", "FP16"])
def test_planted_mistakes_fail(ct):
    m, n, k = 64, 64, 64
    thr = {"FP32": 2.0 ** -13, "FP16": 0.0012}[ct]
    A, Bm, gold, ref = _case(m, n, k, ct)
    good = _fp32_blocked(A, Bm)
    # one dropped term in one element
    dropped = good.copy()
    dropped[5, 7] -= np.float32(A[5, 40] * Bm[40, 7])
    # a sign error in one column
    flipped = good.copy()
    flipped[:, 3] = -flipped[:, 3]
    # transposed operand
    swapped = _fp32_blocked(A.T.copy(), Bm)
    for bad in (dropped, flipped, swapped):
        with pytest.raises(AssertionError):
            check_gemm(_round(bad, ct).T, gold, ct, ref, thr)

"""GPU checks of aclblasQuantizeFp8 and aclblasGemmExScaled (FP8 with per-tensor scales; no counterpart in
the reference, so parity is against closed forms computed on the host):
  - the quantiser's scale equals amax / fp8_max and its codes equal the oracle's round-to-nearest-even,
    saturating FP8 conversion of x * (fp8_max / amax), bit for bit;
  - the scaled GEMM equals alpha * sA * sB * qA qB + beta * C computed in FP64 from the FP8 values, within
    the native FP8 instruction's summation error: |out - gold| <= 2^-11 * |alpha| sA sB sum_k |qa qb|
    + |beta C| 2^-22 + one unit in the last place of the output type (measured error of one 128-product
    block: about 2^-14 of its largest product);
  - end to end, quantise + scaled GEMM reproduces the unquantised product within the FP8 quantisation
    error bound (per operand element eps |x| + half a subnormal step times the scale).
"""
import numpy as np
import pytest
import torch

import aclblas as B
import oracle as O
from common import to_dev, to_host

pytestmark = pytest.mark.gpu

FMAX = {"FP8_E4M3FN": 448.0, "FP8_E5M2": 57344.0}


def _quantize(handle, x, qt):
    dx = to_dev(x)
    q = torch.zeros(x.size, dtype=torch.uint8, device="cuda")
    s = torch.zeros(1, dtype=torch.float32, device="cuda")
    assert B.quantize_fp8(handle.value, x.size, dx.data_ptr(), B.DTYPES[qt], q.data_ptr(), s.data_ptr()) == 0
    torch.cuda.synchronize()
    return q.cpu().numpy(), float(s.cpu().numpy()[0]), q, s


@pytest.mark.parametrize("qt", ["FP8_E4M3FN", "FP8_E5M2"])
@pytest.mark.parametrize("count,scale", [(1, 1.0), (1000, 3.0), (1 << 20, 0.01), (300001, 1e4)])
def test_quantizer_matches_closed_form(handle, qt, count, scale):
    rng = np.random.default_rng(count)
    x = (rng.standard_normal(count) * scale).astype(np.float32)
    if count > 10:
        x[5] = 0.0
        x[7] = -x[6]
    q, s, _, _ = _quantize(handle, x, qt)
    amax = np.float32(np.abs(x).max())
    fmax = np.float32(FMAX[qt])
    assert s == np.float32(amax / fmax)
    inv = np.float32(fmax / amax)
    ref = O.to_fp8((x * inv).astype(np.float32), qt == "FP8_E4M3FN")
    assert np.array_equal(q, ref), int((q != ref).sum())


@pytest.mark.parametrize("xoff,qoff", [(1, 0), (0, 1), (3, 2)])
def test_quantizer_unaligned_operands(handle, xoff, qoff):
    """x not 16-byte or q not 4-byte aligned: the element-wise path, same codes."""
    count = 70001
    rng = np.random.default_rng(xoff * 10 + qoff)
    x = (rng.standard_normal(count + xoff) * 7.0).astype(np.float32)
    dx = to_dev(x)
    q = torch.zeros(count + qoff, dtype=torch.uint8, device="cuda")
    s = torch.zeros(1, dtype=torch.float32, device="cuda")
    qt = "FP8_E5M2"
    assert B.quantize_fp8(handle.value, count, dx.data_ptr() + 4 * xoff, B.DTYPES[qt], q.data_ptr() + qoff,
                          s.data_ptr()) == 0
    torch.cuda.synchronize()
    xs = x[xoff:]
    amax = np.float32(np.abs(xs).max())
    fmax = np.float32(FMAX[qt])
    assert float(s.cpu().numpy()[0]) == np.float32(amax / fmax)
    ref = O.to_fp8((xs * np.float32(fmax / amax)).astype(np.float32), False)
    got = q.cpu().numpy()
    assert np.array_equal(got[qoff:], ref)
    assert not got[:qoff].any()


def test_quantizer_all_zero(handle):
    q, s, _, _ = _quantize(handle, np.zeros(4097, np.float32), "FP8_E4M3FN")
    assert s == 1.0 and not q.any()


@pytest.mark.parametrize("at,bt,ct", [("FP8_E4M3FN", "FP8_E4M3FN", "FP16"), ("FP8_E4M3FN", "FP8_E5M2", "BF16"),
                                      ("FP8_E5M2", "FP8_E4M3FN", "FP32"), ("FP8_E5M2", "FP8_E5M2", "FP16")])
@pytest.mark.parametrize("ta,tb", [("T", "N"), ("N", "N"), ("N", "T"), ("T", "T")])
@pytest.mark.parametrize("shape", [(256, 256, 512), (300, 129, 200), (1024, 768, 640)])
def test_scaled_gemm(handle, at, bt, ct, ta, tb, shape):
    m, n, k = shape
    rng = np.random.default_rng(m + n + k)
    rowsA, colsA = (m, k) if ta == "N" else (k, m)
    rowsB, colsB = (k, n) if tb == "N" else (n, k)
    xa = rng.standard_normal(rowsA * colsA).astype(np.float32)
    xb = (rng.standard_normal(rowsB * colsB) * 5).astype(np.float32)
    qa, sa, dqa, dsa = _quantize(handle, xa, at)
    qb, sb, dqb, dsb = _quantize(handle, xb, bt)
    c0 = rng.standard_normal(m * n).astype(np.float32)
    cq = {"FP32": c0, "FP16": O.to_f16(c0), "BF16": O.to_bf16(c0)}[ct]
    dC = to_dev(cq)
    alpha, beta = 0.75, -0.5
    top = {"N": B.OP_N, "T": B.OP_T}
    st = B.gemm_ex_scaled(handle.value, top[ta], top[tb], m, n, k, alpha, dqa.data_ptr(), B.DTYPES[at], rowsA,
                          dsa.data_ptr(), dqb.data_ptr(), B.DTYPES[bt], rowsB, dsb.data_ptr(), beta, dC.data_ptr(),
                          B.DTYPES[ct], m)
    assert st == 0
    torch.cuda.synchronize()
    out = {"FP32": lambda r: r, "FP16": O.from_f16, "BF16": O.from_bf16}[ct](to_host(dC)).astype(np.float64)
    cv = {"FP32": lambda r: r, "FP16": O.from_f16, "BF16": O.from_bf16}[ct](cq).astype(np.float64)
    va = O.from_fp8(qa, at == "FP8_E4M3FN").astype(np.float64).reshape(colsA, rowsA).T
    vb = O.from_fp8(qb, bt == "FP8_E4M3FN").astype(np.float64).reshape(colsB, rowsB).T
    opA = va if ta == "N" else va.T
    opB = vb if tb == "N" else vb.T
    f = alpha * sa * sb
    gold = f * (opA @ opB) + beta * cv.reshape(n, m).T
    absterm = abs(f) * (np.abs(opA) @ np.abs(opB))
    ulp = {"FP32": 2.0 ** -23, "FP16": 2.0 ** -10, "BF16": 2.0 ** -7}[ct] * np.abs(gold)
    bound = 2.0 ** -11 * absterm + abs(beta) * np.abs(cv.reshape(n, m).T) * 2.0 ** -22 + ulp + 1e-30
    err = np.abs(out.reshape(n, m).T - gold)
    assert (err <= bound).all(), float((err / bound).max())
    # end to end against the unquantised product: each quantised operand is within
    # eps |x| + delta of x (eps = 2^-(mantissa bits + 1), delta = half the subnormal step times
    # the scale), so |out - exact| <= |alpha| (Ea |B| + |A| Eb + Ea Eb) + the bound above
    xa_m = np.abs(xa.astype(np.float64).reshape(colsA, rowsA).T)
    xb_m = np.abs(xb.astype(np.float64).reshape(colsB, rowsB).T)
    XA = xa_m if ta == "N" else xa_m.T
    XB = xb_m if tb == "N" else xb_m.T
    eps = {"FP8_E4M3FN": 2.0 ** -4, "FP8_E5M2": 2.0 ** -3}
    sub = {"FP8_E4M3FN": 2.0 ** -10, "FP8_E5M2": 2.0 ** -17}
    Ea = eps[at] * XA + sub[at] * sa
    Eb = eps[bt] * XB + sub[bt] * sb
    xs = xa.astype(np.float64).reshape(colsA, rowsA).T
    ys = xb.astype(np.float64).reshape(colsB, rowsB).T
    exact = alpha * ((xs if ta == "N" else xs.T) @ (ys if tb == "N" else ys.T)) + beta * cv.reshape(n, m).T
    qbound = abs(alpha) * (Ea @ XB + XA @ Eb + Ea @ Eb) + bound
    err2 = np.abs(out.reshape(n, m).T - exact)
    assert (err2 <= qbound).all(), float((err2 / qbound).max())

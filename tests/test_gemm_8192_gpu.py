"""GPU parity of the large-GEMM kernels at the benchmarked shape, M = N = K = 8192.

At this size aclblasGemmEx (BF16 x BF16 -> FP32) runs the direct-to-LDS 16-bit
kernels (32-k stages when A is m-contiguous, 64-k stages when A is
k-contiguous) and aclblasGemmExScaled (E4M3 x E4M3 with per-tensor scales) the
FP8 direct-to-LDS kernel.  Inputs are N(0,1) (torch's seeded generator on the
GPU), quantised to the operand type before the call, so host and GPU see the
same values.  Three checks per case:

  1. a sample of 256 rows x 256 columns of C against the oracle's GEMM (FP64
     accumulation, one rounding to the output type) on the sampled rows of
     op(A) and columns of op(B):
       - element by element, the bound any correct FP32-accumulating GEMM
         meets (Higham, Accuracy and Stability of Numerical Algorithms, 2nd ed.,
         eq. 3.5): |out - gold| <= gamma_(k+2) sum_l |a_il b_lj| + ulp(gold);
         for the scaled FP8 GEMM the native instruction's summation error
         2^-11 f sum_l |qa qb| + ulp(gold) (f = alpha sA sB), the bound of the
         smaller-shape tests;
       - the relative Frobenius error of the sample, which for N(0,1) inputs
         is about sqrt(k) u / sqrt(k) = u-sized for a correct kernel and jumps to
         about sqrt(32/k) if one 32-k block were dropped or misplaced;
  2. a full-size Freivalds check: C v against A (B v) for a random v, in FP64,
     with the bound |(C - AB) v| <= E |v| where E is the element-wise bound of 1
     (applied through |A| (|B| |v|)).
Running time: one GEMM each plus FP64 matrix-vector products; the oracle's
sampled product is 256 x 256 x 8192 on the host.
"""
import numpy as np
import pytest
import torch

import aclblas as B
import oracle as O

pytestmark = pytest.mark.gpu

M = 8192
SAMPLE = 256
U32 = 2.0 ** -24


def _gamma(k):
    return (k + 2) * U32 / (1 - (k + 2) * U32)


def _randn(seed, count):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(count, device="cuda", generator=g)


def _op_rows(a_vals, trans, rows):
    """Rows `rows` of op(A) for an 8192 x 8192 column-major A (FP32 values on the GPU), as a host
    column-major (len(rows) x K) FP32 array."""
    av = a_vals.view(M, M)              # av[j, i] = A(i, j)
    if trans == "N":
        sub = av[:, rows]               # (K, r): flat j * r + i = column-major r x K
    else:
        sub = av[rows, :].t()           # op(A)(i, l) = A(l, i) = av[i, l]
    return sub.contiguous().cpu().numpy().reshape(-1)


def _op_cols(b_vals, trans, cols):
    """Columns `cols` of op(B) as a host column-major (K x len(cols)) FP32 array."""
    bv = b_vals.view(M, M)              # bv[j, i] = B(i, j)
    if trans == "N":
        sub = bv[cols, :]               # (c, K): column j of B is bv[j, :]
    else:
        sub = bv[:, cols].t()           # op(B)(l, j) = B(j, l) = bv[l, j]
    return sub.contiguous().cpu().numpy().reshape(-1)


def _op_dense64(x_vals, trans):
    """op(X) as a row-major FP64 (M x M) torch tensor on the GPU (checker only)."""
    xv = x_vals.view(M, M).to(torch.float64)   # xv[j, i] = X(i, j)
    return xv.t() if trans == "N" else xv


def _sample_check(out_sub, gold_sub, abs_sub, k, rel_frob_limit, elem_scale):
    err = np.abs(out_sub - gold_sub)
    bound = elem_scale * abs_sub + 2.0 ** -23 * np.abs(gold_sub) + 1e-30
    worst = float((err / bound).max())
    assert worst <= 1.0, "element-wise accumulation bound exceeded by x%.3g" % worst
    rel = float(np.linalg.norm(out_sub - gold_sub) / np.linalg.norm(gold_sub))
    assert rel <= rel_frob_limit, "relative Frobenius error of the sample %.3e" % rel
    return rel


def _freivalds(C64, opA64, opB64, absA64, absB64, elem_scale, out_ulp, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    v = torch.randn(M, dtype=torch.float64, device="cuda", generator=g)
    lhs = C64 @ v
    rhs = opA64 @ (opB64 @ v)
    bound = elem_scale * (absA64 @ (absB64 @ v.abs())) + out_ulp * (C64.abs() @ v.abs()) + 1e-300
    ratio = float(((lhs - rhs).abs() / bound).max())
    assert ratio <= 1.0, "Freivalds bound exceeded by x%.3g" % ratio
    rel = float((lhs - rhs).norm() / rhs.norm())
    return ratio, rel


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
def test_gemm_ex_bf16_fp32_8192(handle, ta, tb):
    a = _randn(11, M * M).to(torch.bfloat16)
    b = _randn(12, M * M).to(torch.bfloat16)
    c = torch.full((M * M,), float("nan"), device="cuda")  # beta = 0: C must be overwritten
    top = {"N": B.OP_N, "T": B.OP_T}
    st = B.gemm_ex(handle.value, top[ta], top[tb], M, M, M, 1.0, a.data_ptr(), B.ACL_BF16, M, b.data_ptr(),
                   B.ACL_BF16, M, 0.0, c.data_ptr(), B.ACL_FLOAT, M, B.COMPUTE["COMPUTE_32F"])
    assert st == 0
    torch.cuda.synchronize()
    assert bool(torch.isfinite(c).all())
    af, bf = a.float(), b.float()

    rng = np.random.default_rng(7)
    rows = torch.from_numpy(np.sort(rng.choice(M, SAMPLE, replace=False))).cuda()
    cols = torch.from_numpy(np.sort(rng.choice(M, SAMPLE, replace=False))).cuda()
    a_sub = _op_rows(af, ta, rows)
    b_sub = _op_cols(bf, tb, cols)
    gold = np.zeros(SAMPLE * SAMPLE, np.float32)
    O.gemm(0, 0, SAMPLE, SAMPLE, M, 1.0, a_sub, O.DT_F32, SAMPLE, b_sub, O.DT_F32, M, 0.0, gold, O.DT_F32,
           SAMPLE)
    absg = np.zeros(SAMPLE * SAMPLE, np.float32)
    O.gemm(0, 0, SAMPLE, SAMPLE, M, 1.0, np.abs(a_sub), O.DT_F32, SAMPLE, np.abs(b_sub), O.DT_F32, M, 0.0, absg,
           O.DT_F32, SAMPLE)
    cv = c.view(M, M)                                    # cv[j, i] = C(i, j)
    out_sub = cv[cols][:, rows].contiguous().cpu().numpy().reshape(-1).astype(np.float64)
    rel = _sample_check(out_sub, gold.astype(np.float64), absg.astype(np.float64), M, 1e-6, _gamma(M))

    opA, opB = _op_dense64(af, ta), _op_dense64(bf, tb)
    C64 = c.view(M, M).to(torch.float64).t()
    ratio, frel = _freivalds(C64, opA, opB, opA.abs(), opB.abs(), _gamma(M), 2.0 ** -23, 13)
    print("bf16 %s%s 8192^3: sample rel %.2e, Freivalds bound ratio %.3g rel %.2e" % (ta, tb, rel, ratio, frel))
    assert frel <= 1e-6


def _quantize(handle, x, q, s):
    assert B.quantize_fp8(handle.value, x.numel(), x.data_ptr(), B.ACL_FLOAT8_E4M3FN, q.data_ptr(),
                          s.data_ptr()) == 0


def test_gemm_ex_scaled_fp8_e4m3_8192(handle):
    e4 = B.ACL_FLOAT8_E4M3FN
    qa = torch.empty(M * M, dtype=torch.uint8, device="cuda")
    qb = torch.empty(M * M, dtype=torch.uint8, device="cuda")
    sc = torch.empty(2, device="cuda")
    xa = _randn(21, M * M)
    xb = _randn(22, M * M)
    _quantize(handle, xa, qa, sc[0:1])
    _quantize(handle, xb, qb, sc[1:2])
    cb = torch.full((M * M,), float("nan"), dtype=torch.bfloat16, device="cuda")
    st = B.gemm_ex_scaled(handle.value, B.OP_N, B.OP_N, M, M, M, 1.0, qa.data_ptr(), e4, M, sc.data_ptr(),
                          qb.data_ptr(), e4, M, sc.data_ptr() + 4, 0.0, cb.data_ptr(), B.ACL_BF16, M)
    assert st == 0
    torch.cuda.synchronize()
    sa, sb = (float(x) for x in sc.cpu().numpy())
    # the scales are amax / 448 of the unquantised operands
    assert sa == np.float32(float(xa.abs().max()) / np.float32(448.0))
    assert sb == np.float32(float(xb.abs().max()) / np.float32(448.0))
    del xa, xb
    # FP8 values as FP32 on the host for the sampled oracle GEMM (the oracle's own conversion)
    rng = np.random.default_rng(8)
    rows = torch.from_numpy(np.sort(rng.choice(M, SAMPLE, replace=False))).cuda()
    cols = torch.from_numpy(np.sort(rng.choice(M, SAMPLE, replace=False))).cuda()
    qa_rows = qa.view(M, M)[:, rows].contiguous().cpu().numpy().reshape(-1)     # column-major 256 x K codes
    qb_cols = qb.view(M, M)[cols, :].contiguous().cpu().numpy().reshape(-1)     # column-major K x 256 codes
    a_sub = O.from_fp8(qa_rows, True)
    b_sub = O.from_fp8(qb_cols, True)
    f = np.float64(sa) * np.float64(sb)
    gold = np.zeros(SAMPLE * SAMPLE, np.float32)
    O.gemm(0, 0, SAMPLE, SAMPLE, M, 1.0, a_sub, O.DT_F32, SAMPLE, b_sub, O.DT_F32, M, 0.0, gold, O.DT_F32,
           SAMPLE, scale_ab=f)
    absg = np.zeros(SAMPLE * SAMPLE, np.float32)
    O.gemm(0, 0, SAMPLE, SAMPLE, M, 1.0, np.abs(a_sub), O.DT_F32, SAMPLE, np.abs(b_sub), O.DT_F32, M, 0.0, absg,
           O.DT_F32, SAMPLE, scale_ab=f)
    raw = cb.view(M, M)[cols][:, rows].contiguous().view(torch.int16).cpu().numpy().reshape(-1).view(np.uint16)
    out_sub = O.from_bf16(raw).astype(np.float64)
    g64 = gold.astype(np.float64)
    err = np.abs(out_sub - g64)
    # native FP8 instruction summation error plus the rounding of the output to BF16 (half an ulp, 2^-8
    # relative, of the exact value; 2^-7 of the computed one covers both)
    bound = 2.0 ** -11 * absg.astype(np.float64) + 2.0 ** -7 * np.abs(g64) + 1e-30
    worst = float((err / bound).max())
    assert worst <= 1.0, "scaled FP8 bound exceeded by x%.3g" % worst
    rel = float(np.linalg.norm(out_sub - g64) / np.linalg.norm(g64))
    assert rel <= 4e-3, rel    # BF16 output rounding: about 2^-9 relative RMS

    # full-size Freivalds in FP64 on the dequantised FP8 values
    lut = torch.from_numpy(O.from_fp8(np.arange(256, dtype=np.uint8), True).astype(np.float64)).cuda()
    A64 = lut[qa.long()].view(M, M).t() * sa           # A(i, j) at [i, j]
    B64 = lut[qb.long()].view(M, M).t() * sb
    C64 = cb.view(M, M).to(torch.float64).t()
    ratio, frel = _freivalds(C64, A64, B64, A64.abs(), B64.abs(), 2.0 ** -11, 2.0 ** -7, 23)
    print("fp8 e4m3 scaled NN 8192^3: sample rel %.2e, Freivalds bound ratio %.3g rel %.2e" % (rel, ratio, frel))
    assert frel <= 4e-3

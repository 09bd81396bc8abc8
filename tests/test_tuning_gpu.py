"""Every launch choice of a tuned call gives the same result, bit for bit.

The tile shape, L2 grouping, persistent grid size and workgroups per CU only
change which workgroup computes which output; each output sums its products in
the same order under every choice.  Each case forces every legal choice through
aclblasSetTuningOverride and compares the outputs with those of the built-in
choice (ragged shapes, so edge tiles of every tile size are exercised).
"""
import numpy as np
import pytest
import torch

import aclblas as B
import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture()
def h():
    handle = B.Handle()
    assert handle.set_stream(torch.cuda.current_stream().cuda_stream) == 0
    yield handle.value
    for op in range(4):
        B.set_tuning_override(handle.value, op, -1)
    torch.cuda.synchronize()
    handle.destroy()


def _all_choices_equal(h, op, choices, run, outputs):
    assert B.set_tuning_override(h, op, -1) == 0
    run()
    torch.cuda.synchronize()
    ref = [o.clone() for o in outputs()]
    for a, b in choices:
        assert B.set_tuning_override(h, op, a, b) == 0
        for o in outputs():
            o.fill_(float("nan")) if o.is_floating_point() else o.fill_(-7)
        run()
        torch.cuda.synchronize()
        for x, y in zip(ref, outputs()):
            assert torch.equal(x, y), "choice (%d, %d) differs" % (a, b)
    assert B.set_tuning_override(h, op, -1) == 0


GEMM_CHOICES = [(t, g) for t in (64, 128, 256) for g in (1, 3, 8, 64)]


@pytest.mark.parametrize("trans", ["NN", "NT", "TN", "TT"])
@pytest.mark.parametrize("shape", [(300, 257, 129), (33, 65, 40), (64, 64, 1000), (97, 31, 290)])
def test_fp32_deep_prefetch_tiles_agree(h, trans, shape):
    """FP32 tiles 32 (eight k-blocks of loads in flight) and 65 (64 x 64, four in flight) run the same
    MFMA sequence per output as the 64 and 128 tiles: bit-identical results, also when the k loop is
    shorter than the prefetch depth (k = 40) or wraps the register ring several times (k = 1000)."""
    m, n, k = shape
    dev = torch.device("cuda")
    gen = torch.Generator(device="cpu").manual_seed(m * 7 + n * 3 + k)
    ta, tb = trans[0] == "T", trans[1] == "T"
    a = (torch.rand(m * k, generator=gen) * 2 - 1).to(dev)
    b = (torch.rand(k * n, generator=gen) * 2 - 1).to(dev)
    c0 = torch.randn(m * n, generator=gen).to(dev)
    c = c0.clone()
    lda, ldb = (k if ta else m), (n if tb else k)

    def run():
        c.copy_(c0)
        assert B.sgemm(h, B.OP_T if ta else B.OP_N, B.OP_T if tb else B.OP_N, m, n, k, 1.25, a.data_ptr(), lda,
                       b.data_ptr(), ldb, -0.5, c.data_ptr(), m) == 0
    _all_choices_equal(h, B.TUNE_GEMM, [(64, 8), (128, 8), (32, 1), (32, 8), (65, 8)], run, lambda: [c])
    ref = (c0.cpu().double().reshape(n, m).T * -0.5 +
           1.25 * (a.cpu().double().reshape(m, k) if ta else a.cpu().double().reshape(k, m).T) @
           (b.cpu().double().reshape(k, n) if tb else b.cpu().double().reshape(n, k).T))
    got = c.cpu().double().reshape(n, m).T
    assert torch.allclose(got, ref, rtol=0, atol=1e-4 * (k ** 0.5))


@pytest.mark.parametrize("kind", ["f32", "bf16", "f8", "f8s"])
@pytest.mark.parametrize("shape", [(300, 257, 129), (520, 130, 700)])
def test_gemm_choices_agree(h, kind, shape):
    m, n, k = shape
    dev = torch.device("cuda")
    gen = torch.Generator(device="cpu").manual_seed(m + n + k)
    if kind == "f32":
        a = (torch.rand(m * k, generator=gen) * 2 - 1).to(dev)
        b = (torch.rand(k * n, generator=gen) * 2 - 1).to(dev)
        c0 = torch.randn(m * n, generator=gen).to(dev)
        c = c0.clone()

        def run():
            c.copy_(c0)
            assert B.sgemm(h, B.OP_T, B.OP_N, m, n, k, 1.5, a.data_ptr(), k, b.data_ptr(), k, 0.5, c.data_ptr(),
                           m) == 0
        choices = [x for x in GEMM_CHOICES if x[0] != 256] + [(t, g) for t in (32, 65) for g in (1, 8)]
    elif kind == "bf16":
        a = torch.randn(m * k, generator=gen).to(dev).to(torch.bfloat16)
        b = torch.randn(k * n, generator=gen).to(dev).to(torch.bfloat16)
        c0 = torch.randn(m * n, generator=gen).to(dev)
        c = c0.clone()

        def run():
            c.copy_(c0)
            assert B.gemm_ex(h, B.OP_N, B.OP_T, m, n, k, 1.0, a.data_ptr(), B.ACL_BF16, m, b.data_ptr(), B.ACL_BF16,
                             n, 0.25, c.data_ptr(), B.ACL_FLOAT, m, B.COMPUTE["COMPUTE_32F"]) == 0
        choices = GEMM_CHOICES
    else:
        rng = np.random.default_rng(m * k)
        # finite codes of moderate magnitude: the FP16 results stay finite
        a = torch.from_numpy(rng.integers(0, 0x50, m * k, dtype=np.uint8)).to(dev)
        b = torch.from_numpy(rng.integers(0, 0x40, k * n, dtype=np.uint8) |
                             (rng.integers(0, 2, k * n, dtype=np.uint8) << 7)).to(dev)
        c = torch.empty(m * n, dtype=torch.float16, device=dev)
        scale = torch.tensor([0.25, 2.0], device=dev)
        e4, e5 = B.ACL_FLOAT8_E4M3FN, B.ACL_FLOAT8_E5M2
        if kind == "f8":
            def run():
                assert B.gemm_ex(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), e4, m, b.data_ptr(), e5, k, 0.0,
                                 c.data_ptr(), B.ACL_FLOAT16, m, B.COMPUTE["COMPUTE_32F"]) == 0
        else:
            def run():
                assert B.gemm_ex_scaled(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), e4, m, scale.data_ptr(),
                                        b.data_ptr(), e4, k, scale.data_ptr() + 4, 0.0, c.data_ptr(),
                                        B.ACL_FLOAT16, m) == 0
        choices = GEMM_CHOICES
    _all_choices_equal(h, B.TUNE_GEMM, choices, run, lambda: [c])


def test_grouped_choices_agree(h):
    dev = torch.device("cuda")
    rng = np.random.default_rng(4)
    mnk = rng.integers(1, 40, size=(12, 3)) * 16 + rng.integers(0, 16, size=(12, 3))
    gen = torch.Generator(device="cpu").manual_seed(4)
    As = [torch.randn(int(m * k), generator=gen).to(dev) for m, n, k in mnk]
    Bs = [torch.randn(int(k * n), generator=gen).to(dev) for m, n, k in mnk]
    C0 = [torch.randn(int(m * n), generator=gen).to(dev) for m, n, k in mnk]
    Cs = [c.clone() for c in C0]
    ms, ns, ks = ([int(x) for x in mnk[:, i]] for i in range(3))
    cnt = len(ms)

    def run():
        for c, c0 in zip(Cs, C0):
            c.copy_(c0)
        assert B.sgemm_grouped_batched(h, [B.OP_N] * cnt, [B.OP_T] * cnt, ms, ns, ks, [1.0] * cnt,
                                       [x.data_ptr() for x in As], ms, [x.data_ptr() for x in Bs], ns, [0.5] * cnt,
                                       [x.data_ptr() for x in Cs], ms, [1] * cnt) == 0
    _all_choices_equal(h, B.TUNE_GROUPED, [(t, w) for t in (64, 128) for w in (1, 2, 3, 4)], run, lambda: Cs)


def test_mixed_choices_agree(h):
    dev = torch.device("cuda")
    systems = W.conditioned_systems(300, seed=9)
    offA, offB = [], []
    oa = ob = 0
    for s in systems:
        offA.append(oa)
        offB.append(ob)
        oa += s["A"].size
        ob += s["B"].size
    dA = torch.from_numpy(np.concatenate([s["A"] for s in systems])).to(dev)
    dB = torch.from_numpy(np.concatenate([s["B"] for s in systems])).to(dev)
    dX = torch.empty_like(dB)
    iters = torch.empty(len(systems), dtype=torch.int32, device=dev)
    berr = torch.empty(len(systems), dtype=torch.float32, device=dev)
    ns = [s["n"] for s in systems]
    rs = [s["nrhs"] for s in systems]
    pa = [dA.data_ptr() + 4 * o for o in offA]
    pb = [dB.data_ptr() + 4 * o for o in offB]
    px = [dX.data_ptr() + 4 * o for o in offB]

    def run():
        assert B.sgesv_batched_mixed(h, ns, rs, pa, ns, pb, ns, px, ns, 1e-9, 6, iters.data_ptr(), berr.data_ptr(),
                                     None) == 0
    _all_choices_equal(h, B.TUNE_MIXED, [(f, s) for f in (1, 2, 4) for s in (1, 3, 4)], run,
                       lambda: [dX, iters, berr])


def _round8(x):
    return (x + 7) // 8 * 8


@pytest.mark.parametrize("kinds", [("BF16", "FP32"), ("FP16", "FP16"), ("BF16", "BF16")])
@pytest.mark.parametrize("trans", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape", [(512, 384, 320), (264, 520, 129), (40, 33, 64), (300, 257, 32)])
def test_direct_to_lds_kernel_matches_register_staged(h, kinds, trans, shape):
    """256 tiles of 16-bit operands with 16-byte aligned leading dimensions run the
    direct-to-LDS kernel; 128 tiles run the register-staged kernel.  Same products,
    same order: the outputs must agree bit for bit, and with an FP32 reference."""
    m, n, k = shape
    at, ct = kinds
    ta, tb = trans
    dev = torch.device("cuda")
    tdt = {"BF16": torch.bfloat16, "FP16": torch.float16, "FP32": torch.float32}
    rowsA, colsA = (k, m) if ta else (m, k)
    rowsB, colsB = (n, k) if tb else (k, n)
    lda, ldb, ldc = _round8(rowsA), _round8(rowsB), m + 3
    gen = torch.Generator(device="cpu").manual_seed(m * 7 + n * 3 + k + 31 * ta + 17 * tb)
    A = torch.randn(lda * colsA, generator=gen).to(tdt[at]).to(dev)
    Bm = torch.randn(ldb * colsB, generator=gen).to(tdt[at]).to(dev)
    C0 = torch.randn(ldc * n, generator=gen).to(tdt[ct]).to(dev)
    outs = []
    for tile in (256, 128):
        assert B.set_tuning_override(h, B.TUNE_GEMM, tile, 8) == 0
        C = C0.clone()
        assert B.gemm_ex(h, B.OP_T if ta else B.OP_N, B.OP_T if tb else B.OP_N, m, n, k, 0.75, A.data_ptr(),
                         B.DTYPES[at], lda, Bm.data_ptr(), B.DTYPES[at], ldb, -0.5, C.data_ptr(), B.DTYPES[ct], ldc,
                         B.COMPUTE["COMPUTE_32F"]) == 0
        torch.cuda.synchronize()
        outs.append(C)
    assert torch.equal(outs[0], outs[1]), "direct-to-LDS and register-staged kernels differ"
    # FP32 reference (column-major views)
    Af = A.float().cpu().reshape(colsA, lda)[:, :rowsA].t()      # rowsA x colsA
    Bf = Bm.float().cpu().reshape(colsB, ldb)[:, :rowsB].t()
    opA = Af.t() if ta else Af
    opB = Bf.t() if tb else Bf
    ref = 0.75 * (opA.double() @ opB.double()) - 0.5 * C0.float().cpu().reshape(n, ldc)[:, :m].t().double()
    got = outs[0].float().cpu().reshape(n, ldc)[:, :m].t().double()
    tol = {"FP32": 1e-5, "FP16": 2e-3, "BF16": 1e-2}[ct] * max(1.0, ref.abs().max().item())
    assert (got - ref).abs().max().item() <= tol
    # padding rows of C untouched
    pad_before = C0.cpu().reshape(n, ldc)[:, m:]
    pad_after = outs[0].cpu().reshape(n, ldc)[:, m:]
    assert torch.equal(pad_before, pad_after)


@pytest.mark.parametrize("n", [9, 16, 24, 32, 40, 64])
def test_lu_rows_per_lane_agree(h, n):
    """One or two matrix rows per lane: identical factors, pivots and info (ragged batch, ties)."""
    dev = torch.device("cuda")
    batch = 37
    rng = np.random.default_rng(n)
    A0 = rng.integers(-2, 3, size=(batch, n, n)).astype(np.float32)    # many exact pivot ties
    A0[5] = rng.standard_normal((n, n)).astype(np.float32)
    A0[7][:, 3] = 0.0                                                   # zero pivot
    dA0 = torch.from_numpy(A0).to(dev)
    LU = dA0.clone()
    piv = torch.empty((batch, n), dtype=torch.int32, device=dev)
    info = torch.empty(batch, dtype=torch.int32, device=dev)
    ptrs = B.ptr_array(LU)

    def run():
        LU.copy_(dA0)
        piv.fill_(-7)
        assert B.sgetrf_batched(h, n, ptrs.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch) == 0
    _all_choices_equal(h, B.TUNE_LU, [(1, 1), (2, 1)], run, lambda: [LU, piv, info])


@pytest.mark.parametrize("kinds", [("FP8_E4M3FN", "FP8_E4M3FN", "FP16"), ("FP8_E5M2", "FP8_E4M3FN", "FP32"),
                                   ("FP8_E4M3FN", "FP8_E5M2", "BF16")])
@pytest.mark.parametrize("trans", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("shape", [(512, 384, 640), (264, 520, 300), (40, 33, 100)])
def test_direct_to_lds_fp8_kernel_matches_register_staged(h, kinds, trans, shape):
    """aclblasGemmExScaled with 256 tiles and 16-byte aligned operands runs the FP8 direct-to-LDS kernel,
    with 128 tiles the register-staged kernel: the same MFMA sums, bit-identical outputs."""
    m, n, k = shape
    at, bt, ct = kinds
    ta, tb = trans
    dev = torch.device("cuda")
    rowsA, colsA = (k, m) if ta else (m, k)
    rowsB, colsB = (n, k) if tb else (k, n)
    r16 = lambda x: (x + 15) // 16 * 16
    lda, ldb, ldc = r16(rowsA), r16(rowsB), m + 5
    rng = np.random.default_rng(m + 3 * n + 7 * k + 11 * ta + 13 * tb)
    codes = lambda cnt: (rng.integers(0, 0x40, cnt, dtype=np.uint8) | (rng.integers(0, 2, cnt, dtype=np.uint8) << 7))
    A = torch.from_numpy(codes(lda * colsA)).to(dev)
    Bm = torch.from_numpy(codes(ldb * colsB)).to(dev)
    tdt = {"FP16": torch.float16, "BF16": torch.bfloat16, "FP32": torch.float32}[ct]
    C0 = torch.randn(ldc * n, generator=torch.Generator(device="cpu").manual_seed(k)).to(tdt).to(dev)
    scale = torch.tensor([0.5, 0.25], device=dev)
    outs = []
    for tile in (256, 128):
        assert B.set_tuning_override(h, B.TUNE_GEMM, tile, 8) == 0
        C = C0.clone()
        assert B.gemm_ex_scaled(h, B.OP_T if ta else B.OP_N, B.OP_T if tb else B.OP_N, m, n, k, 1.25, A.data_ptr(),
                                B.DTYPES[at], lda, scale.data_ptr(), Bm.data_ptr(), B.DTYPES[bt], ldb,
                                scale.data_ptr() + 4, -0.5, C.data_ptr(), B.DTYPES[ct], ldc) == 0
        torch.cuda.synchronize()
        outs.append(C)
    assert torch.equal(outs[0], outs[1]), "direct-to-LDS and register-staged FP8 kernels differ"
    pad_before = C0.cpu().reshape(n, ldc)[:, m:]
    assert torch.equal(pad_before, outs[0].cpu().reshape(n, ldc)[:, m:])

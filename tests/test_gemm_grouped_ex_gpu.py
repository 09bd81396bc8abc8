"""GPU parity of aclblasGemmGroupedBatchedEx (grouped FP16 / BF16 / FP8 GEMM).

Case list: the reference's CSV table (tests/golden/cases/gemm_grouped_batched_ex.csv, from
test/gemm_grouped_batched_ex/arch35/gemm_grouped_batched_ex_test.csv), each problem checked against the
oracle's FP64 GEMM with the table's MERE / MARE thresholds (and, where that criterion is too coarse, the
FP32-accumulation bound, as for the other GEMM entries).  Beside it: the epilogue's FP16 rounding (ties
away from zero, reference gemm_grouped_batched_ex_epilogue_kernel.cpp:146-148) on exact ties, and BF16
groups bit-identical to aclblasGemmBatchedEx on the same problems.
"""
import ctypes

import numpy as np
import pytest

import aclblas as B
import oracle as O
from common import load_cases, status_code, to_dev, to_host
from test_gemm_gpu import _as_f32, check_gemm, parse_fill, quantize

pytestmark = pytest.mark.gpu

TOP = {"N": B.OP_N, "T": B.OP_T, "C": B.OP_C}


def _arr(row, key, count, conv):
    text = row[key].strip()
    parts = text.split(";") if text else []
    if len(parts) == 1 and count > 1:
        parts = parts * count
    return [conv(p) for p in parts]


def _scalars(values, compute):
    """alpha / beta host array: float32, or FP16 bit patterns under COMPUTE_16F; and the values used."""
    if compute == "COMPUTE_16F":
        h = np.array(values, np.float16)
        arr = (ctypes.c_uint16 * max(1, len(values)))(*h.view(np.uint16).tolist())
        return arr, [float(x) for x in h.astype(np.float32)]
    return [float(v) for v in values], [float(np.float32(v)) for v in values]


def _ptr_table(tensors):
    import torch
    return torch.tensor([t.data_ptr() for t in tensors] or [0], dtype=torch.int64, device="cuda")


CASES = [r for r in load_cases("gemm_grouped_batched_ex.csv")]


@pytest.mark.parametrize("row", CASES, ids=[r["case_name"] for r in CASES])
def test_grouped_ex_reference_cases(handle, row):
    expect = status_code(row["expect_result"])
    if expect != 0:
        pytest.skip("argument-check row: covered by tests/test_abi.py on the CPU")
    gc = int(row["group_count"])
    at, bt, ct, comp = row["Atype"], row["Btype"], row["Ctype"], row["computeType"]
    ta = _arr(row, "transa_array", gc, str)
    tb = _arr(row, "transb_array", gc, str)
    ms, ns, ks = (_arr(row, key, gc, int) for key in ("m_array", "n_array", "k_array"))
    sizes = _arr(row, "group_size", gc, int)
    alpha_arr, alphas = _scalars(_arr(row, "alpha_array", gc, float), comp)
    beta_arr, betas = _scalars(_arr(row, "beta_array", gc, float), comp)
    lda_l, ldb_l, ldc_l = [], [], []
    for g in range(gc):
        rowsA = ms[g] if ta[g] == "N" else ks[g]
        rowsB = ks[g] if tb[g] == "N" else ns[g]
        lda_l.append(_arr(row, "lda_array", gc, int)[g] if row["lda_array"] else max(1, rowsA))
        ldb_l.append(_arr(row, "ldb_array", gc, int)[g] if row["ldb_array"] else max(1, rowsB))
        ldc_l.append(_arr(row, "ldc_array", gc, int)[g] if row["ldc_array"] else max(1, ms[g]))
    print("grouped-ex case", row["case_name"], gc, at, bt, ct, comp, flush=True)
    As, Bs, Cs, checks = [], [], [], []
    seed = int(row["random_seed"] or 0)
    for g in range(gc):
        m, n, k, lda, ldb, ldc = ms[g], ns[g], ks[g], lda_l[g], ldb_l[g], ldc_l[g]
        colsA = k if ta[g] == "N" else m
        colsB = n if tb[g] == "N" else k
        for p in range(sizes[g]):
            s = seed + 3 * len(As)
            aq, av = quantize(parse_fill(row["a_fill"], max(1, lda * colsA), s), at)
            bq, bv = quantize(parse_fill(row["b_fill"], max(1, ldb * colsB), s + 1), bt)
            cq, cv = quantize(parse_fill(row["c_fill"], max(1, ldc * n), s + 2), ct)
            gold = cq.copy()
            if m > 0 and n > 0:
                dt = O.DT_F16 if ct == "FP16" else O.DT_BF16
                O.gemm(ta[g] != "N", tb[g] != "N", m, n, k, alphas[g], av, O.DT_F32, lda, bv, O.DT_F32, ldb,
                       betas[g], gold, dt, ldc)
            ref = (av, bv, cv, m, n, k, lda, ldb, ldc, ta[g], tb[g], alphas[g], betas[g])
            As.append(to_dev(aq)); Bs.append(to_dev(bq)); Cs.append(to_dev(cq))
            checks.append((g, cq, gold, ref))
    pa, pb, pc = _ptr_table(As), _ptr_table(Bs), _ptr_table(Cs)
    st = B.gemm_grouped_batched_ex(handle.value, [TOP[x] for x in ta], [TOP[x] for x in tb], ms, ns, ks, alpha_arr,
                                   pa.data_ptr(), B.DTYPES[at], lda_l, pb.data_ptr(), B.DTYPES[bt], ldb_l,
                                   beta_arr, pc.data_ptr(), B.DTYPES[ct], ldc_l, sizes, B.COMPUTE[comp],
                                   group_count=gc)
    assert st == 0
    thr = float(row["mere_threshold"] or 0.0)
    mult = float(row["mare_multiplier"] or 10.0)
    for i, (g, cq, gold, ref) in enumerate(checks):
        m, n, ldc = ref[3], ref[4], ref[8]
        out_raw = to_host(Cs[i])
        if m == 0 or n == 0:
            assert np.array_equal(out_raw, cq), "group %d problem %d: C changed" % (g, i)
            continue
        out = _as_f32(out_raw, ct).reshape(n, ldc)[:, :m]
        gf = _as_f32(gold, ct).reshape(n, ldc)[:, :m]
        # rows below m of each column are untouched
        assert np.array_equal(out_raw.reshape(n, ldc)[:, m:], cq.reshape(n, ldc)[:, m:])
        try:
            check_gemm(out, gf, ct, ref, thr if thr > 0 else 2.0 ** -10, mult,
                       case="gemm_grouped_batched_ex:%s#%d" % (row["case_name"], i))
        except AssertionError as e:
            raise AssertionError("group %d problem %d: %s" % (g, i, e))


def test_fp16_epilogue_rounds_ties_away_from_zero(handle):
    # 1 x 1 x 1 products 1 * 1 scaled by alpha: 1 + 2^-11 lies halfway between the
    # FP16 neighbours 1 and 1 + 2^-10 (nearest-even would give 1)
    import torch
    alphas = [1.0 + 2.0 ** -11, -(1.0 + 2.0 ** -11), 1.0 + 2.0 ** -12, 1.0 + 3 * 2.0 ** -11, 2.0 ** -25]
    expect = [1.0 + 2.0 ** -10, -(1.0 + 2.0 ** -10), 1.0, 1.0 + 2.0 ** -9, 2.0 ** -24]
    g = len(alphas)
    one = torch.tensor([np.float16(1.0).view(np.uint16)], dtype=torch.int16, device="cuda")
    As = [one.clone() for _ in range(g)]
    Bs = [one.clone() for _ in range(g)]
    Cs = [torch.zeros(1, dtype=torch.int16, device="cuda") for _ in range(g)]
    pa, pb, pc = _ptr_table(As), _ptr_table(Bs), _ptr_table(Cs)
    st = B.gemm_grouped_batched_ex(handle.value, [B.OP_N] * g, [B.OP_N] * g, [1] * g, [1] * g, [1] * g, alphas,
                                   pa.data_ptr(), B.ACL_FLOAT16, [1] * g, pb.data_ptr(), B.ACL_FLOAT16, [1] * g,
                                   [0.0] * g, pc.data_ptr(), B.ACL_FLOAT16, [1] * g, [1] * g,
                                   B.COMPUTE["COMPUTE_32F"])
    assert st == 0
    got = [float(to_host(c).view(np.float16)[0]) for c in Cs]
    assert got == expect, got


def test_bf16_groups_match_gemm_batched_ex(handle):
    import torch
    rng = np.random.default_rng(3)
    shapes = [(40, 24, 64, "N", "T", 1.0, 0.0), (128, 96, 32, "T", "N", -0.5, 0.25), (17, 33, 200, "N", "N", 2.0, 1.0)]
    sizes = [2, 3, 1]
    outs_g, outs_b = [], []
    As, Bs, Cg, Cb = [], [], [], []
    for (m, n, k, ta, tb, alpha, beta), size in zip(shapes, sizes):
        for _ in range(size):
            a = O.to_bf16(rng.standard_normal(m * k).astype(np.float32))
            b = O.to_bf16(rng.standard_normal(k * n).astype(np.float32))
            c = O.to_bf16(rng.standard_normal(m * n).astype(np.float32))
            As.append(to_dev(a)); Bs.append(to_dev(b)); Cg.append(to_dev(c)); Cb.append(to_dev(c))
    ta = [TOP[s[3]] for s in shapes]
    tb = [TOP[s[4]] for s in shapes]
    ms, ns, ks = [s[0] for s in shapes], [s[1] for s in shapes], [s[2] for s in shapes]
    lda = [m if t == "N" else k for (m, n, k, t, _, _, _) in shapes]
    ldb = [k if t == "N" else n for (m, n, k, _, t, _, _) in shapes]
    pa, pb, pcg, pcb = _ptr_table(As), _ptr_table(Bs), _ptr_table(Cg), _ptr_table(Cb)
    st = B.gemm_grouped_batched_ex(handle.value, ta, tb, ms, ns, ks, [s[5] for s in shapes], pa.data_ptr(),
                                   B.ACL_BF16, lda, pb.data_ptr(), B.ACL_BF16, ldb, [s[6] for s in shapes],
                                   pcg.data_ptr(), B.ACL_BF16, ms, sizes, B.COMPUTE["COMPUTE_32F"])
    assert st == 0
    start = 0
    for gi, ((m, n, k, _, _, alpha, beta), size) in enumerate(zip(shapes, sizes)):
        st = B.gemm_batched_ex(handle.value, ta[gi], tb[gi], m, n, k, alpha, pa.data_ptr() + 8 * start,
                               B.ACL_BF16, lda[gi], pb.data_ptr() + 8 * start, B.ACL_BF16, ldb[gi], beta,
                               pcb.data_ptr() + 8 * start, B.ACL_BF16, m, size, B.COMPUTE["COMPUTE_32F"])
        assert st == 0
        start += size
    for x, y in zip(Cg, Cb):
        assert torch.equal(x, y)

"""BLASLt on the CPU: descriptor objects, algorithm queries, aclblasLtMatmul's argument rows, and the
oracle's BLASLt golden pinned.

Object rules follow the reference (blasLt/api/aclblaslt_layout.cpp, aclblaslt_matmul_desc.cpp,
aclblaslt_matmul_preference.cpp, aclblaslt_version.cpp); the argument rows are every row of the reference's
table (test/blasLtMatmul/arch35/blasLtMatmul_test.csv) whose status is decided before any device work: error
rows and empty problems.  On a host without a GPU the handle has no device, so nothing is ever launched here.

The golden (oracle.lt_matmul, the reference's test/blasLtMatmul/blasLtMatmul_golden.h restated) is pinned by:
its E4M3 decoding against torch's float8_e4m3fn (every byte below the exponent-15 range the golden maps to
zero), its E2M1 decoding against the OCP value table, E8M0 against 2^(s-127); the FP32 path against numpy in
FP64; and, for the two scale layouts and the operand transposes, invariances a wrong index breaks: the same
logical product from transposed storage with the K-major-pairs scale layout and from K-contiguous storage
with the row layout, scaling by one scale step doubling the product, and a single block's scale changing
exactly the entries it covers.
"""
import ctypes

import numpy as np
import pytest

import aclblas as B
import oracle as O
from common import load_cases, status_code

LT_TYPES = {"FP32": B.ACL_FLOAT, "BF16": B.ACL_BF16, "MXFP8_E4M3FN": B.ACL_FLOAT8_E4M3FN,
            "MXFP4_E2M1": B.ACL_FLOAT4_E2M1}


@pytest.fixture(scope="module")
def lt():
    st, h = B.lt_create()
    assert st == 0
    yield h
    assert B.load().aclblasLtDestroy(h) == 0


def test_version_and_properties():
    L = B.load()
    v = ctypes.c_size_t()
    assert L.aclblasLtGetVersion(ctypes.byref(v)) == 0 and v.value == (1 << 24)
    assert L.aclblasLtGetVersion(None) == status_code("INVALID_VALUE")
    x = ctypes.c_int(-1)
    for prop, want in ((0, 1), (1, 0), (2, 0)):
        assert L.aclblasLtGetProperty(prop, ctypes.byref(x)) == 0 and x.value == want
    assert L.aclblasLtGetProperty(3, ctypes.byref(x)) == status_code("INVALID_VALUE")


def test_layout_attributes():
    L = B.load()
    inv = status_code("INVALID_VALUE")
    lay = ctypes.c_void_p()
    assert L.aclblasLtMatrixLayoutCreate(None, 0, 4, 4, 4) == inv
    assert L.aclblasLtMatrixLayoutCreate(ctypes.byref(lay), 0, 4, 4, -1) == inv
    assert L.aclblasLtMatrixLayoutCreate(ctypes.byref(lay), 0, 7, 5, 0) == 0
    try:
        got = ctypes.c_int64()
        n = ctypes.c_size_t()
        assert L.aclblasLtMatrixLayoutGetAttribute(lay, B.LT_LAYOUT_LD, ctypes.byref(got), 8, ctypes.byref(n)) == 0
        assert got.value == 7 and n.value == 8      # ld 0 means rows
        order = ctypes.c_int32()
        assert L.aclblasLtMatrixLayoutGetAttribute(lay, B.LT_LAYOUT_ORDER, ctypes.byref(order), 4, None) == 0
        assert order.value == B.LT_ORDER_COL        # default order
        rows = ctypes.c_uint64(9)
        assert L.aclblasLtMatrixLayoutSetAttribute(lay, B.LT_LAYOUT_ROWS, ctypes.byref(rows), 8) == 0
        assert L.aclblasLtMatrixLayoutSetAttribute(lay, B.LT_LAYOUT_ROWS, ctypes.byref(rows), 4) == inv
        small = ctypes.c_int32()
        assert L.aclblasLtMatrixLayoutGetAttribute(lay, B.LT_LAYOUT_ROWS, ctypes.byref(small), 4,
                                                   ctypes.byref(n)) == inv and n.value == 8
        out = ctypes.c_uint64()
        assert L.aclblasLtMatrixLayoutGetAttribute(lay, B.LT_LAYOUT_ROWS, ctypes.byref(out), 8, None) == 0
        assert out.value == 9
        t = ctypes.c_uint32()
        assert L.aclblasLtMatrixLayoutGetAttribute(lay, B.LT_LAYOUT_TYPE, ctypes.byref(t), 4, None) == 0
        assert t.value == B.ACL_FLOAT
        assert L.aclblasLtMatrixLayoutSetAttribute(lay, 7, ctypes.byref(rows), 8) == inv
        assert L.aclblasLtMatrixLayoutSetAttribute(lay, B.LT_LAYOUT_ROWS, None, 8) == inv
    finally:
        assert L.aclblasLtMatrixLayoutDestroy(lay) == 0
    assert L.aclblasLtMatrixLayoutDestroy(None) == inv


def test_matmul_desc_attributes():
    L = B.load()
    inv, nsup = status_code("INVALID_VALUE"), status_code("NOT_SUPPORTED")
    st, d = B.lt_desc(transA=B.OP_T)
    assert st == 0
    try:
        v = ctypes.c_int32()
        assert L.aclblasLtMatmulDescGetAttribute(d, B.LT_DESC_TRANSA, ctypes.byref(v), 4, None) == 0
        assert v.value == B.OP_T
        assert L.aclblasLtMatmulDescGetAttribute(d, B.LT_DESC_TRANSB, ctypes.byref(v), 4, None) == 0
        assert v.value == B.OP_N
        assert L.aclblasLtMatmulDescSetAttribute(d, B.LT_DESC_TRANSA, ctypes.byref(v), 8) == inv
        e = ctypes.c_uint32()
        assert L.aclblasLtMatmulDescGetAttribute(d, B.LT_DESC_EPILOGUE, ctypes.byref(e), 4, None) == 0
        assert e.value == B.LT_EPILOGUE_DEFAULT
        p = ctypes.c_void_p(0x1000)
        assert L.aclblasLtMatmulDescSetAttribute(d, B.LT_DESC_A_SCALE_POINTER, ctypes.byref(p), 8) == 0
        q = ctypes.c_void_p()
        assert L.aclblasLtMatmulDescGetAttribute(d, B.LT_DESC_A_SCALE_POINTER, ctypes.byref(q), 8, None) == 0
        assert q.value == 0x1000
        mode = ctypes.c_int32(1)
        assert L.aclblasLtMatmulDescSetAttribute(d, B.LT_DESC_A_SCALE_MODE, ctypes.byref(mode), 4) == 0
        assert L.aclblasLtMatmulDescSetAttribute(d, 13, ctypes.byref(mode), 4) == nsup   # pointer mode
        assert L.aclblasLtMatmulDescGetAttribute(d, 13, ctypes.byref(mode), 4, None) == nsup
        assert L.aclblasLtMatmulDescInit(d, 2, 0) == 0                                    # back to defaults
        assert L.aclblasLtMatmulDescGetAttribute(d, B.LT_DESC_TRANSA, ctypes.byref(v), 4, None) == 0
        assert v.value == B.OP_N
    finally:
        assert L.aclblasLtMatmulDescDestroy(d) == 0
    assert L.aclblasLtMatmulDescCreate(None, 2, 0) == inv


def test_preference_and_algorithms(lt):
    L = B.load()
    inv, nsup = status_code("INVALID_VALUE"), status_code("NOT_SUPPORTED")
    pref = ctypes.c_void_p()
    assert L.aclblasLtMatmulPreferenceCreate(ctypes.byref(pref)) == 0
    ws = ctypes.c_uint64(1 << 20)
    assert L.aclblasLtMatmulPreferenceSetAttribute(pref, B.LT_PREF_MAX_WORKSPACE_BYTES, ctypes.byref(ws), 8) == 0
    got = ctypes.c_uint64()
    assert L.aclblasLtMatmulPreferenceGetAttribute(pref, B.LT_PREF_MAX_WORKSPACE_BYTES, ctypes.byref(got), 8,
                                                   None) == 0 and got.value == 1 << 20
    assert L.aclblasLtMatmulPreferenceSetAttribute(pref, 5, ctypes.byref(ws), 8) == inv
    ids = (ctypes.c_int * 4)()
    cnt = ctypes.c_int(-1)
    F, E4, F4, BF = B.ACL_FLOAT, B.ACL_FLOAT8_E4M3FN, B.ACL_FLOAT4_E2M1, B.ACL_BF16
    assert L.aclblasLtMatmulAlgoGetIds(lt, 2, F, F, F, F, F, ids, 4, ctypes.byref(cnt)) == 0 and cnt.value == 1
    assert L.aclblasLtMatmulAlgoGetIds(lt, 2, F, B.ACL_FLOAT16, B.ACL_FLOAT16, F, F, ids, 4,
                                       ctypes.byref(cnt)) == 0 and cnt.value == 0
    assert L.aclblasLtMatmulAlgoGetIds(None, 2, F, F, F, F, F, ids, 4, ctypes.byref(cnt)) == \
        status_code("NOT_INITIALIZED")
    algo = B.LtAlgo()
    assert L.aclblasLtMatmulAlgoInit(lt, 2, F, E4, E4, F, BF, 0, ctypes.byref(algo)) == 0
    assert algo.max_workspace_bytes == 0
    assert L.aclblasLtMatmulAlgoInit(lt, 2, F, E4, F4, F, BF, 0, ctypes.byref(algo)) == nsup
    assert L.aclblasLtMatmulAlgoInit(lt, 2, F, F, F, F, F, 1, ctypes.byref(algo)) == inv
    results = (B.LtHeuristicResult * 3)()
    st, d = B.lt_desc()
    lays = [B.lt_layout(F4, 16, 16, 8)[1], B.lt_layout(F4, 16, 16, 8)[1], B.lt_layout(F, 16, 16, 16)[1],
            B.lt_layout(BF, 16, 16, 16)[1]]
    assert L.aclblasLtMatmulAlgoGetHeuristic(lt, d, *lays, pref, 3, results, ctypes.byref(cnt)) == 0
    assert cnt.value == 1 and results[0].state == 0 and results[0].workspaceSize == 0
    bad = B.lt_layout(B.ACL_FLOAT16, 16, 16, 16)[1]
    assert L.aclblasLtMatmulAlgoGetHeuristic(lt, d, bad, bad, lays[2], lays[3], pref, 3, results,
                                             ctypes.byref(cnt)) == nsup and cnt.value == 0
    for x in lays + [bad]:
        L.aclblasLtMatrixLayoutDestroy(x)
    L.aclblasLtMatmulDescDestroy(d)
    assert L.aclblasLtMatmulPreferenceDestroy(pref) == 0


# ------------------------------------------------------------------ argument rows of the reference's table
def _case_layouts(row):
    at, bt = LT_TYPES[row["dtypeA"]], LT_TYPES[row["dtypeB"]]
    ct, dt = LT_TYPES[row["dtypeC"]], LT_TYPES[row["dtypeD"]]
    M, N, K = int(row["M"]), int(row["N"]), int(row["K"])
    ta, tb = row["transA"], row["transB"]
    ra, ca = (M, K) if ta == "N" else (K, M)
    rb, cb = (K, N) if tb == "N" else (N, K)
    lda, ldb = int(row["lda"]), int(row["ldb"])
    if at == B.ACL_FLOAT4_E2M1:     # the reference's callers give FP4 layouts the bytes per row
        lda, ldb = (lda + 1) // 2, (ldb + 1) // 2
    return [B.lt_layout(at, ra, ca, lda)[1], B.lt_layout(bt, rb, cb, ldb)[1],
            B.lt_layout(ct, M, N, int(row["ldc"]))[1], B.lt_layout(dt, M, N, int(row["ldd"]))[1]]


CPU_ROWS = [r for r in load_cases("blaslt_matmul.csv")
            if status_code(r["expect_result"]) != 0 or int(r["M"]) == 0 or int(r["N"]) == 0]


@pytest.mark.parametrize("row", CPU_ROWS, ids=[r["case_name"] for r in CPU_ROWS])
def test_reference_argument_rows(lt, row):
    L = B.load()
    lays = _case_layouts(row)
    st, d = B.lt_desc(transA=B.OPS[row["transA"]], transB=B.OPS[row["transB"]],
                      scaleA=0x100 if row["scaleA_fill"] != "none" else None,
                      scaleB=0x100 if row["scaleB_fill"] != "none" else None)
    assert st == 0
    fake = 0x1000  # never dereferenced: every row here returns during argument checking
    h = None if row["handle_null"] == "true" else lt
    desc = None if row["computeDesc_null"] == "true" else d
    alpha = None if row["alpha_null"] == "true" else float(row["alpha"])
    A = None if row["A_null"] == "true" else fake
    C = None if row["C_is_null"] == "true" else fake
    try:
        got = L.aclblasLtMatmul(h, desc, B._fp(alpha), A, lays[0], fake, lays[1], B._fp(float(row["beta"])), C,
                                lays[2], fake, lays[3], None, None, 0, None)
        assert got == status_code(row["expect_result"]), (row["case_name"], got)
    finally:
        for x in lays:
            L.aclblasLtMatrixLayoutDestroy(x)
        L.aclblasLtMatmulDescDestroy(d)


def test_matmul_checks_beyond_the_table(lt):
    """Rules of the reference's engine the table does not exercise (matmul_engine.cpp:100-161), and the
    shape checks this build adds before launching."""
    L = B.load()
    inv, nsup = status_code("INVALID_VALUE"), status_code("NOT_SUPPORTED")
    F, E4 = B.ACL_FLOAT, B.ACL_FLOAT8_E4M3FN
    fake = 0x1000

    def run(at=F, dt=F, scales=(fake, fake), beta=0.0, C=fake, shapes=None, epi=None, ws=None, algo=None):
        shp = shapes or ((16, 32), (32, 16), (16, 16), (16, 16))
        lays = [B.lt_layout(at, *shp[0], shp[0][1])[1], B.lt_layout(at, *shp[1], shp[1][1])[1],
                B.lt_layout(F, *shp[2], shp[2][1])[1], B.lt_layout(dt, *shp[3], shp[3][1])[1]]
        st, d = B.lt_desc(scaleA=scales[0], scaleB=scales[1])
        if epi is not None:
            e = ctypes.c_uint32(epi)
            L.aclblasLtMatmulDescSetAttribute(d, B.LT_DESC_EPILOGUE, ctypes.byref(e), 4)
        try:
            return B.lt_matmul(lt, d, 1.0, fake, lays[0], fake, lays[1], beta, C, lays[2], fake, lays[3],
                               algo=algo, workspace=ws)
        finally:
            for x in lays:
                L.aclblasLtMatrixLayoutDestroy(x)
            L.aclblasLtMatmulDescDestroy(d)

    assert run(at=E4, dt=B.ACL_BF16, scales=(None, fake)) == inv          # MX without scales
    assert run(beta=1.0, C=None) == inv                                   # beta with a null C
    assert run(ws=fake + 8) == inv                                        # misaligned workspace
    algo = B.LtAlgo()
    algo.max_workspace_bytes = 64
    assert run(algo=algo) == inv                                          # workspace below the algorithm's
    assert run(dt=B.ACL_BF16) == nsup                                     # FP32 inputs, BF16 D
    assert run(epi=B.LT_EPILOGUE_RELU) == nsup                            # epilogues other than DEFAULT
    assert run(shapes=((16, 32), (31, 16), (16, 16), (16, 16))) == inv    # B rows != K
    # a valid problem reaches the launch, which a host without a GPU cannot make (the pointers here are
    # fake: never launch them where a GPU is visible)
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        assert run() == status_code("EXECUTION_FAILED")


# ------------------------------------------------------------------ the golden, pinned
def test_golden_decodings():
    torch = pytest.importorskip("torch")
    # E4M3: every byte with exponent field < 15 (the golden maps 15 to zero), via an MX product with one
    # non-zero element and unit scales
    codes = np.array([b for b in range(256) if ((b >> 3) & 0xF) < 15], dtype=np.uint8)
    want = torch.from_numpy(codes.copy()).view(torch.float8_e4m3fn).float().numpy()
    got = np.empty(codes.size, np.float32)
    one_b = np.full((32, 1), 0x38, np.uint8)            # B: k x n = 32 x 1 of 1.0 (transB N)
    s = np.full(2 * 1, 127, np.uint8)
    for i, c in enumerate(codes):
        a = np.zeros((1, 32), np.uint8)
        a[0, 0] = c
        got[i] = O.lt_matmul(O.LT_MXFP8, False, 0, 0, 1, 1, 32, 1.0, a, 32, one_b, 1, s, s, 0.0, None, 1, 1)[0, 0]
    np.testing.assert_array_equal(got, want)
    # E2M1: the OCP table, low nibble first
    table = [0.0, 0.5, 1.0, 1.5, 2.0, 3.0, 4.0, 6.0]
    for nib in range(16):
        a = np.zeros((1, 16), np.uint8)
        a[0, 0] = nib                                   # element 0 (low nibble)
        b = np.zeros((32, 1), np.uint8)
        b[0, 0] = 0x2                                   # B element (k=0, n=0) = 1.0 (low nibble of byte 0)
        v = O.lt_matmul(O.LT_MXFP4, False, 0, 0, 1, 1, 32, 1.0, a, 16, b, 1, s, s, 0.0, None, 1, 1)[0, 0]
        expect = (-1.0 if nib & 8 else 1.0) * table[nib & 7]
        assert v == expect, (nib, v)
    # E8M0 scales: 2^(s-127), 0 -> 0 (the golden's decoding)
    a = np.full((1, 32), 0x38, np.uint8)
    for sc, expect in ((127, 32.0), (130, 256.0), (120, 32.0 / 128), (0, 0.0)):
        sa = np.array([sc, 127], np.uint8)
        v = O.lt_matmul(O.LT_MXFP8, False, 0, 0, 1, 1, 32, 1.0, a, 32, one_b, 1, sa, s, 0.0, None, 1, 1)[0, 0]
        assert v == expect, (sc, v)


def test_golden_fp32_against_numpy():
    rng = np.random.default_rng(3)
    for ta in (0, 1):
        for tb in (0, 1):
            M, N, K = 7, 9, 13
            A = rng.standard_normal((K, M) if ta else (M, K)).astype(np.float32)
            Bm = rng.standard_normal((N, K) if tb else (K, N)).astype(np.float32)
            C = rng.standard_normal((M, N)).astype(np.float32)
            D = O.lt_matmul(O.LT_FP32, False, ta, tb, M, N, K, 0.5, A, A.shape[1], Bm, Bm.shape[1], None, None,
                            -2.0, C, N, N)
            opA = A.T if ta else A
            opB = Bm.T if tb else Bm
            want = (0.5 * (opA.astype(np.float64) @ opB.astype(np.float64)) - 2.0 * C).astype(np.float32)
            np.testing.assert_allclose(D, want, rtol=1e-6, atol=1e-6)


def _mx_data(rng, rows, cols, fp4):
    if fp4:
        nib = rng.integers(3, 9, size=(rows, cols)).astype(np.uint8)
        return nib, (nib[:, 0::2] | (nib[:, 1::2] << 4)).astype(np.uint8)
    b = rng.integers(0x38, 0x50, size=(rows, cols)).astype(np.uint8)
    return b, b


def _pack_scales_rowwise(s):       # s: [rows][K/32] -> [rows][2 ceil(K/64)]
    rows, nb = s.shape
    out = np.full((rows, 2 * ((nb + 1) // 2)), 127, np.uint8)
    out[:, :nb] = s
    return out.reshape(-1)


def _pack_scales_kpairs(s):        # s: [rows][K/32] -> [ceil(K/64)][rows][2]
    rows, nb = s.shape
    out = np.full(((nb + 1) // 2, rows, 2), 127, np.uint8)
    for b in range(nb):
        out[b // 2, :, b % 2] = s[:, b]
    return out.reshape(-1)


@pytest.mark.parametrize("fp4", [False, True])
def test_golden_layout_invariances(fp4):
    """One logical product M x K by K x N from every storage: A K-contiguous (transA N, row scales) or
    transposed (transA T, K-major pairs); B transposed (transB N, pairs) or K-contiguous (transB T, rows)."""
    rng = np.random.default_rng(11 + fp4)
    M, N, K = 5, 6, 96
    typ = O.LT_MXFP4 if fp4 else O.LT_MXFP8
    a_el, _ = _mx_data(rng, M, K, fp4)          # logical A[m][k] element codes
    b_el, _ = _mx_data(rng, N, K, fp4)          # logical B^T[n][k]
    sa = rng.integers(125, 131, size=(M, K // 32)).astype(np.uint8)
    sb = rng.integers(125, 131, size=(N, K // 32)).astype(np.uint8)

    def store(el, kcontig):
        """(rows x K) element codes -> stored bytes and ld: K-contiguous or transposed."""
        x = el if kcontig else el.T.copy()
        if fp4:
            if x.shape[1] % 2:
                x = np.concatenate([x, np.zeros((x.shape[0], 1), np.uint8)], axis=1)
            x = (x[:, 0::2] | (x[:, 1::2] << 4)).astype(np.uint8)
        return np.ascontiguousarray(x), x.shape[1]

    results = []
    for ta in (0, 1):
        for tb in (0, 1):
            A, lda = store(a_el, ta == 0)
            Bm, ldb = store(b_el, tb == 1)
            sA = _pack_scales_rowwise(sa) if ta == 0 else _pack_scales_kpairs(sa)
            sB = _pack_scales_kpairs(sb) if tb == 0 else _pack_scales_rowwise(sb)
            results.append(O.lt_matmul(typ, False, ta, tb, M, N, K, 1.0, A, lda, Bm, ldb, sA, sB, 0.0, None, 1,
                                       N))
    for r in results[1:]:
        np.testing.assert_array_equal(r, results[0])
    # one scale step on every A block doubles the product
    A, lda = store(a_el, True)
    Bm, ldb = store(b_el, True)
    base = O.lt_matmul(typ, False, 0, 1, M, N, K, 1.0, A, lda, Bm, ldb, _pack_scales_rowwise(sa),
                       _pack_scales_rowwise(sb), 0.0, None, 1, N)
    up = O.lt_matmul(typ, False, 0, 1, M, N, K, 1.0, A, lda, Bm, ldb, _pack_scales_rowwise(sa + 1),
                     _pack_scales_rowwise(sb), 0.0, None, 1, N)
    np.testing.assert_array_equal(up, 2 * base)
    # zeroing one block's scale changes exactly that row of D (and not by a transposed index)
    z = sa.copy()
    z[2, 1] = 0
    zr = O.lt_matmul(typ, False, 0, 1, M, N, K, 1.0, A, lda, Bm, ldb, _pack_scales_rowwise(z),
                     _pack_scales_rowwise(sb), 0.0, None, 1, N)
    changed = np.argwhere(zr != base)
    assert set(changed[:, 0]) == {2}

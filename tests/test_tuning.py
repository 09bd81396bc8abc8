"""Launch tuning on the CPU: bucket boundaries, override semantics and the measured table.

A GPU-less handle reports no compute units, so the bucket functions use their
256-CU default (the MI355X count).  The table rows are parsed from
ops-blas_amd/csrc/tuning_table.inc and must be what the library returns.
"""
import os
import re

import pytest

import aclblas as B
from common import status_code

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "ops-blas_amd", "csrc", "tuning_table.inc")
CUS = 256


@pytest.fixture(scope="module")
def h():
    handle = B.Handle()
    yield handle.value
    handle.destroy()


def _bucket(h, op, cls, units, k=0):
    st, b = B.tuning_bucket(h, op, cls, units, k)
    assert st == 0
    return b


def test_gemm_bucket_boundaries(h):
    # size edges at cus/4, cus, 4 cus, 16 cus tiles; k edges at 256 and 2048
    edges = [CUS // 4, CUS, 4 * CUS, 16 * CUS]
    for cls in range(4):
        for s, lo in enumerate([0] + edges):
            for kb, k in enumerate([1, 256, 2048]):
                expect = (cls * 5 + s) * 3 + kb
                assert _bucket(h, B.TUNE_GEMM, cls, lo, k) == expect
                if s < 4:
                    assert _bucket(h, B.TUNE_GEMM, cls, edges[s] - 1, k) == expect
            assert _bucket(h, B.TUNE_GEMM, cls, 1, 255) == (cls * 5 + 0) * 3 + 0
            assert _bucket(h, B.TUNE_GEMM, cls, 1, 2047) == (cls * 5 + 0) * 3 + 1


def test_grouped_and_mixed_buckets(h):
    for cls in (0, 1):
        assert [_bucket(h, B.TUNE_GROUPED, cls, u) for u in (0, CUS - 1, CUS, 4 * CUS - 1, 4 * CUS, 10 ** 9)] == \
            [3 * cls + b for b in (0, 0, 1, 1, 2, 2)]
    assert B.tuning_bucket(h, B.TUNE_GROUPED, 2, 10)[0] == status_code("INVALID_VALUE")
    for cls in (0, 1):
        assert [_bucket(h, B.TUNE_MIXED, cls, u) for u in (1, CUS - 1, CUS, 4 * CUS, 16 * CUS - 1, 16 * CUS)] == \
            [4 * cls + b for b in (0, 0, 1, 2, 2, 3)]
    assert B.tuning_bucket(h, B.TUNE_MIXED, 2, 10)[0] == status_code("INVALID_VALUE")


def test_lu_buckets(h):
    # S index x waves of a one-row-per-lane launch (batch * S / 64) per SIMD: < 2, < 4, < 8, more
    simds = 4 * CUS
    for n, si in ((1, 0), (8, 0), (9, 1), (16, 1), (17, 2), (32, 2), (33, 3), (64, 3)):
        S = 8 << si
        for w, waves in enumerate((simds, 3 * simds, 6 * simds, 12 * simds)):
            assert _bucket(h, B.TUNE_LU, n, waves * 64 // S) == si * 4 + w
        assert _bucket(h, B.TUNE_LU, n, 2 * simds * 64 // S) == si * 4 + 1
    assert B.tuning_bucket(h, B.TUNE_LU, 0, 10)[0] == status_code("INVALID_VALUE")
    assert B.tuning_bucket(h, B.TUNE_LU, 65, 10)[0] == status_code("INVALID_VALUE")


def test_bucket_argument_checks(h):
    inv = status_code("INVALID_VALUE")
    assert B.tuning_bucket(None, 0, 0, 1, 1)[0] == status_code("HANDLE_IS_NULLPTR")
    assert B.tuning_bucket(h, 4, 0, 1, 1)[0] == inv
    assert B.tuning_bucket(h, 0, 4, 1, 1)[0] == inv
    assert B.tuning_bucket(h, 0, 0, -1, 1)[0] == inv
    assert B.tuning_bucket(h, 0, 0, 1, -1)[0] == inv


def _table_rows():
    rows = []
    for line in open(TABLE):
        m = re.match(r"\s*\{\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*(-?\d+)\s*\}\s*,", line)
        if m:
            rows.append(tuple(int(x) for x in m.groups()))
    return rows


def test_table_rows_are_legal_and_used(h):
    rows = _table_rows()
    seen = set()
    for op, bucket, a, b in rows:
        assert (op, bucket) not in seen, "duplicate row"
        seen.add((op, bucket))
        assert 0 <= op < 4
        st, ga, gb = B.get_tuning_choice(h, op, bucket)
        assert st == 0
        assert (ga, gb) == (a, b), "row %r not honoured" % ((op, bucket, a, b),)
        if op == B.TUNE_GEMM:
            assert a in (64, 128, 256) or (a in (32, 65) and bucket // 15 == B.CLASS_F32)
            assert 1 <= b <= 64
            assert not (a == 256 and bucket // 15 == B.CLASS_F32)
        elif op == B.TUNE_GROUPED:
            assert a in (64, 128) and 1 <= b <= 4
        elif op == B.TUNE_MIXED:
            assert 1 <= a <= 4 and 1 <= b <= 4
        else:
            assert a == 1 or (a == 2 and bucket >= 4)


def test_override_semantics(h):
    inv = status_code("INVALID_VALUE")
    f32_bucket = _bucket(h, B.TUNE_GEMM, B.CLASS_F32, 10 ** 6, 4096)
    bf16_bucket = _bucket(h, B.TUNE_GEMM, B.CLASS_16, 10 ** 6, 4096)
    before_f32 = B.get_tuning_choice(h, B.TUNE_GEMM, f32_bucket)
    try:
        assert B.set_tuning_override(h, B.TUNE_GEMM, 256, 4) == 0
        assert B.get_tuning_choice(h, B.TUNE_GEMM, bf16_bucket) == (0, 256, 4)
        # FP32 cannot take a 256 tile: its bucket keeps the table / built-in choice
        assert B.get_tuning_choice(h, B.TUNE_GEMM, f32_bucket) == before_f32
        assert B.set_tuning_override(h, B.TUNE_GEMM, 64, 16) == 0
        assert B.get_tuning_choice(h, B.TUNE_GEMM, f32_bucket) == (0, 64, 16)
        assert B.set_tuning_override(h, B.TUNE_GEMM, 96, 8) == inv
        assert B.set_tuning_override(h, B.TUNE_GEMM, 128, 0) == inv
        assert B.set_tuning_override(h, B.TUNE_GROUPED, 256, 1) == inv
        assert B.set_tuning_override(h, B.TUNE_MIXED, 5, 1) == inv
        assert B.set_tuning_override(h, 4, 1, 1) == inv
        assert B.set_tuning_override(None, 0, 64, 8) == status_code("HANDLE_IS_NULLPTR")
        assert B.set_tuning_override(h, B.TUNE_MIXED, 1, 2) == 0
        assert B.get_tuning_choice(h, B.TUNE_MIXED, 3) == (0, 1, 2)
        # LU: two rows per lane only above n = 8
        assert B.set_tuning_override(h, B.TUNE_LU, 3, 1) == inv
        assert B.set_tuning_override(h, B.TUNE_LU, 2, 1) == 0
        assert B.get_tuning_choice(h, B.TUNE_LU, _bucket(h, B.TUNE_LU, 32, 4096))[1] == 2
        assert B.get_tuning_choice(h, B.TUNE_LU, _bucket(h, B.TUNE_LU, 8, 4096))[1] != 2
    finally:
        for op in range(4):
            assert B.set_tuning_override(h, op, -1) == 0
    assert B.get_tuning_choice(h, B.TUNE_GEMM, f32_bucket) == before_f32
    assert B.get_tuning_choice(h, 0, 60)[0] == inv
    assert B.get_tuning_choice(h, 2, -1)[0] == inv


def test_built_in_choices_without_table_rows(h):
    rows = {(op, b) for op, b, _, _ in _table_rows()}
    # buckets with no row fall back to a zero choice in the query (the calls use
    # their own built-in shape); buckets with a row report it
    for op, count in ((0, 60), (1, 3), (2, 4), (3, 16)):
        for bucket in range(count):
            st, a, b = B.get_tuning_choice(h, op, bucket)
            assert st == 0
            if (op, bucket) not in rows:
                assert (a, b) == (0, 0)

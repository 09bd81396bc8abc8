"""aclblasQuantizeFp8NoAlloc (the quantiser whose per-block maxima live in a file-scope device array instead of
the handle's workspace) against aclblasQuantizeFp8 through the C ABI: two calls in a row, the second on a
smaller input, codes and scale bit-identical to the workspace entry's, for both FP8 formats and for aligned
and unaligned operands.  Also: nothing is allocated (the handle's workspace size is unchanged and the call
works with a borrowed workspace too small for the workspace entry).
"""
import ctypes

import numpy as np
import pytest
import torch

import aclblas as B

pytestmark = pytest.mark.gpu


def _no_alloc():
    return B.quantize_fp8_no_alloc


def _run(fn, h, x, off, qt):
    count = x.numel() - off
    q = torch.full((count + 3,), 0xA5, dtype=torch.uint8, device="cuda")
    s = torch.full((1,), -1.0, device="cuda")
    st = fn(h, count, x.data_ptr() + 4 * off, qt, q.data_ptr() + 1, s.data_ptr())
    assert st == 0, st
    return q, s


@pytest.mark.parametrize("qt", ["FP8_E4M3FN", "FP8_E5M2"])
@pytest.mark.parametrize("off", [0, 1])
def test_no_alloc_quantizer_matches_workspace_entry(handle, qt, off):
    noalloc = _no_alloc()
    code = B.DTYPES[qt]
    g = torch.Generator(device="cuda").manual_seed(17 + off)
    for count, scale in ((3_000_001, 5.0), (70_001, 0.25)):   # second call smaller: stale maxima must not leak
        x = torch.randn(count + off, device="cuda", generator=g) * scale
        qa, sa = _run(noalloc, handle.value, x, off, code)
        qb, sb = _run(lambda *a: B.quantize_fp8(*a), handle.value, x, off, code)
        torch.cuda.synchronize()
        assert torch.equal(qa, qb), int((qa != qb).sum())
        assert torch.equal(sa, sb), (float(sa), float(sb))
        amax = float(x[off:].abs().max())
        fmax = {"FP8_E4M3FN": 448.0, "FP8_E5M2": 57344.0}[qt]
        assert float(sa) == np.float32(np.float32(amax) / np.float32(fmax))


def test_no_alloc_quantizer_needs_no_workspace():
    """With a borrowed 64-byte workspace the workspace entry fails (it needs 4 KiB) and the no-allocation entry
    still gives the codes computed earlier by the workspace entry."""
    noalloc = _no_alloc()
    h = B.Handle()
    stream = torch.cuda.current_stream()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        x = torch.randn(123_457, device="cuda") * 3.0
        q0, s0 = _run(lambda *a: B.quantize_fp8(*a), h.value, x, 0, B.ACL_FLOAT8_E4M3FN)
        tiny = torch.zeros(16, device="cuda")
        assert B.load().aclblasSetWorkspace(h.value, ctypes.c_void_p(tiny.data_ptr()), ctypes.c_size_t(64)) == 0
        q = torch.zeros(x.numel() + 3, dtype=torch.uint8, device="cuda")
        s = torch.zeros(1, device="cuda")
        assert B.quantize_fp8(h.value, x.numel(), x.data_ptr(), B.ACL_FLOAT8_E4M3FN, q.data_ptr() + 1,
                              s.data_ptr()) == B.STATUS_CODE["ALLOC_FAILED"]
        q1, s1 = _run(noalloc, h.value, x, 0, B.ACL_FLOAT8_E4M3FN)
        torch.cuda.synchronize()
        assert torch.equal(q0, q1) and torch.equal(s0, s1)
    finally:
        torch.cuda.synchronize()
        h.destroy()


def test_no_alloc_quantizer_argument_checks(handle):
    noalloc = _no_alloc()
    D = 16
    assert noalloc(None, 4, D, B.ACL_FLOAT8_E4M3FN, D, D) == B.STATUS_CODE["HANDLE_IS_NULLPTR"]
    assert noalloc(handle.value, -1, D, B.ACL_FLOAT8_E4M3FN, D, D) == B.STATUS_CODE["INVALID_VALUE"]
    assert noalloc(handle.value, 4, D, B.ACL_BF16, D, D) == B.STATUS_CODE["INVALID_VALUE"]
    assert noalloc(handle.value, 4, D, B.ACL_FLOAT8_E4M3FN, D, None) == B.STATUS_CODE["INVALID_VALUE"]

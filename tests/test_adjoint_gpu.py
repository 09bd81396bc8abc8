"""GPU parity of aclblasSgetrsBatchedAdjoint, wrapped in a torch.autograd.Function.

The Function lives here, in the test harness only: forward is
aclblasSgetrfBatched + aclblasSgetrsBatched (X = A^-1 B), backward is one call
of aclblasSgetrsBatchedAdjoint (dB = A^-T dX by a transposed solve with the
saved factors and pivots, dA = -dB X^T by a batched GEMM).  Checks:
  - dB bit-identical to the oracle's transposed solve on the same factors;
  - dA against the oracle's FP64-accumulated -dB X^T (relative 1e-5);
  - dA and dB against a CPU FP64 gradient of torch.linalg.solve, with a
    tolerance scaled by the condition number;
  - a few entries against FP64 central finite differences.
Workload of the adjoint benchmark case: 2048 systems, n = 128, nrhs = 8,
A ~ N(0,1) + 8I, B ~ N(0,1), dX ~ N(0,1).
"""
import ctypes

import numpy as np
import pytest
import torch

import aclblas as B
import oracle as O

pytestmark = pytest.mark.gpu


class LUSolve(torch.autograd.Function):
    """X = A^-1 B for a batch; A (batch, n, n) and B (batch, n, nrhs) row-major torch tensors."""

    @staticmethod
    def forward(ctx, A, Bm, handle):
        batch, n, _ = A.shape
        nrhs = Bm.shape[2]
        LU = A.transpose(1, 2).contiguous()           # column-major storage [b, j, i]
        X = Bm.transpose(1, 2).contiguous()           # (batch, nrhs, n)
        piv = torch.zeros((batch, n), dtype=torch.int32, device=A.device)
        info = torch.zeros(batch, dtype=torch.int32, device=A.device)
        pLU, pX = B.ptr_array(LU), B.ptr_array(X)
        assert B.sgetrf_batched(handle, n, pLU.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch) == 0
        hinfo = ctypes.c_int(-1)
        assert B.sgetrs_batched(handle, B.OP_N, n, nrhs, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                hinfo, batch) == 0
        ctx.save_for_backward(LU, piv, X)
        ctx.handle = handle
        return X.transpose(1, 2)

    @staticmethod
    def backward(ctx, gX):
        LU, piv, X = ctx.saved_tensors
        batch, n, _ = LU.shape
        nrhs = X.shape[1]
        dB = gX.transpose(1, 2).contiguous()          # column-major n x nrhs, overwritten with A^-T dX
        dA = torch.empty((batch, n, n), dtype=torch.float32, device=LU.device)
        pLU, pX, pdB, pdA = B.ptr_array(LU), B.ptr_array(X), B.ptr_array(dB), B.ptr_array(dA)
        st = B.sgetrs_batched_adjoint(ctx.handle, n, nrhs, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                      pdB.data_ptr(), n, pdA.data_ptr(), n, batch)
        assert st == 0, st
        ctx.keep = (pLU, pX, pdB, pdA)
        return dA.transpose(1, 2), dB.transpose(1, 2), None


def _inputs(batch, n, nrhs, seed):
    A = torch.from_numpy(O.normal((batch, n, n), seed, n=n, batch=batch, shift=8.0)).transpose(1, 2).contiguous()
    Bm = torch.from_numpy(O.normal((batch, n, nrhs), seed + 1))
    dX = torch.from_numpy(O.normal((batch, n, nrhs), seed + 2))
    return A, Bm, dX


@pytest.mark.parametrize("batch,n,nrhs", [(2048, 128, 8), (96, 32, 3), (17, 5, 1), (9, 64, 16), (4, 200, 8)])
def test_adjoint_autograd(handle, batch, n, nrhs):
    A, Bm, dX = _inputs(batch, n, nrhs, 11 + n)
    dev = torch.device("cuda")
    Ad = A.to(dev).requires_grad_(True)
    Bd = Bm.to(dev).requires_grad_(True)
    X = LUSolve.apply(Ad, Bd, handle.value)
    (X * dX.to(dev)).sum().backward()
    torch.cuda.synchronize()
    gA = Ad.grad.cpu().double()
    gB = Bd.grad.cpu().double()

    # oracle on the same factors (the GPU's: for 64 < n <= 256 its LU is blocked and rounds differently from
    # the oracle's unblocked order): dB bit-exact, dA to FP64-accumulation accuracy
    LUd = A.transpose(1, 2).contiguous().to(dev)
    pivd = torch.zeros((batch, n), dtype=torch.int32, device=dev)
    infod = torch.zeros(batch, dtype=torch.int32, device=dev)
    assert B.sgetrf_batched(handle.value, n, B.ptr_array(LUd).data_ptr(), n, pivd.data_ptr(), infod.data_ptr(),
                            batch) == 0
    torch.cuda.synchronize()
    LUo, pivo = LUd.cpu().numpy(), pivd.cpu().numpy()
    Xo = O.getrs_batched(0, LUo, pivo, Bm.transpose(1, 2).contiguous().numpy())
    dBo, dAo = O.getrs_adjoint(LUo, pivo, Xo, dX.transpose(1, 2).contiguous().numpy())
    assert np.array_equal(gB.float().transpose(1, 2).numpy(), dBo), "dB differs from the oracle's transposed solve"
    dAo_rm = torch.from_numpy(dAo).transpose(1, 2).double()
    scale = dAo_rm.abs().amax(dim=(1, 2), keepdim=True).clamp_min(1e-30)
    assert ((gA - dAo_rm).abs() / scale).max().item() < 1e-5

    # CPU FP64 gradient of torch.linalg.solve on the FP32 inputs
    A64 = A.double().requires_grad_(True)
    B64 = Bm.double().requires_grad_(True)
    (torch.linalg.solve(A64, B64) * dX.double()).sum().backward()
    cond = torch.linalg.cond(A.double(), p=1)
    tolA = (cond * 2.0 ** -24 * 64).clamp_min(1e-6)
    errA = ((gA - A64.grad).abs().amax(dim=(1, 2)) / A64.grad.abs().amax(dim=(1, 2)))
    errB = ((gB - B64.grad).abs().amax(dim=(1, 2)) / B64.grad.abs().amax(dim=(1, 2)))
    assert (errA <= tolA).all(), (errA / tolA).max().item()
    assert (errB <= tolA).all(), (errB / tolA).max().item()

    # FP64 central finite differences on a few entries of the first system
    A0, B0, d0 = A[0].double().numpy(), Bm[0].double().numpy(), dX[0].double().numpy()
    h = 1e-6 * np.abs(A0).max()
    rng = np.random.default_rng(n)
    for _ in range(4):
        i, j = rng.integers(0, n, 2)
        E = np.zeros_like(A0)
        E[i, j] = h
        fd = (np.sum(d0 * np.linalg.solve(A0 + E, B0)) - np.sum(d0 * np.linalg.solve(A0 - E, B0))) / (2 * h)
        assert abs(gA[0, i, j].item() - fd) <= 1e-3 * max(1.0, abs(fd)), (i, j, gA[0, i, j].item(), fd)


def test_adjoint_without_dA(handle):
    """dAarray = NULL: only dB = A^-T dX, bit-identical to a plain transposed getrs."""
    batch, n, nrhs = 33, 48, 4
    A, Bm, dX = _inputs(batch, n, nrhs, 5)
    dev = torch.device("cuda")
    LU = A.transpose(1, 2).contiguous().to(dev)
    piv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    pLU = B.ptr_array(LU)
    assert B.sgetrf_batched(handle.value, n, pLU.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch) == 0
    X = Bm.transpose(1, 2).contiguous().to(dev)
    d1 = dX.transpose(1, 2).contiguous().to(dev)
    d2 = d1.clone()
    p1, p2, pX = B.ptr_array(d1), B.ptr_array(d2), B.ptr_array(X)
    assert B.sgetrs_batched_adjoint(handle.value, n, nrhs, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                    p1.data_ptr(), n, None, n, batch) == 0
    hinfo = ctypes.c_int(-1)
    assert B.sgetrs_batched(handle.value, B.OP_T, n, nrhs, pLU.data_ptr(), n, piv.data_ptr(), p2.data_ptr(), n,
                            hinfo, batch) == 0
    torch.cuda.synchronize()
    assert torch.equal(d1, d2)

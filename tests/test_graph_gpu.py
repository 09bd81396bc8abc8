"""HIP graph capture of the calls that stage a per-problem table (mixed-order solve, grouped GEMM), of a
plain LU + solve sequence and of the batched inverse.

Each call is captured once on a dedicated stream (the handle bound to it before capture, so capture sees
no stream synchronisation and no workspace growth), then the graph is replayed several times with new
right-hand sides / operands written between replays by ordinary stream work; every replay must equal an
eager call on the same inputs bit for bit.  No memset is recorded: the captured work is one host-to-device
copy of the table and the kernels.
"""
import ctypes

import numpy as np
import pytest
import torch

import aclblas as B
import workloads as W

pytestmark = pytest.mark.gpu


def _pack(systems, dev):
    offA, offB, a_parts, b_parts = [], [], [], []
    oa = ob = 0
    for s in systems:
        offA.append(oa)
        offB.append(ob)
        a_parts.append(s["A"])
        b_parts.append(s["B"])
        oa += s["A"].size
        ob += s["B"].size
    return (torch.from_numpy(np.concatenate(a_parts)).to(dev), torch.from_numpy(np.concatenate(b_parts)).to(dev),
            offA, offB)


def test_mixed_solve_graph_replays():
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        systems = W.conditioned_systems(48, seed=21)
        dA, dB, offA, offB = _pack(systems, dev)
        dX = torch.zeros_like(dB)
        ns = [s["n"] for s in systems]
        rs = [s["nrhs"] for s in systems]
        pa = [dA.data_ptr() + 4 * o for o in offA]
        pb = [dB.data_ptr() + 4 * o for o in offB]
        px = [dX.data_ptr() + 4 * o for o in offB]
        iters = torch.zeros(len(systems), dtype=torch.int32, device=dev)

        def solve(xptrs, it):
            return B.sgesv_batched_mixed(h.value, ns, rs, pa, ns, pb, ns, xptrs, ns, 1e-9, 8, it.data_ptr(), None,
                                         None)

        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            assert solve(px, iters) == 0          # eager first call: sizes the workspace
        stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream, capture_error_mode="relaxed"):
            assert solve(px, iters) == 0
        dX2 = torch.zeros_like(dB)
        px2 = [dX2.data_ptr() + 4 * o for o in offB]
        it2 = torch.zeros_like(iters)
        rng = np.random.default_rng(5)
        for rep in range(4):
            newB = torch.from_numpy(rng.standard_normal(dB.numel()).astype(np.float32)).to(dev)
            with torch.cuda.stream(stream):
                dB.copy_(newB)
                dX.zero_()
                g.replay()
                assert solve(px2, it2) == 0       # eager reference on the same inputs
            stream.synchronize()
            assert torch.equal(dX, dX2), "replay %d differs from the eager call" % rep
            assert torch.equal(iters, it2)
    finally:
        torch.cuda.synchronize()
        h.destroy()


def test_grouped_gemm_graph_replays():
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        shapes = [(300, 129, 77), (64, 64, 512), (129, 257, 33)]
        As = [torch.randn(m * k, device=dev) for m, n, k in shapes]
        Bs = [torch.randn(k * n, device=dev) for m, n, k in shapes]
        Cs = [torch.randn(m * n, device=dev) for m, n, k in shapes]
        C0 = [c.clone() for c in Cs]
        ms, nn, ks = [s[0] for s in shapes], [s[1] for s in shapes], [s[2] for s in shapes]

        def call(cs):
            return B.sgemm_grouped_batched(h.value, [B.OP_N] * 3, [B.OP_N] * 3, ms, nn, ks, [1.5] * 3,
                                           [a.data_ptr() for a in As], ms, [b.data_ptr() for b in Bs], ks,
                                           [0.5] * 3, [c.data_ptr() for c in cs], ms, [1] * 3)

        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            assert call(Cs) == 0
        stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream, capture_error_mode="relaxed"):
            assert call(Cs) == 0
        Ref = [torch.empty_like(c) for c in Cs]
        for rep in range(3):
            with torch.cuda.stream(stream):
                for a in As:
                    a.normal_()
                for c, c0, r in zip(Cs, C0, Ref):
                    c.copy_(c0)
                    r.copy_(c0)
                g.replay()
                assert call(Ref) == 0
            stream.synchronize()
            for c, r in zip(Cs, Ref):
                assert torch.equal(c, r), "replay %d differs" % rep
    finally:
        torch.cuda.synchronize()
        h.destroy()


def test_lu_solve_sequence_graph_replays():
    """aclblasSgetrfBatched + aclblasSgetrsBatched captured as one factor + solve sequence."""
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        batch, n = 512, 32
        A0, B0 = W.lu_batch(batch, n, 2, seed=3)
        dA0 = torch.from_numpy(A0).to(dev)
        dB0 = torch.from_numpy(B0).to(dev)
        LU, X = dA0.clone(), dB0.clone()
        piv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
        info = torch.zeros(batch, dtype=torch.int32, device=dev)
        pLU, pX = B.ptr_array(LU), B.ptr_array(X)
        hinfo = ctypes.c_int(0)

        def seq():
            st = B.sgetrf_batched(h.value, n, pLU.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch)
            st |= B.sgetrs_batched(h.value, B.OP_N, n, 2, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                   hinfo, batch)
            return st

        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            assert seq() == 0
        LU2, X2 = dA0.clone(), dB0.clone()
        piv2 = torch.zeros_like(piv)
        pLU2, pX2 = B.ptr_array(LU2), B.ptr_array(X2)
        for rep in range(3):
            with torch.cuda.stream(stream):
                LU.copy_(dA0 + rep)
                X.copy_(dB0 * (rep + 1))
                LU2.copy_(dA0 + rep)
                X2.copy_(dB0 * (rep + 1))
                g.replay()
                st = B.sgetrf_batched(h.value, n, pLU2.data_ptr(), n, piv2.data_ptr(), info.data_ptr(), batch)
                st |= B.sgetrs_batched(h.value, B.OP_N, n, 2, pLU2.data_ptr(), n, piv2.data_ptr(), pX2.data_ptr(),
                                       n, hinfo, batch)
                assert st == 0
            stream.synchronize()
            assert torch.equal(LU, LU2) and torch.equal(piv, piv2) and torch.equal(X, X2), rep
    finally:
        torch.cuda.synchronize()
        h.destroy()


def test_matinv_graph_replays():
    # the batched inverse allocates nothing and records no memset: two kernels
    # (out-of-place LU into the handle's workspace, inverse into Ainv)
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        batch, n = 512, 32
        rng = np.random.default_rng(9)
        dA = torch.from_numpy((rng.standard_normal((batch, n, n)) + 4 * np.eye(n)).astype(np.float32)).to(dev)
        C = torch.zeros_like(dA)
        C2 = torch.zeros_like(dA)
        info = torch.full((batch,), -1, dtype=torch.int32, device=dev)
        info2 = torch.full_like(info, -1)
        pa, pc, pc2 = B.ptr_array(dA), B.ptr_array(C), B.ptr_array(C2)

        def inv(out, inf):
            return B.smatinv_batched(h.value, n, pa.data_ptr(), n, out.data_ptr(), n, inf.data_ptr(), batch)

        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream, capture_error_mode="relaxed"):
            assert inv(pc, info) == 0
        for rep in range(4):
            newA = rng.standard_normal((batch, n, n)) + 4 * np.eye(n)
            if rep == 2:
                newA[7, 3, :] = 0.0                 # a singular member: info 4
            with torch.cuda.stream(stream):
                dA.copy_(torch.from_numpy(newA.astype(np.float32)).to(dev))
                C.zero_()
                g.replay()
                assert inv(pc2, info2) == 0        # eager reference on the same inputs
            stream.synchronize()
            assert torch.equal(C, C2), "replay %d differs from the eager call" % rep
            assert torch.equal(info, info2)
            assert int(info[7]) == (4 if rep == 2 else 0)
    finally:
        torch.cuda.synchronize()
        h.destroy()


def test_graph_replays_after_workspace_growth():
    """A graph that captured the mixed solve (its table and factors live in the handle's arena) keeps working
    after a later eager call grew the arena: the captured arena is kept until aclblasDestroy, so the replays
    touch no freed memory and still equal eager calls bit for bit."""
    dev = torch.device("cuda")
    stream = torch.cuda.Stream()
    h = B.Handle()
    assert h.set_stream(stream.cuda_stream) == 0
    try:
        systems = W.conditioned_systems(24, seed=31)
        dA, dB, offA, offB = _pack(systems, dev)
        dX = torch.zeros_like(dB)
        ns = [s["n"] for s in systems]
        rs = [s["nrhs"] for s in systems]
        pa = [dA.data_ptr() + 4 * o for o in offA]
        pb = [dB.data_ptr() + 4 * o for o in offB]
        px = [dX.data_ptr() + 4 * o for o in offB]
        iters = torch.zeros(len(systems), dtype=torch.int32, device=dev)

        def solve(xptrs, it, count=len(systems)):
            return B.sgesv_batched_mixed(h.value, ns[:count], rs[:count], pa[:count], ns[:count], pb[:count],
                                         ns[:count], xptrs[:count], ns[:count], 1e-9, 8, it.data_ptr(), None, None)

        torch.cuda.synchronize()
        with torch.cuda.stream(stream):
            assert solve(px, iters) == 0
        stream.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream, capture_error_mode="relaxed"):
            assert solve(px, iters) == 0
        # an eager call whose workspace needs exceed the arena: 400 systems of order 256 (about 110 MiB of
        # factors) make the handle replace its 32 MiB arena
        big = W.conditioned_systems(400, seed=32, nmin=256, nmax=256, rmin=1, rmax=1)
        bA, bB, boA, boB = _pack(big, dev)
        bX = torch.zeros_like(bB)
        bn = [256] * len(big)
        br = [1] * len(big)
        with torch.cuda.stream(stream):
            assert B.sgesv_batched_mixed(h.value, bn, br, [bA.data_ptr() + 4 * o for o in boA], bn,
                                         [bB.data_ptr() + 4 * o for o in boB], bn,
                                         [bX.data_ptr() + 4 * o for o in boB], bn, 1e-6, 4, None, None, None) == 0
        stream.synchronize()
        dX2 = torch.zeros_like(dB)
        px2 = [dX2.data_ptr() + 4 * o for o in offB]
        it2 = torch.zeros_like(iters)
        rng = np.random.default_rng(6)
        for rep in range(3):
            newB = torch.from_numpy(rng.standard_normal(dB.numel()).astype(np.float32)).to(dev)
            with torch.cuda.stream(stream):
                dB.copy_(newB)
                dX.zero_()
                g.replay()
                assert solve(px2, it2) == 0
            stream.synchronize()
            assert torch.equal(dX, dX2), "replay %d after the arena grew differs from the eager call" % rep
            assert torch.equal(iters, it2)
    finally:
        torch.cuda.synchronize()
        h.destroy()

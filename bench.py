"""Benchmark of the aclblas MI355X build on one GPU.

Primary workload (one "step", the `value` of the JSON line): FP32 LU with
partial pivoting and solve of 4096 systems of order 32 with one right-hand
side, A ~ N(0,1) + 4I, b ~ N(0,1), through aclblasSgetrfBatched +
aclblasSgetrsBatched.  Inputs are resident in HBM before the timed region;
every step factors its own fresh copy of the batch (getrf works in place).

The same JSON line carries
  roofline      the dominant kernel (getrf) against HBM bandwidth: algorithmic
                bytes per launch / its average duration measured with HIP events
                on the stream it runs on; traffic = HBM bytes per launch from the
                committed FETCH_SIZE / WRITE_SIZE summary of this workload
  cpu_baseline  the CPU restatement of the reference's algorithm (oracle/),
                timed on this host's cores on the same workload
  secondary     the other measured workloads (skipped with --primary-only):
                mixed-order solve with BF16 matrix-core LU and FP64 refinement
                against the plain FP32 getrf+getrs path, GEMM TFLOP/s (FP32
                512^3, BF16 -> FP32 8192^3, grouped FP32), adjoint of a batched
                solve.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ops-blas_amd"))

METRIC = "batched LU solves/s BF16+refine vs FP32 at equal backward error; GEMM TFLOP/s"
BATCH, N, NRHS = 4096, 32, 1
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic bytes per system: getrf reads A and writes LU (n*n*4 each), writes
# n pivots and one info word, reads one 8-byte pointer
GETRF_BYTES = 2 * N * N * 4 + N * 4 + 4 + 8
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "getrf_wave32_batch4096_traffic.md")


def measured_traffic():
    """HBM bytes per getrf launch from the committed FETCH_SIZE/WRITE_SIZE summary of this workload."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--primary-only", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--profile-gemm", action="store_true",
                    help="only run three BF16 x BF16 -> FP32 8192^3 GEMMs (for counter collection) and exit")
    ap.add_argument("--profile-fp8", action="store_true",
                    help="only quantise two 8192^2 FP32 operands to E4M3 and run three scaled FP8 8192^3 GEMMs "
                         "(BF16 C) and exit")
    ap.add_argument("--profile-mixed", action="store_true",
                    help="only run the mixed-order solve of 10000 systems three times (for profiling) and exit")
    ap.add_argument("--profile-matinv", action="store_true",
                    help="only run the 4096 x 32 batched inverse (21 calls, for profiling) and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before warm-up, run 5 elementwise negations of a MIB-MiB FP32 tensor (read and write "
                         "MIB MiB each): the known-size reference that scales hardware byte counters")
    return ap.parse_args()


def cpu_baseline(A, Bv):
    """Times the oracle's getrf+getrs (the reference's algorithm restated in C++) on this host."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    cores = min(16, os.cpu_count() or 1, O.threads())
    t1 = O.time_getrf_getrs(A, Bv, cores, reps=1)
    reps = max(3, min(200, int(10.0 / max(t1, 1e-6))))
    t = O.time_getrf_getrs(A, Bv, cores, reps=reps)
    return {"value": A.shape[0] / t, "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": "full workload (%d systems, n=%d, nrhs=%d), median of %d passes" % (A.shape[0], N, NRHS,
                                                                                       reps)}


def backward_error_batch(A, Bv, X):
    """max normwise backward error over a fixed-order batch (FP64, numpy); layouts as lu_batch."""
    import numpy as np
    Am = np.transpose(A.astype(np.float64), (0, 2, 1))            # row-major (batch, n, n)
    Bm = np.transpose(Bv.astype(np.float64), (0, 2, 1))           # (batch, n, nrhs)
    Xm = np.transpose(X.astype(np.float64), (0, 2, 1))
    R = Bm - Am @ Xm
    an = np.abs(Am).sum(axis=2).max(axis=1)[:, None]
    den = an * np.abs(Xm).max(axis=1) + np.abs(Bm).max(axis=1)
    return float((np.abs(R).max(axis=1) / den).max())


class Timer:
    """HIP events on the given stream around a callable."""

    def __init__(self, torch, stream):
        self.torch = torch
        self.stream = stream

    def __call__(self, fn, reps=1):
        t = self.torch
        e0, e1 = t.cuda.Event(enable_timing=True), t.cuda.Event(enable_timing=True)
        t.cuda.synchronize()
        e0.record(self.stream)
        for _ in range(reps):
            fn()
        e1.record(self.stream)
        t.cuda.synchronize()
        return e0.elapsed_time(e1) / reps


def bench_mixed(torch, B, W, h, timer, plain=True):
    """10000 mixed-order systems (n 16..256, nrhs 1..16, cond 1e2..1e6): BF16 matrix-core LU + FP64
    refinement in two persistent launches, against the plain FP32 path (aclblasSgetrfBatched +
    aclblasSgetrsBatched per order, right-hand sides padded to the order's largest nrhs)."""
    import numpy as np
    systems = W.conditioned_systems(10000, seed=0)
    dev = torch.device("cuda")
    count = len(systems)
    offA, offB = [], []
    oa = ob = 0
    for s in systems:
        offA.append(oa)
        offB.append(ob)
        oa += s["n"] * s["n"]
        ob += s["n"] * 16
    hA = np.zeros(oa, np.float32)
    hB = np.zeros(ob, np.float32)
    for s, a, b in zip(systems, offA, offB):
        hA[a:a + s["A"].size] = s["A"]
        hB[b:b + s["B"].size] = s["B"]
    dA = torch.from_numpy(hA).to(dev)
    dB = torch.from_numpy(hB).to(dev)
    dX = torch.zeros_like(dB)
    ns = [s["n"] for s in systems]
    rs = [s["nrhs"] for s in systems]
    pa = [dA.data_ptr() + 4 * o for o in offA]
    pb = [dB.data_ptr() + 4 * o for o in offB]
    px = [dX.data_ptr() + 4 * o for o in offB]
    iters = torch.zeros(count, dtype=torch.int32, device=dev)
    berr = torch.zeros(count, dtype=torch.float32, device=dev)
    info = torch.zeros(count, dtype=torch.int32, device=dev)

    # host arrays built once, as a C caller holds them (the timed call still sorts the systems, builds and
    # stages the per-system table)
    hns, hrs = B.host_ints(ns), B.host_ints(rs)
    hpa, hpb, hpx = B.host_ptrs(pa), B.host_ptrs(pb), B.host_ptrs(px)

    def mixed():
        st = B.sgesv_batched_mixed(h.value, hns, hrs, hpa, hns, hpb, hns, hpx, hns, 1e-6, 10, iters.data_ptr(),
                                   berr.data_ptr(), info.data_ptr())
        assert st == 0, st

    mixed()
    mixed_ms = min(timer(mixed) for _ in range(3))
    X = dX.cpu().numpy()
    be_mixed = max(W.normwise_backward_error(s["A"], s["n"], s["B"], X[o:o + s["n"] * s["nrhs"]], s["nrhs"])
                   for s, o in zip(systems, offB))

    if not plain:
        return {"systems": count, "mixed_ms": mixed_ms, "mixed_max_backward_error": be_mixed}
    # plain FP32 path: per order, copies of A (getrf is in place) and padded right-hand sides
    groups = {}
    for i, s in enumerate(systems):
        groups.setdefault(s["n"], []).append(i)
    plan = []
    for n, members in sorted(groups.items()):
        r = max(systems[i]["nrhs"] for i in members)
        LU = torch.empty((len(members), n, n), dtype=torch.float32, device=dev)
        Xg = torch.zeros((len(members), r, n), dtype=torch.float32, device=dev)
        srcA = torch.stack([dA[offA[i]:offA[i] + n * n].view(n, n) for i in members])
        srcB = torch.zeros((len(members), r, n), dtype=torch.float32, device=dev)
        for t, i in enumerate(members):
            k = systems[i]["nrhs"]
            srcB[t, :k] = dB[offB[i]:offB[i] + n * k].view(k, n)
        piv = torch.zeros((len(members), n), dtype=torch.int32, device=dev)
        inf = torch.zeros(len(members), dtype=torch.int32, device=dev)
        plan.append((n, r, members, LU, Xg, srcA, srcB, piv, inf, B.ptr_array(LU), B.ptr_array(Xg)))
    hinfo = ctypes.c_int(0)

    def refresh():
        for (n, r, members, LU, Xg, srcA, srcB, piv, inf, pLU, pX) in plan:
            LU.copy_(srcA)
            Xg.copy_(srcB)

    def plain():
        for (n, r, members, LU, Xg, srcA, srcB, piv, inf, pLU, pX) in plan:
            st = B.sgetrf_batched(h.value, n, pLU.data_ptr(), n, piv.data_ptr(), inf.data_ptr(), len(members))
            st |= B.sgetrs_batched(h.value, B.OP_N, n, r, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                   hinfo, len(members))
            assert st == 0, st

    refresh()
    plain()
    times = []
    for _ in range(3):
        refresh()
        times.append(timer(plain))
    plain_ms = min(times)
    be_plain = 0.0
    for (n, r, members, LU, Xg, srcA, srcB, piv, inf, pLU, pX) in plan:
        Xh = Xg.cpu().numpy()
        for t, i in enumerate(members):
            s = systems[i]
            be_plain = max(be_plain, W.normwise_backward_error(s["A"], n, s["B"], Xh[t, :s["nrhs"]].reshape(-1),
                                                               s["nrhs"]))
    it = iters.cpu().numpy()
    return {"systems": count, "n": "16..256", "nrhs": "1..16", "cond": "1e2..1e6", "tol": 1e-6,
            "mixed_ms": mixed_ms, "mixed_solves_per_s": count / (mixed_ms * 1e-3),
            "mixed_max_backward_error": be_mixed, "mixed_mean_corrections": float(it.mean()),
            "plain_fp32_ms": plain_ms, "plain_fp32_solves_per_s": count / (plain_ms * 1e-3),
            "plain_fp32_max_backward_error": be_plain, "speedup": plain_ms / mixed_ms,
            "plain_fp32_launches": 2 * len(plan)}


def bench_gemm(torch, B, W, h, timer):
    import numpy as np
    dev = torch.device("cuda")
    out = {}
    # FP32 512^3 NN, alpha 1, beta 0.5, U[-1,1]
    m = 512
    a = (torch.rand(m * m, device=dev) * 2 - 1)
    b = (torch.rand(m * m, device=dev) * 2 - 1)
    c = (torch.rand(m * m, device=dev) * 2 - 1)
    f = lambda: B.sgemm(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, a.data_ptr(), m, b.data_ptr(), m, 0.5,
                        c.data_ptr(), m)
    f()
    ms = timer(f, 200)
    out["sgemm_512"] = {"us": ms * 1e3, "tflops": 2 * m ** 3 / (ms * 1e-3) / 1e12}
    # BF16 x BF16 -> FP32, 8192^3, N(0,1)
    m = 8192
    a = torch.randn(m * m, device=dev).to(torch.bfloat16)
    b = torch.randn(m * m, device=dev).to(torch.bfloat16)
    c = torch.zeros(m * m, device=dev)
    f = lambda: B.gemm_ex(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, a.data_ptr(), B.ACL_BF16, m, b.data_ptr(),
                          B.ACL_BF16, m, 0.0, c.data_ptr(), B.ACL_FLOAT, m, B.COMPUTE["COMPUTE_32F"])
    f()
    ms = timer(f, 10)
    tf = 2 * m ** 3 / (ms * 1e-3) / 1e12
    out["bf16_fp32_8192"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / BF16_PEAK_TFLOPS}
    # FP8 E4M3 with per-tensor amax scales, 8192^3: quantise A and B (timed separately), scaled GEMM -> BF16
    af = torch.randn(m * m, device=dev)
    bf = torch.randn(m * m, device=dev)
    qa = torch.empty(m * m, dtype=torch.uint8, device=dev)
    qb = torch.empty(m * m, dtype=torch.uint8, device=dev)
    sc = torch.empty(2, device=dev)
    cb = torch.empty(m * m, dtype=torch.bfloat16, device=dev)
    e4 = B.ACL_FLOAT8_E4M3FN

    def quant():
        assert B.quantize_fp8(h.value, m * m, af.data_ptr(), e4, qa.data_ptr(), sc.data_ptr()) == 0
        assert B.quantize_fp8(h.value, m * m, bf.data_ptr(), e4, qb.data_ptr(), sc.data_ptr() + 4) == 0

    f = lambda: B.gemm_ex_scaled(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, qa.data_ptr(), e4, m, sc.data_ptr(),
                                 qb.data_ptr(), e4, m, sc.data_ptr() + 4, 0.0, cb.data_ptr(), B.ACL_BF16, m)
    quant()
    f()
    q_ms = timer(quant, 10)
    ms = timer(f, 10)
    tf = 2 * m ** 3 / (ms * 1e-3) / 1e12
    out["fp8_scaled_8192"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / FP8_PEAK_TFLOPS,
                              "quantize_both_ms": q_ms}
    del af, bf, qa, qb, cb
    # FP32 8192^3 (aclblasSgemm), N(0,1)
    a = torch.randn(m * m, device=dev)
    b = torch.randn(m * m, device=dev)
    f = lambda: B.sgemm(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, a.data_ptr(), m, b.data_ptr(), m, 0.0,
                        c.data_ptr(), m)
    f()
    ms = timer(f, 5)
    tf = 2 * m ** 3 / (ms * 1e-3) / 1e12
    out["fp32_8192"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / FP32_PEAK_TFLOPS}
    del a, b, c
    # grouped FP32: 64 groups, m, n, k in {64, 80, ..., 2048}, seed 0, one problem per group
    rng = np.random.default_rng(0)
    mnk = rng.integers(0, 125, size=(64, 3)) * 16 + 64
    As, Bs, Cs = [], [], []
    flops = 0
    for mm, nn, kk in mnk:
        As.append(torch.rand(int(mm * kk), device=dev))
        Bs.append(torch.rand(int(kk * nn), device=dev))
        Cs.append(torch.rand(int(mm * nn), device=dev))
        flops += 2 * int(mm) * int(nn) * int(kk)
    g = [int(x) for x in mnk[:, 0]], [int(x) for x in mnk[:, 1]], [int(x) for x in mnk[:, 2]]
    f = lambda: B.sgemm_grouped_batched(h.value, [B.OP_N] * 64, [B.OP_N] * 64, g[0], g[1], g[2], [1.0] * 64,
                                        [x.data_ptr() for x in As], g[0], [x.data_ptr() for x in Bs], g[2],
                                        [0.5] * 64, [x.data_ptr() for x in Cs], g[0], [1] * 64)
    f()
    ms = timer(f, 5)
    tf = flops / (ms * 1e-3) / 1e12
    out["grouped_fp32_64"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / FP32_PEAK_TFLOPS, "flop": flops}
    # the same 64 shapes in BF16 (BF16 C) through aclblasGemmGroupedBatchedEx
    del As, Bs, Cs
    Ah = [torch.randn(int(mm * kk), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    Bh = [torch.randn(int(kk * nn), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    Ch = [torch.randn(int(mm * nn), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    ptr = lambda ts: torch.tensor([x.data_ptr() for x in ts], dtype=torch.int64, device=dev)
    pa, pb, pc = ptr(Ah), ptr(Bh), ptr(Ch)
    f = lambda: B.gemm_grouped_batched_ex(h.value, [B.OP_N] * 64, [B.OP_N] * 64, g[0], g[1], g[2], [1.0] * 64,
                                          pa.data_ptr(), B.ACL_BF16, g[0], pb.data_ptr(), B.ACL_BF16, g[2],
                                          [0.5] * 64, pc.data_ptr(), B.ACL_BF16, g[0], [1] * 64,
                                          B.COMPUTE["COMPUTE_32F"])
    assert f() == 0
    ms = timer(f, 5)
    tf = flops / (ms * 1e-3) / 1e12
    out["grouped_bf16_ex_64"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / BF16_PEAK_TFLOPS, "flop": flops}
    return out


def bench_adjoint(torch, B, W, h, timer):
    """2048 systems n=128 nrhs=8: forward getrf+getrs and the adjoint (transposed solve + batched GEMM)."""
    dev = torch.device("cuda")
    batch, n, r = 2048, 128, 8
    A, Bv = W.lu_batch(batch, n, r, seed=1, shift=8.0)
    dX = torch.randn((batch, r, n), device=dev)
    A0 = torch.from_numpy(A).to(dev)
    B0 = torch.from_numpy(Bv).to(dev)
    LU, X = A0.clone(), B0.clone()
    G = dX.clone()
    dA = torch.empty_like(A0)
    piv = torch.zeros((batch, n), dtype=torch.int32, device=dev)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    pLU, pX, pG, pdA = B.ptr_array(LU), B.ptr_array(X), B.ptr_array(G), B.ptr_array(dA)
    hinfo = ctypes.c_int(0)

    def fwd():
        st = B.sgetrf_batched(h.value, n, pLU.data_ptr(), n, piv.data_ptr(), info.data_ptr(), batch)
        st |= B.sgetrs_batched(h.value, B.OP_N, n, r, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n, hinfo,
                               batch)
        assert st == 0, st

    def bwd():
        st = B.sgetrs_batched_adjoint(h.value, n, r, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                      pG.data_ptr(), n, pdA.data_ptr(), n, batch)
        assert st == 0, st

    LU.copy_(A0)
    X.copy_(B0)
    fwd_ms = timer(fwd)
    G.copy_(dX)
    bwd_ms = timer(bwd)
    return {"batch": batch, "n": n, "nrhs": r, "forward_ms": fwd_ms, "adjoint_ms": bwd_ms,
            "forward_solves_per_s": batch / (fwd_ms * 1e-3)}


def bench_matinv(torch, B, W, h, timer):
    """4096 inverses of order 32 (aclblasSmatinvBatched): stage + LU + inverse launches."""
    dev = torch.device("cuda")
    batch, n = 4096, 32
    A, _ = W.lu_batch(batch, n, 1, seed=3, shift=4.0)
    dA = torch.from_numpy(A).to(dev)
    C = torch.empty_like(dA)
    info = torch.zeros(batch, dtype=torch.int32, device=dev)
    pA, pC = B.ptr_array(dA), B.ptr_array(C)

    def inv():
        st = B.smatinv_batched(h.value, n, pA.data_ptr(), n, pC.data_ptr(), n, info.data_ptr(), batch)
        assert st == 0, st

    inv()
    ms = timer(inv, reps=20)
    assert int(info.abs().sum().item()) == 0
    # algorithmic bytes: A read, Ainv written, info, two pointers per matrix
    per = 2 * n * n * 4 + 4 + 16
    return {"batch": batch, "n": n, "ms": ms, "inverses_per_s": batch / (ms * 1e-3),
            "hbm_frac_of_algorithmic_bytes": per * batch / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import aclblas as B
    import workloads as W

    B.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.profile_mixed:
        h = B.Handle()
        stream = torch.cuda.current_stream()
        h.set_stream(stream.cuda_stream)
        res = bench_mixed(torch, B, W, h, Timer(torch, stream), plain=False)
        h.destroy()
        print(json.dumps(res))
        return
    if args.profile_matinv:
        h = B.Handle()
        stream = torch.cuda.current_stream()
        h.set_stream(stream.cuda_stream)
        res = bench_matinv(torch, B, W, h, Timer(torch, stream))
        h.destroy()
        print(json.dumps(res))
        return
    if args.profile_fp8:
        h = B.Handle()
        h.set_stream(torch.cuda.current_stream().cuda_stream)
        m = 8192
        e4 = B.ACL_FLOAT8_E4M3FN
        qa = torch.empty(m * m, dtype=torch.uint8, device=dev)
        qb = torch.empty(m * m, dtype=torch.uint8, device=dev)
        sc = torch.empty(2, device=dev)
        cb = torch.empty(m * m, dtype=torch.bfloat16, device=dev)
        for q, seed in ((qa, 1), (qb, 2)):
            x = torch.randn(m * m, device=dev, generator=torch.Generator(device=dev).manual_seed(seed))
            assert B.quantize_fp8(h.value, m * m, x.data_ptr(), e4, q.data_ptr(),
                                  sc.data_ptr() + (0 if q is qa else 4)) == 0
            del x
        for _ in range(3):
            st = B.gemm_ex_scaled(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, qa.data_ptr(), e4, m, sc.data_ptr(),
                                  qb.data_ptr(), e4, m, sc.data_ptr() + 4, 0.0, cb.data_ptr(), B.ACL_BF16, m)
            assert st == 0, st
        torch.cuda.synchronize()
        h.destroy()
        print("fp8 profile run done")
        return
    if args.profile_gemm:
        h = B.Handle()
        h.set_stream(torch.cuda.current_stream().cuda_stream)
        if args.calibrate_copy > 0:
            src = torch.ones(args.calibrate_copy * (1 << 18), device=dev)
            dst = torch.empty_like(src)
            for _ in range(5):
                torch.neg(src, out=dst)
            torch.cuda.synchronize()
            del src, dst
        m = 8192
        a = torch.randn(m * m, device=dev).to(torch.bfloat16)
        b = torch.randn(m * m, device=dev).to(torch.bfloat16)
        c = torch.zeros(m * m, device=dev)
        for _ in range(3):
            st = B.gemm_ex(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, a.data_ptr(), B.ACL_BF16, m, b.data_ptr(),
                           B.ACL_BF16, m, 0.0, c.data_ptr(), B.ACL_FLOAT, m, B.COMPUTE["COMPUTE_32F"])
            assert st == 0, st
        torch.cuda.synchronize()
        h.destroy()
        print("gemm profile run done")
        return
    A, Bv = W.lu_batch(BATCH, N, NRHS, seed=0, shift=4.0)
    total = args.warmup + args.steps
    dA0 = torch.from_numpy(A).to(dev)
    dB0 = torch.from_numpy(Bv).to(dev)
    # one private copy of the inputs per step, all resident before timing
    dA = [dA0.clone() for _ in range(total)]
    dB = [dB0.clone() for _ in range(total)]
    pA = [B.ptr_array(x) for x in dA]
    pB = [B.ptr_array(x) for x in dB]
    piv = torch.zeros((BATCH, N), dtype=torch.int32, device=dev)
    info = torch.zeros(BATCH, dtype=torch.int32, device=dev)
    h = B.Handle()
    stream = torch.cuda.current_stream()
    h.set_stream(stream.cuda_stream)
    inf = ctypes.c_int(0)

    def step(i, ev=None):
        if ev is not None:
            ev[0].record(stream)
        st = B.sgetrf_batched(h.value, N, pA[i].data_ptr(), N, piv.data_ptr(), info.data_ptr(), BATCH)
        if ev is not None:
            ev[1].record(stream)
        st |= B.sgetrs_batched(h.value, B.OP_N, N, NRHS, pA[i].data_ptr(), N, piv.data_ptr(), pB[i].data_ptr(), N,
                               inf, BATCH)
        if st != 0:
            raise RuntimeError("aclblas call failed with status %d" % st)

    if args.calibrate_copy > 0:
        src = torch.ones(args.calibrate_copy * (1 << 18), device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(args.warmup + s, evs[s])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0

    getrf_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / len(evs)
    berr = backward_error_batch(A, Bv, dB[-1].cpu().numpy())
    solves = BATCH * args.steps / elapsed
    achieved = GETRF_BYTES * BATCH / (getrf_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": solves,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (A ~ N(0,1) + 4I, b ~ N(0,1), numpy seed 0)",
        "config": {"workload": "FP32 aclblasSgetrfBatched + aclblasSgetrsBatched, batch 4096 x 32x32, nrhs 1",
                   "batch": BATCH, "n": N, "nrhs": NRHS, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "getrf (n<=32, two systems per wave)", "kernel_ms": getrf_ms,
                     "algorithmic_bytes_per_launch": GETRF_BYTES * BATCH},
        "max_backward_error": berr,
    }
    if not args.primary_only:
        timer = Timer(torch, stream)
        sec = {}
        sec["mixed_solve"] = bench_mixed(torch, B, W, h, timer)
        sec["gemm"] = bench_gemm(torch, B, W, h, timer)
        sec["adjoint"] = bench_adjoint(torch, B, W, h, timer)
        sec["matinv"] = bench_matinv(torch, B, W, h, timer)
        out["secondary"] = sec
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, Bv)
    h.destroy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

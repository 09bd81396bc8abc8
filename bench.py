"""Benchmark of the aclblas MI355X build on one GPU.

Primary workload (one "step", the `value` of the JSON line): the mixed-size
batched solve with refinement - 10 000 systems, n uniform in [16, 256],
A = U diag(s) V^T with U, V from the QR of a Gaussian per system, cond(A)
log-uniform in [1e2, 1e6], nrhs uniform in [1, 16], B ~ N(0,1) - solved by
aclblasSgesvBatchedMixedEx with two-plane BF16 trailing updates on the matrix
cores, FP64-residual refinement to normwise backward error <= 1e-6 and the FP32
fallback for the systems refinement cannot finish.  Inputs are resident in HBM
before the timed region; one step is one call (A and B are not modified).

The same JSON line carries
  roofline      the dominant kernel of that call against HBM bandwidth:
                algorithmic bytes per launch / its average duration from HIP
                events recorded around each launch on the handle's stream
                (aclblasSetLaunchEvents); traffic = HBM bytes per launch from the
                committed FETCH_SIZE / WRITE_SIZE summary of this workload
  comparison    the same kernels with FP32 trailing updates on the same inputs
                and at the same tolerance, the other update precisions, and the
                plain FP32 public path (aclblasSgetrfBatched +
                aclblasSgetrsBatched per order)
  cpu_baseline  the CPU restatement of the reference's algorithm (oracle/:
                unblocked getrf + getrs per system), timed on this host's cores
                on the same systems
  secondary     the other measured workloads (skipped with --primary-only): the
                4096 x 32 FP32 getrf + getrs batch, GEMM TFLOP/s (FP32 512^3,
                BF16 -> FP32 8192^3, FP8 scaled 8192^3, FP32 8192^3, grouped),
                the adjoint of a batched solve, the batched inverse.
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
MIXED_COUNT, MIXED_TOL, MIXED_MAXITER = 10000, 1e-6, 10
C2_BATCH, C2_N, C2_NRHS = 4096, 32, 1
HBM_PEAK_GBS = 8000.0
BF16_PEAK_TFLOPS = 2500.0
FP8_PEAK_TFLOPS = 5000.0
FP32_PEAK_TFLOPS = 157.3
# algorithmic bytes per system of the 4096 x 32 getrf: A read and LU written (n*n*4 each), n pivots, one info
# word, one 8-byte pointer
C2_GETRF_BYTES = 2 * C2_N * C2_N * 4 + C2_N * 4 + 4 + 8
PROFILES = os.path.join(ROOT, "profiles")


def measured_traffic(name):
    """HBM bytes per launch from a committed FETCH_SIZE/WRITE_SIZE summary (line 'traffic_bytes_per_launch: X')."""
    try:
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return float(line.split(":")[1])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--update", default="BF16X2", choices=["FP32", "BF16", "BF16X2", "BF16X3"],
                    help="trailing-update precision of the headline mixed solve")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--primary-only", action="store_true", help="skip the comparison and secondary workloads")
    ap.add_argument("--profile-gemm", action="store_true",
                    help="only run three BF16 x BF16 -> FP32 8192^3 GEMMs (for counter collection) and exit")
    ap.add_argument("--profile-grouped", action="store_true",
                    help="only run the grouped GEMM workload three times in FP32 and three times in BF16 and exit")
    ap.add_argument("--profile-sgemm512", action="store_true",
                    help="only run 200 FP32 512^3 aclblasSgemm calls (NN, alpha 1, beta 0.5, U[-1,1]) and exit")
    ap.add_argument("--profile-fp8", action="store_true",
                    help="only quantise two 8192^2 FP32 operands to E4M3 and run three scaled FP8 8192^3 GEMMs "
                         "(BF16 C) and exit")
    ap.add_argument("--profile-mixed", action="store_true",
                    help="only run the mixed solve of the 10000 systems three times with --update (for profiling) "
                         "and exit")
    ap.add_argument("--profile-matinv", action="store_true",
                    help="only run the 4096 x 32 batched inverse (21 calls, for profiling) and exit")
    ap.add_argument("--profile-c2", action="store_true",
                    help="only run the 4096 x 32 FP32 getrf + getrs steps (for profiling) and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="before warm-up, run 5 elementwise negations of a MIB-MiB FP32 tensor (read and write "
                         "MIB MiB each): the known-size reference that scales hardware byte counters")
    return ap.parse_args()


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def cpu_cores():
    """Host threads the lease grants: the process's CPU affinity, capped by OMP_NUM_THREADS when the environment
    sets it (the per-GPU CPU share on a shared box)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def calibrate_copy(torch, dev, mib):
    src = torch.ones(mib * (1 << 18), device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    del src, dst


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


# ------------------------------------------------------------------ mixed-size solve (headline)
class MixedWorkload:
    """The 10 000 conditioned systems, packed on the device; host arrays built once, as a C caller holds them."""

    def __init__(self, torch, B, W, count=MIXED_COUNT, seed=0):
        import numpy as np
        self.torch, self.B, self.W = torch, B, W
        t0 = time.perf_counter()
        self.systems = W.conditioned_systems(count, seed=seed)
        log("generated %d systems in %.1f s" % (count, time.perf_counter() - t0))
        dev = torch.device("cuda")
        self.count = count
        offA, offB = [], []
        oa = ob = 0
        for s in self.systems:
            offA.append(oa)
            offB.append(ob)
            oa += s["n"] * s["n"]
            ob += s["n"] * s["nrhs"]
        hA = np.zeros(oa, np.float32)
        hB = np.zeros(ob, np.float32)
        for s, a, b in zip(self.systems, offA, offB):
            hA[a:a + s["A"].size] = s["A"]
            hB[b:b + s["B"].size] = s["B"]
        self.hA, self.hB, self.offA, self.offB = hA, hB, offA, offB
        self.dA = torch.from_numpy(hA).to(dev)
        self.dB = torch.from_numpy(hB).to(dev)
        self.dX = torch.zeros_like(self.dB)
        self.ns = [s["n"] for s in self.systems]
        self.rs = [s["nrhs"] for s in self.systems]
        self.hns, self.hrs = B.host_ints(self.ns), B.host_ints(self.rs)
        self.hpa = B.host_ptrs([self.dA.data_ptr() + 4 * o for o in offA])
        self.hpb = B.host_ptrs([self.dB.data_ptr() + 4 * o for o in offB])
        self.hpx = B.host_ptrs([self.dX.data_ptr() + 4 * o for o in offB])
        self.iters = torch.zeros(count, dtype=torch.int32, device=dev)
        self.berr = torch.zeros(count, dtype=torch.float32, device=dev)
        self.info = torch.zeros(count, dtype=torch.int32, device=dev)
        self.fallback = torch.zeros(count, dtype=torch.int32, device=dev)
        n = np.array(self.ns, dtype=np.float64)
        r = np.array(self.rs, dtype=np.float64)
        self.lu_flop = float((2.0 * n ** 3 / 3.0).sum())
        # factor kernel, per system: A read and the factors written (4 n^2 bytes each) plus the pivot order
        # (one int per row and panel) and the 32 x 32 inverse of each panel's unit lower diagonal block that the
        # solves use (4 KiB per panel)
        panels = np.ceil(n / 32.0)
        rowmap = np.array([sum(int(m) - k0 for k0 in range(0, int(m), 32)) for m in n], dtype=np.float64)
        self.factor_bytes = float((8.0 * n * n + 4.0 * rowmap + 4096.0 * panels).sum())
        # solve kernel (one solve and one FP64 residual, the pass every system makes): factors, inverses and A read,
        # B read twice, X written
        self.refine_bytes = float((8.0 * n * n + 4096.0 * panels + 4.0 * rowmap + 12.0 * n * r).sum())

    def call(self, update, events=None):
        B = self.B
        if events is not None:
            assert B.set_launch_events(self.h, events, len(events)) == 0
        st = B.sgesv_batched_mixed_ex(self.h, self.hns, self.hrs, self.hpa, self.hns, self.hpb, self.hns, self.hpx,
                                      self.hns, MIXED_TOL, MIXED_MAXITER, B.MIXED_UPDATE[update],
                                      self.iters.data_ptr(), self.berr.data_ptr(), self.info.data_ptr(),
                                      self.fallback.data_ptr())
        if st != 0:
            raise RuntimeError("aclblasSgesvBatchedMixedEx failed with status %d" % st)

    def max_backward_error(self):
        """Worst normwise backward error of the current X over all systems, recomputed on the host in FP64."""
        X = self.dX.cpu().numpy()
        W = self.W
        return max(W.normwise_backward_error(s["A"], s["n"], s["B"], X[o:o + s["n"] * s["nrhs"]], s["nrhs"])
                   for s, o in zip(self.systems, self.offB))

    def run(self, update, steps, warmup):
        """warmup + timed calls; per-launch HIP events on every timed call.  Returns the measurements."""
        import ctypes
        import numpy as np
        torch, B = self.torch, self.B
        stream = torch.cuda.current_stream()
        for _ in range(warmup):
            self.call(update)
        torch.cuda.synchronize()
        sets = []
        for _ in range(steps):
            evs = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
            for e in evs:
                e.record(stream)          # creates the event
            sets.append((evs, (ctypes.c_void_p * 8)(*[e.cuda_event for e in evs])))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for evs, arr in sets:
            self.call(update, arr)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        assert B.set_launch_events(self.h, None, 0) == 0
        fb = self.fallback.cpu().numpy()
        it = self.iters.cpu().numpy()
        info = self.info.cpu().numpy()
        ms = lambda a, b: float(np.mean([e[a].elapsed_time(e[b]) for e, _ in sets]))
        out = {"update": update, "ms_per_call": elapsed * 1e3 / steps,
               "solves_per_s": self.count * steps / elapsed,
               "factor_ms": ms(0, 1), "refine_ms": ms(2, 3),
               "fallback_systems": int(fb.sum()), "fallback_rate": float(fb.mean()),
               "mean_corrections": float(it.mean()), "max_corrections": int(it.max()),
               "nonzero_info": int((info != 0).sum()),
               "reported_max_backward_error": float(self.berr.max().item())}
        if update != "FP32":
            out["fallback_factor_ms"] = ms(4, 5)
            out["fallback_refine_ms"] = ms(6, 7)
        out["max_backward_error"] = self.max_backward_error()
        return out

    def plain_public_path(self, timer):
        """aclblasSgetrfBatched + aclblasSgetrsBatched per order (right-hand sides padded to the order's largest
        nrhs), the plain FP32 path a caller of the reference's API would take."""
        torch, B, W = self.torch, self.B, self.W
        dev = torch.device("cuda")
        groups = {}
        for i, s in enumerate(self.systems):
            groups.setdefault(s["n"], []).append(i)
        plan = []
        for n, members in sorted(groups.items()):
            r = max(self.systems[i]["nrhs"] for i in members)
            LU = torch.empty((len(members), n, n), dtype=torch.float32, device=dev)
            Xg = torch.zeros((len(members), r, n), dtype=torch.float32, device=dev)
            srcA = torch.stack([self.dA[self.offA[i]:self.offA[i] + n * n].view(n, n) for i in members])
            srcB = torch.zeros((len(members), r, n), dtype=torch.float32, device=dev)
            for t, i in enumerate(members):
                k = self.systems[i]["nrhs"]
                srcB[t, :k] = self.dB[self.offB[i]:self.offB[i] + n * k].view(k, n)
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
                st = B.sgetrf_batched(self.h, n, pLU.data_ptr(), n, piv.data_ptr(), inf.data_ptr(), len(members))
                st |= B.sgetrs_batched(self.h, B.OP_N, n, r, pLU.data_ptr(), n, piv.data_ptr(), pX.data_ptr(), n,
                                       hinfo, len(members))
                assert st == 0, st

        refresh()
        plain()
        times = []
        for _ in range(2):
            refresh()
            times.append(timer(plain))
        ms = min(times)
        be = 0.0
        for (n, r, members, LU, Xg, srcA, srcB, piv, inf, pLU, pX) in plan:
            Xh = Xg.cpu().numpy()
            for t, i in enumerate(members):
                s = self.systems[i]
                be = max(be, W.normwise_backward_error(s["A"], n, s["B"], Xh[t, :s["nrhs"]].reshape(-1),
                                                       s["nrhs"]))
        return {"ms_per_call": ms, "solves_per_s": self.count / (ms * 1e-3), "max_backward_error": be,
                "launches": 2 * len(plan)}

    def cpu_baseline(self):
        """The oracle's getrf + getrs (the reference's algorithm restated in C++) on a sample of the same systems,
        on every host core the lease grants.  The sample is the largest prefix whose single pass is estimated to
        take about 10 s (all 10 000 systems when that is faster)."""
        import numpy as np
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        cores = cpu_cores()
        args = lambda k: (self.hA, self.offA[:k], self.hB, self.offB[:k], self.ns[:k], self.rs[:k], cores)
        probe = min(500, self.count)
        t_probe = O.time_getrf_getrs_mixed(*args(probe), reps=1)
        k = self.count if t_probe * self.count / probe <= 10.0 else max(probe, int(probe * 10.0 / t_probe))
        t = O.time_getrf_getrs_mixed(*args(k), reps=3)
        return {"value": k / t, "unit": "solves/s", "cores": cores, "kind": "port",
                "sample": "first %d of the %d systems (same generator, n 16..256, nrhs 1..16), one getrf + getrs "
                          "pass per system, median of 3 passes" % (k, self.count)}


def bench_c2(torch, B, W, h, steps, warmup):
    """4096 FP32 systems of order 32, one right-hand side, aclblasSgetrfBatched + aclblasSgetrsBatched; every step
    on its own input copy (getrf works in place)."""
    dev = torch.device("cuda")
    A, Bv = W.lu_batch(C2_BATCH, C2_N, C2_NRHS, seed=0, shift=4.0)
    total = warmup + steps
    dA0 = torch.from_numpy(A).to(dev)
    dB0 = torch.from_numpy(Bv).to(dev)
    dA = [dA0.clone() for _ in range(total)]
    dB = [dB0.clone() for _ in range(total)]
    pA = [B.ptr_array(x) for x in dA]
    pB = [B.ptr_array(x) for x in dB]
    piv = torch.zeros((C2_BATCH, C2_N), dtype=torch.int32, device=dev)
    info = torch.zeros(C2_BATCH, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()
    inf = ctypes.c_int(0)

    def step(i, ev=None):
        if ev is not None:
            ev[0].record(stream)
        st = B.sgetrf_batched(h.value, C2_N, pA[i].data_ptr(), C2_N, piv.data_ptr(), info.data_ptr(), C2_BATCH)
        if ev is not None:
            ev[1].record(stream)
        st |= B.sgetrs_batched(h.value, B.OP_N, C2_N, C2_NRHS, pA[i].data_ptr(), C2_N, piv.data_ptr(),
                               pB[i].data_ptr(), C2_N, inf, C2_BATCH)
        if st != 0:
            raise RuntimeError("aclblas call failed with status %d" % st)

    for i in range(warmup):
        step(i)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        step(warmup + s, evs[s])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    getrf_ms = sum(e[0].elapsed_time(e[1]) for e in evs) / len(evs)
    berr = backward_error_batch(A, Bv, dB[-1].cpu().numpy())
    achieved = C2_GETRF_BYTES * C2_BATCH / (getrf_ms * 1e-3) / 1e9
    return {"workload": "FP32 aclblasSgetrfBatched + aclblasSgetrsBatched, batch 4096 x 32x32, nrhs 1",
            "solves_per_s": C2_BATCH * steps / elapsed, "ms_per_step": elapsed * 1e3 / steps,
            "max_backward_error": berr,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS,
                         "traffic": measured_traffic("getrf_wave32_batch4096_traffic.md"),
                         "kernel": "getrf (n<=32, two systems per wave)", "kernel_ms": getrf_ms,
                         "algorithmic_bytes_per_launch": C2_GETRF_BYTES * C2_BATCH}}


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
    f32_call, bf16_call, flops, keep = grouped_legs(torch, B, h, dev)
    f32_call()
    ms = timer(f32_call, 5)
    tf = flops / (ms * 1e-3) / 1e12
    out["grouped_fp32_64"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / FP32_PEAK_TFLOPS, "flop": flops}
    assert bf16_call() == 0
    ms = timer(bf16_call, 5)
    tf = flops / (ms * 1e-3) / 1e12
    out["grouped_bf16_ex_64"] = {"ms": ms, "tflops": tf, "frac_of_peak": tf / BF16_PEAK_TFLOPS, "flop": flops}
    del keep
    return out


def grouped_legs(torch, B, h, dev):
    """The grouped workload: 64 groups of one problem, m, n, k in {64, 80, ..., 2048} (numpy seed 0), alpha 1,
    beta 0.5.  Returns the FP32 aclblasSgemmGroupedBatched call, the BF16 (BF16 C) aclblasGemmGroupedBatchedEx
    call on the same shapes, the flop count and the tensors to keep alive."""
    import numpy as np
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
    f32_call = lambda: B.sgemm_grouped_batched(h.value, [B.OP_N] * 64, [B.OP_N] * 64, g[0], g[1], g[2],
                                               [1.0] * 64, [x.data_ptr() for x in As], g[0],
                                               [x.data_ptr() for x in Bs], g[2], [0.5] * 64,
                                               [x.data_ptr() for x in Cs], g[0], [1] * 64)
    Ah = [torch.randn(int(mm * kk), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    Bh = [torch.randn(int(kk * nn), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    Ch = [torch.randn(int(mm * nn), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
    ptr = lambda ts: torch.tensor([x.data_ptr() for x in ts], dtype=torch.int64, device=dev)
    pa, pb, pc = ptr(Ah), ptr(Bh), ptr(Ch)
    bf16_call = lambda: B.gemm_grouped_batched_ex(h.value, [B.OP_N] * 64, [B.OP_N] * 64, g[0], g[1], g[2],
                                                  [1.0] * 64, pa.data_ptr(), B.ACL_BF16, g[0], pb.data_ptr(),
                                                  B.ACL_BF16, g[2], [0.5] * 64, pc.data_ptr(), B.ACL_BF16, g[0],
                                                  [1] * 64, B.COMPUTE["COMPUTE_32F"])
    return f32_call, bf16_call, flops, (As, Bs, Cs, Ah, Bh, Ch, pa, pb, pc)


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
    import torch
    import aclblas as B
    import workloads as W

    B.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    h = B.Handle()
    stream = torch.cuda.current_stream()
    h.set_stream(stream.cuda_stream)
    timer = Timer(torch, stream)
    if args.calibrate_copy > 0:
        calibrate_copy(torch, dev, args.calibrate_copy)
    if args.profile_mixed:
        wl = MixedWorkload(torch, B, W)
        wl.h = h.value
        for _ in range(3):
            wl.call(args.update)
        torch.cuda.synchronize()
        h.destroy()
        print("mixed profile run done (%s)" % args.update)
        return
    if args.profile_c2:
        res = bench_c2(torch, B, W, h, args.steps, args.warmup)
        h.destroy()
        print(json.dumps(res))
        return
    if args.profile_matinv:
        res = bench_matinv(torch, B, W, h, timer)
        h.destroy()
        print(json.dumps(res))
        return
    if args.profile_fp8:
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
    if args.profile_grouped:
        f32_call, bf16_call, _, keep = grouped_legs(torch, B, h, dev)
        for _ in range(3):
            assert f32_call() == 0
        for _ in range(3):
            assert bf16_call() == 0
        torch.cuda.synchronize()
        del keep
        h.destroy()
        print("grouped profile run done")
        return

    if args.profile_sgemm512:
        m = 512
        a = (torch.rand(m * m, device=dev) * 2 - 1)
        b = (torch.rand(m * m, device=dev) * 2 - 1)
        c = (torch.rand(m * m, device=dev) * 2 - 1)
        for _ in range(200):
            st = B.sgemm(h.value, B.OP_N, B.OP_N, m, m, m, 1.0, a.data_ptr(), m, b.data_ptr(), m, 0.5, c.data_ptr(), m)
            assert st == 0, st
        torch.cuda.synchronize()
        h.destroy()
        print("sgemm 512 profile run done")
        return

    if args.profile_gemm:
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

    wl = MixedWorkload(torch, B, W)
    wl.h = h.value
    log("headline: mixed solve, %s updates, %d warm-up + %d timed calls" % (args.update, args.warmup, args.steps))
    head = wl.run(args.update, args.steps, args.warmup)
    log("headline: %.3f ms per call, fallback %d, max backward error %.2e" %
        (head["ms_per_call"], head["fallback_systems"], head["max_backward_error"]))
    kernels = {"factor": (head["factor_ms"], wl.factor_bytes), "refine": (head["refine_ms"], wl.refine_bytes)}
    dom = max(kernels, key=lambda k: kernels[k][0])
    kms, kbytes = kernels[dom]
    achieved = kbytes / (kms * 1e-3) / 1e9
    name = {"factor": "GesvFactorKernel (%s updates, all systems)" % args.update,
            "refine": "GesvRefineKernel (solve + FP64 refinement, all systems)"}[dom]
    traffic_file = {"factor": "gesv_factor_%s_10000_traffic.md" % args.update.lower(),
                    "refine": "gesv_refine_10000_traffic.md"}[dom]
    out = {
        "metric": METRIC,
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_call"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16",
        "data": "synthetic (A = U diag(s) V^T per system, U and V from the QR of Gaussians, numpy seed 0; "
                "B ~ N(0,1))",
        "config": {"workload": "mixed-size batched solve with refinement: 10000 systems, n 16..256, "
                               "cond 1e2..1e6, nrhs 1..16, backward error <= 1e-6, aclblasSgesvBatchedMixedEx "
                               "(%s trailing updates, FP32 fallback)" % args.update,
                   "systems": MIXED_COUNT, "n": "16..256", "nrhs": "1..16", "cond": "1e2..1e6",
                   "tol": MIXED_TOL, "update": args.update, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(traffic_file),
                     "kernel": name, "kernel_ms": kms, "algorithmic_bytes_per_launch": kbytes,
                     "lu_tflops_in_factor": wl.lu_flop / (head["factor_ms"] * 1e-3) / 1e12},
        "max_backward_error": head["max_backward_error"],
        "headline": head,
    }
    if not args.primary_only:
        comp = {}
        for upd in ("FP32", "BF16", "BF16X3"):
            if upd == args.update:
                continue
            log("comparison: %s updates" % upd)
            comp[upd] = wl.run(upd, 5 if upd == "FP32" else 3, 1)
        if "FP32" in comp:
            comp["speedup_vs_fp32_same_kernels"] = comp["FP32"]["ms_per_call"] / head["ms_per_call"]
        log("comparison: plain public path")
        comp["plain_fp32_public_path"] = wl.plain_public_path(timer)
        comp["speedup_vs_plain_public_path"] = comp["plain_fp32_public_path"]["ms_per_call"] / head["ms_per_call"]
        out["comparison"] = comp
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = wl.cpu_baseline()
    del wl
    if not args.primary_only:
        sec = {}
        log("secondary: 4096 x 32 getrf + getrs")
        sec["lu_4096x32"] = bench_c2(torch, B, W, h, 20, 5)
        log("secondary: GEMM")
        sec["gemm"] = bench_gemm(torch, B, W, h, timer)
        log("secondary: adjoint")
        sec["adjoint"] = bench_adjoint(torch, B, W, h, timer)
        log("secondary: batched inverse")
        sec["matinv"] = bench_matinv(torch, B, W, h, timer)
        out["secondary"] = sec
    h.destroy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark: batched LU solves/s on one MI355X (aclblasSgetrfBatched + aclblasSgetrsBatched).

Workload (one "step"): FP32 LU with partial pivoting and solve of 4096 systems of
order 32 with one right-hand side, A ~ N(0,1) + 4I, b ~ N(0,1), through the C ABI
of libops_blas.so.  Inputs are resident in HBM before the timed region; every
step factors its own fresh copy of the batch (getrf works in place), so no step
re-uses a previous step's output or cache state beyond what the hardware keeps.

Prints one JSON line (see the keys below) with
  value         solves/s over the timed steps (whole job, one GPU)
  roofline      the dominant kernel (getrf) against HBM bandwidth: algorithmic
                bytes per launch / its average duration measured with HIP events
                on the stream it runs on
  cpu_baseline  the CPU restatement of the reference's algorithm (oracle/),
                timed on this host's cores on the same workload
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
# algorithmic bytes per system: getrf reads A and writes LU (n*n*4 each), writes
# n pivots and one info word, reads one 8-byte pointer
GETRF_BYTES = 2 * N * N * 4 + N * 4 + 4 + 8
# getrs reads LU, pivots and b, writes x, reads two pointers
GETRS_BYTES = N * N * 4 + N * 4 + 2 * N * NRHS * 4 + 16


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


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (use --gpus 1)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import aclblas as B
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    B.load()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    A = O.normal((BATCH, N, N), 0, n=N, batch=BATCH, shift=4.0)
    Bv = O.normal((BATCH, NRHS, N), 1)
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
    # correctness of the last step against the FP64 residual oracle
    X = dB[-1].cpu().numpy()
    berr = float(O.backward_error(A, Bv, X).max())
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
        "data": "synthetic (A ~ N(0,1) + 4I, b ~ N(0,1), seeds 0/1)",
        "config": {"workload": "FP32 aclblasSgetrfBatched + aclblasSgetrsBatched, batch 4096 x 32x32, nrhs 1",
                   "batch": BATCH, "n": N, "nrhs": NRHS, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": "getrf (n<=32 wave kernel)", "kernel_ms": getrf_ms,
                     "algorithmic_bytes_per_launch": GETRF_BYTES * BATCH},
        "max_backward_error": berr,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(A, Bv)
    h.destroy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

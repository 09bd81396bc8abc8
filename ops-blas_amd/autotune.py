"""Measures the launch choices of libops_blas.so per size bucket and writes the choice table.

    python ops-blas_amd/autotune.py --table ops-blas_amd/csrc/tuning_table.inc --report tune.md

For every bucket of every tuned operation (see aclblasSetTuningOverride in
include/cann_ops_blas.h) the tuner builds one representative problem, forces
each candidate choice on a handle through the override, times it with HIP
events on the handle's stream (median over repetitions) and keeps the fastest.
Every candidate must produce output bit-identical to the first one: the choice
changes only which workgroup computes which output, never the order in which
an output sums its products, so a difference is a library bug and stops the run.

Buckets and representative problems:
  GEMM, 4 element classes x 5 tile-count buckets x 3 k buckets: square
    m = n giving cus/8, cus/2, 2 cus, 8 cus, 32 cus tiles of 128 x 128, and
    k = 128, 1024, 4096; candidates tile 64/128/256 (FP32: 32/65/64/128, where
    32 and 65 are 32 x 32 and 64 x 64 tiles with deep load prefetch) x L2
    group height 4/8/16.
  grouped GEMM, 2 x 3 buckets: FP32 (SgemmGroupedBatched) and BF16
    (GemmGroupedBatchedEx) problems with m, n, k uniform in steps of 16 (16
    problems up to 512, 40 up to 768, 64 up to 2048); candidates tile 64/128 x
    1..4 workgroups per CU.
  mixed solve, 8 buckets: BF16 or FP32 updates x 128, 512, 2048, 10000 systems, n 16..256, nrhs
    1..16, cond 1e2..1e6; candidates 1..3 factor x 1..4 solve workgroups per CU.
  batched LU (n <= 64), 16 buckets: n = 16, 32, 64 with batches giving 1, 3,
    6 and 12 one-row-per-lane waves per SIMD (n <= 8 has one choice);
    candidates 1 or 2 matrix rows per lane.
After writing the table, rebuild the library (make -C ops-blas_amd).
"""
import argparse
import datetime
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import aclblas as B  # noqa: E402
import workloads as W  # noqa: E402


def _median_ms(torch, stream, fn, min_reps=3, budget_ms=40.0):
    """Median per-call time (HIP events on `stream`) after one untimed call."""
    fn()
    times = []
    t_start = time.perf_counter()
    while len(times) < min_reps or (sum(times) < budget_ms and len(times) < 25):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        e1.synchronize()
        times.append(e0.elapsed_time(e1))
        if time.perf_counter() - t_start > 5.0:
            break
    times.sort()
    return times[len(times) // 2]


class Tuner:
    def __init__(self, torch, quick=False, log=print):
        self.torch = torch
        self.quick = quick
        self.log = log
        self.dev = torch.device("cuda")
        self.stream = torch.cuda.current_stream()
        self.h = B.Handle()
        assert self.h.set_stream(self.stream.cuda_stream) == 0
        self.cus = torch.cuda.get_device_properties(0).multi_processor_count
        self.rows = []      # (op, bucket, a, b)
        self.report = []    # markdown lines

    def close(self):
        for op in range(4):
            B.set_tuning_override(self.h.value, op, -1)
        self.h.destroy()

    def bucket(self, op, cls, units, k=0):
        st, b = B.tuning_bucket(self.h.value, op, cls, units, k)
        assert st == 0, st
        return b

    def _sweep(self, op, bucket, candidates, run, outputs, what):
        """Times every candidate; outputs() returns tensors compared bit for bit."""
        torch = self.torch
        results = []
        ref = None
        for a, b in candidates:
            assert B.set_tuning_override(self.h.value, op, a, b) == 0
            ms = _median_ms(torch, self.stream, run, min_reps=2 if self.quick else 3,
                            budget_ms=15.0 if self.quick else 40.0)
            torch.cuda.synchronize()
            outs = [o.clone() for o in outputs()]
            if ref is None:
                ref = outs
            else:
                for x, y in zip(ref, outs):
                    if not torch.equal(x, y):
                        raise RuntimeError("op %d bucket %d: choice (%d, %d) changed the result (%s)"
                                           % (op, bucket, a, b, what))
            results.append(((a, b), ms))
        B.set_tuning_override(self.h.value, op, -1)
        st, da, db = B.get_tuning_choice(self.h.value, op, bucket)
        best, best_ms = min(results, key=lambda r: r[1])
        self.rows.append((op, bucket, best[0], best[1]))
        cells = ", ".join("%d/%d: %.3f" % (c[0], c[1], t) for c, t in results)
        self.report.append("| %d | %d | %s | %d/%d | %.3f | %s |" % (op, bucket, what, best[0], best[1], best_ms,
                                                                     cells))
        self.log("op %d bucket %2d %-40s best %d/%d %.3f ms" % (op, bucket, what, best[0], best[1], best_ms))

    # ------------------------------------------------------------------ GEMM
    def tune_gemm(self, classes=None):
        torch = self.torch
        targets = [self.cus / 8, self.cus / 2, 2 * self.cus, 8 * self.cus, 32 * self.cus]
        ks = [128, 1024, 4096]
        if classes is None:
            classes = [B.CLASS_F32, B.CLASS_16, B.CLASS_F8, B.CLASS_F8_SCALED]
        groups = [8] if self.quick else [4, 8, 16]
        rng = np.random.default_rng(0)
        for cls in classes:
            # FP32 32 and 65: 32 x 32 and 64 x 64 tiles with deep load prefetch
            tiles = [32, 65, 64, 128] if cls == B.CLASS_F32 else [64, 128, 256]
            for s, tgt in enumerate(targets):
                m = max(64, int(round(128 * math.sqrt(tgt) / 16)) * 16)
                for kb, k in enumerate(ks):
                    bucket = self.bucket(B.TUNE_GEMM, cls, ((m + 127) // 128) ** 2, k)
                    assert bucket == (cls * 5 + s) * 3 + kb, (cls, s, kb, bucket)
                    run, outs, what = self._gemm_problem(cls, m, k, rng)
                    cands = [(t, g) for t in tiles for g in groups]
                    self._sweep(B.TUNE_GEMM, bucket, cands, run, outs, what)
                    del run, outs
                    torch.cuda.empty_cache()

    def _gemm_problem(self, cls, m, k, rng):
        torch = self.torch
        dev = self.dev
        h = self.h.value
        n = m
        if cls == B.CLASS_F32:
            a = torch.rand(m * k, device=dev) * 2 - 1
            b = torch.rand(k * n, device=dev) * 2 - 1
            c = torch.empty(m * n, device=dev)
            run = lambda: B.sgemm(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), m, b.data_ptr(), k, 0.0,
                                  c.data_ptr(), m)
            return run, lambda: [c], "FP32 %dx%dx%d" % (m, n, k)
        if cls == B.CLASS_16:
            a = torch.randn(m * k, device=dev).to(torch.bfloat16)
            b = torch.randn(k * n, device=dev).to(torch.bfloat16)
            c = torch.empty(m * n, device=dev)
            run = lambda: B.gemm_ex(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), B.ACL_BF16, m, b.data_ptr(),
                                    B.ACL_BF16, k, 0.0, c.data_ptr(), B.ACL_FLOAT, m, B.COMPUTE["COMPUTE_32F"])
            return run, lambda: [c], "BF16->FP32 %dx%dx%d" % (m, n, k)
        # FP8 E4M3 codes without NaN (0x7F / 0xFF), small magnitudes
        def codes(count):
            x = rng.integers(0, 0x60, size=count, dtype=np.uint8) | (rng.integers(0, 2, size=count,
                                                                                     dtype=np.uint8) << 7)
            return torch.from_numpy(x).to(dev)
        a = codes(m * k)
        b = codes(k * n)
        c = torch.empty(m * n, device=dev, dtype=torch.float16)
        e4 = B.ACL_FLOAT8_E4M3FN
        if cls == B.CLASS_F8:
            run = lambda: B.gemm_ex(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), e4, m, b.data_ptr(), e4, k, 0.0,
                                    c.data_ptr(), B.ACL_FLOAT16, m, B.COMPUTE["COMPUTE_32F"])
            return run, lambda: [c], "FP8->FP16 %dx%dx%d" % (m, n, k)
        scale = torch.full((2,), 0.5, device=dev)
        run = lambda: B.gemm_ex_scaled(h, B.OP_N, B.OP_N, m, n, k, 1.0, a.data_ptr(), e4, m, scale.data_ptr(),
                                       b.data_ptr(), e4, k, scale.data_ptr() + 4, 0.0, c.data_ptr(),
                                       B.ACL_FLOAT16, m)
        return run, lambda: [c], "FP8 scaled->FP16 %dx%dx%d" % (m, n, k)

    # ----------------------------------------------------------- grouped GEMM
    def tune_grouped(self):
        torch = self.torch
        dev = self.dev
        h = self.h.value
        specs = [(16, 512), (40, 768), (64, 2048)]
        rng = np.random.default_rng(1)
        for want, (count, top) in enumerate(specs):
            mnk = rng.integers(0, (top - 64) // 16 + 1, size=(count, 3)) * 16 + 64
            tiles = int(sum(((mm + 127) // 128) * ((nn + 127) // 128) for mm, nn, _ in mnk))
            bucket = self.bucket(B.TUNE_GROUPED, 0, tiles)
            if bucket != want:
                raise RuntimeError("grouped representative %d has %d tiles: bucket %d" % (want, tiles, bucket))
            As = [torch.rand(int(mm * kk), device=dev) for mm, nn, kk in mnk]
            Bs = [torch.rand(int(kk * nn), device=dev) for mm, nn, kk in mnk]
            Cs = [torch.empty(int(mm * nn), device=dev) for mm, nn, kk in mnk]
            ms_, ns_, ks_ = ([int(x) for x in mnk[:, i]] for i in range(3))
            pa = [x.data_ptr() for x in As]
            pb = [x.data_ptr() for x in Bs]
            pc = [x.data_ptr() for x in Cs]

            def run():
                st = B.sgemm_grouped_batched(h, [B.OP_N] * count, [B.OP_N] * count, ms_, ns_, ks_, [1.0] * count,
                                             pa, ms_, pb, ks_, [0.0] * count, pc, ms_, [1] * count)
                assert st == 0, st

            cands = [(t, w) for t in (64, 128) for w in (1, 2, 3, 4)]
            self._sweep(B.TUNE_GROUPED, bucket, cands, run, lambda: Cs,
                        "%d FP32 problems up to %d (%d tiles)" % (count, top, tiles))
            del As, Bs, Cs

            # the same shapes in BF16 through aclblasGemmGroupedBatchedEx
            bucket = self.bucket(B.TUNE_GROUPED, 1, tiles)
            Ah = [torch.randn(int(mm * kk), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
            Bh = [torch.randn(int(kk * nn), device=dev).to(torch.bfloat16) for mm, nn, kk in mnk]
            Ch = [torch.empty(int(mm * nn), device=dev, dtype=torch.bfloat16) for mm, nn, kk in mnk]
            ptr = lambda ts: torch.tensor([x.data_ptr() for x in ts], dtype=torch.int64, device=dev)
            qa, qb, qc = ptr(Ah), ptr(Bh), ptr(Ch)

            def run_ex():
                st = B.gemm_grouped_batched_ex(h, [B.OP_N] * count, [B.OP_N] * count, ms_, ns_, ks_, [1.0] * count,
                                               qa.data_ptr(), B.ACL_BF16, ms_, qb.data_ptr(), B.ACL_BF16, ks_,
                                               [0.0] * count, qc.data_ptr(), B.ACL_BF16, ms_, [1] * count,
                                               B.COMPUTE["COMPUTE_32F"])
                assert st == 0, st

            self._sweep(B.TUNE_GROUPED, bucket, cands, run_ex, lambda: Ch,
                        "%d BF16 problems up to %d (%d tiles)" % (count, top, tiles))
            del Ah, Bh, Ch

    # ------------------------------------------------------------ mixed solve
    def tune_mixed(self):
        torch = self.torch
        dev = self.dev
        h = self.h.value
        counts = [128, 512, 2048, 2500 if self.quick else 10000]
        # class 0: BF16 trailing updates (the default two planes), class 1: FP32 updates
        for cls, q, count in [(c, q, n) for c in (0, 1) for q, n in enumerate(counts)]:
            systems = W.conditioned_systems(count, seed=100 + q)
            bucket = self.bucket(B.TUNE_MIXED, cls, count)
            if bucket != 4 * cls + q:
                raise RuntimeError("mixed representative %d, class %d: bucket %d" % (q, cls, bucket))
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
            # host arrays built once, as a C caller holds them
            ns = B.host_ints([s["n"] for s in systems])
            rs = B.host_ints([s["nrhs"] for s in systems])
            pa = B.host_ptrs([dA.data_ptr() + 4 * o for o in offA])
            pb = B.host_ptrs([dB.data_ptr() + 4 * o for o in offB])
            px = B.host_ptrs([dX.data_ptr() + 4 * o for o in offB])
            iters = torch.empty(count, dtype=torch.int32, device=dev)

            update = B.MIXED_UPDATE["BF16X2" if cls == 0 else "FP32"]

            def run():
                st = B.sgesv_batched_mixed_ex(h, ns, rs, pa, ns, pb, ns, px, ns, 1e-6, 10, update, iters.data_ptr())
                assert st == 0, st

            cands = [(f, s) for f in (1, 2, 3) for s in (1, 2, 3, 4)]
            self._sweep(B.TUNE_MIXED, bucket, cands, run, lambda: [dX, iters],
                        "%d mixed systems, %s updates" % (count, "BF16x2" if cls == 0 else "FP32"))

    # ------------------------------------------------------------ batched LU
    def tune_lu(self):
        torch = self.torch
        dev = self.dev
        h = self.h.value
        simds = 4 * self.cus
        for S in (16, 32, 64):
            for want, waves in enumerate((simds, 3 * simds, 6 * simds, 12 * simds)):
                batch = waves * 64 // S
                bucket = self.bucket(B.TUNE_LU, S, batch)
                if bucket != (S.bit_length() - 4) * 4 + want:
                    raise RuntimeError("LU representative n=%d batch=%d: bucket %d" % (S, batch, bucket))
                A0, _ = W.lu_batch(batch, S, 1, seed=S + want)
                dA0 = torch.from_numpy(A0).to(dev)
                LU = dA0.clone()
                piv = torch.empty((batch, S), dtype=torch.int32, device=dev)
                info = torch.empty(batch, dtype=torch.int32, device=dev)
                ptrs = B.ptr_array(LU)

                def run():
                    LU.copy_(dA0)
                    assert B.sgetrf_batched(h, S, ptrs.data_ptr(), S, piv.data_ptr(), info.data_ptr(), batch) == 0

                self._sweep(B.TUNE_LU, bucket, [(1, 1), (2, 1)], run, lambda: [LU, piv, info],
                            "LU n=%d batch=%d" % (S, batch))
                del dA0, LU, piv, info, ptrs
                torch.cuda.empty_cache()

    # ------------------------------------------------------------------ output
    def write_table(self, path, keep_from=None):
        """Writes the measured rows plus, from `keep_from`, the rows of buckets not tuned in this run."""
        props = self.torch.cuda.get_device_properties(0)
        dev_name = "%s (%s)" % (props.name, getattr(props, "gcnArchName", "").split(":")[0])
        tuned = {(op, bucket) for op, bucket, _, _ in self.rows}
        if keep_from and os.path.exists(keep_from):
            for op, bucket, a, b in read_table(keep_from):
                if (op, bucket) not in tuned:
                    self.rows.append((op, bucket, a, b))
        with open(path, "w") as f:
            f.write("// Launch-choice table {op, bucket, a, b}: written by ops-blas_amd/autotune.py\n")
            f.write("// from timings on %s, %d CUs, %s.\n" % (dev_name, self.cus,
                                                             datetime.date.today().isoformat()))
            for op, bucket, a, b in sorted(self.rows):
                f.write("{%d, %d, %d, %d},\n" % (op, bucket, a, b))

    def write_report(self, path, seconds):
        props = self.torch.cuda.get_device_properties(0)
        dev_name = "%s %s" % (props.name, getattr(props, "gcnArchName", ""))
        with open(path, "w") as f:
            f.write("# Launch tuning on %s (%d CUs)\n\n" % (dev_name, self.cus))
            f.write("Tuner run time %.0f s.  Times are medians per call in ms; `a/b` is the choice "
                    "(GEMM: tile/L2 group rows; grouped: tile/workgroups per CU; mixed: factor/solve "
                    "workgroups per CU).\n\n" % seconds)
            f.write("| op | bucket | problem | best | ms | all candidates (a/b: ms) |\n|---|---|---|---|---|---|\n")
            for line in self.report:
                f.write(line + "\n")


def read_table(path):
    """(op, bucket, a, b) rows of a table file."""
    rows = []
    for line in open(path):
        line = line.strip()
        if line.startswith("{") and line.endswith("},"):
            rows.append(tuple(int(x) for x in line[1:-2].split(",")))
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--table", default=os.path.join(HERE, "csrc", "tuning_table.inc"))
    ap.add_argument("--report", default=None, help="markdown summary of every measurement")
    ap.add_argument("--ops", default="gemm,grouped,mixed,lu", help="comma list of gemm, grouped, mixed, lu")
    ap.add_argument("--gemm-classes", default="f32,16,f8,f8s",
                    help="GEMM element classes to tune: comma list of f32, 16, f8, f8s")
    ap.add_argument("--quick", action="store_true", help="one L2 group height, fewer repetitions")
    ap.add_argument("--keep-from", default=os.path.join(HERE, "csrc", "tuning_table.inc"),
                    help="table whose rows for buckets not tuned in this run are carried over")
    args = ap.parse_args(argv)
    import torch
    if not torch.cuda.is_available():
        raise SystemExit("autotune needs a GPU")
    B.load()
    t0 = time.perf_counter()
    tuner = Tuner(torch, quick=args.quick, log=lambda s: print(s, flush=True))
    try:
        ops = set(args.ops.split(","))
        if "gemm" in ops:
            names = {"f32": B.CLASS_F32, "16": B.CLASS_16, "f8": B.CLASS_F8, "f8s": B.CLASS_F8_SCALED}
            tuner.tune_gemm([names[c] for c in args.gemm_classes.split(",")])
        if "grouped" in ops:
            tuner.tune_grouped()
        if "mixed" in ops:
            tuner.tune_mixed()
        if "lu" in ops:
            tuner.tune_lu()
    finally:
        tuner.close()
    tuner.write_table(args.table, args.keep_from)
    if args.report:
        tuner.write_report(args.report, time.perf_counter() - t0)
    print("wrote %d rows to %s" % (len(tuner.rows), args.table), flush=True)


if __name__ == "__main__":
    main()

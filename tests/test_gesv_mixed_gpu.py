"""GPU checks of aclblasSgesvBatchedMixed (mixed-order batched solve, BF16 matrix-core LU updates, FP64
refinement).  The reference has no such entry, so there is no bit-level oracle: every system is checked
against FP64 on the host.
  - normwise backward error |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) of the returned FP32 X, recomputed
    in FP64, <= 1e-6 (the stopping tolerance), and the reported one <= tol;
  - forward error against an FP64 solve <= 64 * cond(A) * 2^-24 (relative, max norm);
  - a tight tolerance (1e-12) is reached by refinement;
  - an exactly singular system reports info = k+1 for its zero column and NaN X, without disturbing the others.
Systems: A = U diag(s) V^T with cond(A) log-uniform in [1e2, 1e6], n in [16, 256], nrhs in [1, 16].
"""
import numpy as np
import pytest
import torch

import aclblas as B
import workloads as W

pytestmark = pytest.mark.gpu


def _upload(systems):
    """Packs A, B (and a zero X) of every system into three device buffers; returns pointer lists."""
    dev = torch.device("cuda")
    offA, offB, a_parts, b_parts = [], [], [], []
    oa = ob = 0
    for s in systems:
        offA.append(oa)
        offB.append(ob)
        a_parts.append(s["A"])
        b_parts.append(s["B"])
        oa += s["A"].size + (-s["A"].size) % 4
        ob += s["B"].size + (-s["B"].size) % 4
        a_parts.append(np.zeros((-s["A"].size) % 4, np.float32))
        b_parts.append(np.zeros((-s["B"].size) % 4, np.float32))
    dA = torch.from_numpy(np.concatenate(a_parts)).to(dev)
    dB = torch.from_numpy(np.concatenate(b_parts)).to(dev)
    dX = torch.zeros_like(dB)
    pa = [dA.data_ptr() + 4 * o for o in offA]
    pb = [dB.data_ptr() + 4 * o for o in offB]
    px = [dX.data_ptr() + 4 * o for o in offB]
    return dA, dB, dX, pa, pb, px, offB


def _run(handle, systems, tol, max_iter):
    dA, dB, dX, pa, pb, px, offB = _upload(systems)
    count = len(systems)
    iters = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    berr = torch.full((count,), -1.0, dtype=torch.float32, device="cuda")
    info = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    ns = [s["n"] for s in systems]
    rs = [s["nrhs"] for s in systems]
    st = B.sgesv_batched_mixed(handle.value, ns, rs, pa, ns, pb, ns, px, ns, tol, max_iter, iters.data_ptr(),
                               berr.data_ptr(), info.data_ptr())
    assert st == 0, st
    torch.cuda.synchronize()
    X = dX.cpu().numpy()
    xs = [X[o:o + s["n"] * s["nrhs"]] for o, s in zip(offB, systems)]
    return xs, iters.cpu().numpy(), berr.cpu().numpy(), info.cpu().numpy()


def test_mixed_solve_conditioned_systems(handle):
    systems = W.conditioned_systems(600, seed=3)
    xs, iters, berr, info = _run(handle, systems, 1e-6, 10)
    assert (info == 0).all()
    assert (berr <= 1e-6).all(), berr.max()
    for s, x in zip(systems, xs):
        n, r = s["n"], s["nrhs"]
        be = W.normwise_backward_error(s["A"], n, s["B"], x, r)
        assert be <= 1e-6, (n, r, s["cond"], be)
        A = s["A"].astype(np.float64).reshape(n, n).T
        Bm = s["B"].astype(np.float64).reshape(r, n).T
        ref = np.linalg.solve(A, Bm)
        xm = x.astype(np.float64).reshape(r, n).T
        fe = np.abs(xm - ref).max() / np.abs(ref).max()
        assert fe <= 64 * s["cond"] * 2.0 ** -24, (n, r, s["cond"], fe)


def test_mixed_solve_refines_to_tight_tolerance(handle):
    systems = W.conditioned_systems(64, seed=11)
    xs, iters, berr, info = _run(handle, systems, 1e-12, 10)
    assert (info == 0).all()
    assert (berr <= 1e-12).all(), berr.max()
    assert (iters >= 1).all() and (iters <= 4).all(), iters


def test_mixed_solve_singular_member(handle):
    systems = W.conditioned_systems(9, seed=5, nmin=16, nmax=80)
    bad = 4
    n = systems[bad]["n"]
    A = systems[bad]["A"].copy()
    A[3 * n:4 * n] = 0.0          # column 3 is zero: exact zero pivot at column 3
    systems[bad]["A"] = A
    xs, iters, berr, info = _run(handle, systems, 1e-6, 5)
    assert info[bad] == 4, info
    assert np.isnan(xs[bad]).all()
    others = [i for i in range(len(systems)) if i != bad]
    assert (info[others] == 0).all() and (berr[others] <= 1e-6).all()

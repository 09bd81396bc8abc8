"""GPU checks of aclblasSgesvBatchedMixed / aclblasSgesvBatchedMixedEx (mixed-order batched solve, LU
trailing updates on matrix cores in FP32 or split BF16, FP64 refinement, FP32 fallback).  The reference has no such entry, so there is no bit-level oracle: every system is checked
against FP64 on the host.
  - normwise backward error |b - A x|_inf / (|A|_inf |x|_inf + |b|_inf) of the returned FP32 X, recomputed
    in FP64, <= 1e-6 (the stopping tolerance), and the reported one <= tol;
  - forward error against an FP64 solve <= 64 * cond(A) * 2^-24 (relative, max norm);
  - a tight tolerance (1e-12) is reached by refinement;
  - an exactly singular system reports info = k+1 for its zero column and NaN X, without disturbing the others;
  - the systems the BF16 modes cannot refine are solved again with FP32 updates and flagged.
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


def _run(handle, systems, tol, max_iter, update=None):
    dA, dB, dX, pa, pb, px, offB = _upload(systems)
    count = len(systems)
    iters = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    berr = torch.full((count,), -1.0, dtype=torch.float32, device="cuda")
    info = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    fb = torch.full((count,), -1, dtype=torch.int32, device="cuda")
    ns = [s["n"] for s in systems]
    rs = [s["nrhs"] for s in systems]
    if update is None:
        st = B.sgesv_batched_mixed(handle.value, ns, rs, pa, ns, pb, ns, px, ns, tol, max_iter, iters.data_ptr(),
                                   berr.data_ptr(), info.data_ptr())
    else:
        st = B.sgesv_batched_mixed_ex(handle.value, ns, rs, pa, ns, pb, ns, px, ns, tol, max_iter,
                                      B.MIXED_UPDATE[update], iters.data_ptr(), berr.data_ptr(), info.data_ptr(),
                                      fb.data_ptr())
    assert st == 0, st
    torch.cuda.synchronize()
    X = dX.cpu().numpy()
    xs = [X[o:o + s["n"] * s["nrhs"]] for o, s in zip(offB, systems)]
    return xs, iters.cpu().numpy(), berr.cpu().numpy(), info.cpu().numpy(), fb.cpu().numpy()


def _check_against_fp64(systems, xs, tol):
    for s, x in zip(systems, xs):
        n, r = s["n"], s["nrhs"]
        be = W.normwise_backward_error(s["A"], n, s["B"], x, r)
        assert be <= tol, (n, r, s["cond"], be)
        A = s["A"].astype(np.float64).reshape(n, n).T
        Bm = s["B"].astype(np.float64).reshape(r, n).T
        ref = np.linalg.solve(A, Bm)
        xm = x.astype(np.float64).reshape(r, n).T
        fe = np.abs(xm - ref).max() / np.abs(ref).max()
        assert fe <= 64 * s["cond"] * 2.0 ** -24, (n, r, s["cond"], fe)


@pytest.mark.parametrize("update", ["FP32", "BF16", "BF16X2", "BF16X3"])
def test_mixed_solve_conditioned_systems(handle, update):
    """Every trailing-update precision, 300 systems of the benchmark's distribution: every system ends at
    backward error <= 1e-6 (recomputed on the host in FP64), by refinement or by the FP32 fallback."""
    systems = W.conditioned_systems(300, seed=3)
    xs, iters, berr, info, fb = _run(handle, systems, 1e-6, 10, update)
    assert (info == 0).all()
    assert (berr <= 1e-6).all(), berr.max()
    assert set(np.unique(fb)) <= {0, 1}
    if update == "FP32":
        assert not fb.any()
    if update == "BF16":
        # one BF16 plane carries 8 significant bits: refinement cannot reach 1e-6 once cond(A) is large,
        # so some systems must take the FP32 path
        assert fb.any()
    conds = np.array([s["cond"] for s in systems])
    print("%s: fallback %d of %d (cond of those >= %.3g), mean corrections %.2f" %
          (update, int(fb.sum()), len(systems), conds[fb == 1].min() if fb.any() else 0.0, iters.mean()))
    _check_against_fp64(systems, xs, 1e-6)


def test_mixed_solve_default_entry(handle):
    """aclblasSgesvBatchedMixed is the two-plane mode."""
    systems = W.conditioned_systems(64, seed=4)
    xs, iters, berr, info, _ = _run(handle, systems, 1e-6, 10)
    xs2, iters2, berr2, info2, _ = _run(handle, systems, 1e-6, 10, "BF16X2")
    assert all(np.array_equal(a, b) for a, b in zip(xs, xs2))
    assert np.array_equal(iters, iters2) and np.array_equal(info, info2)
    _check_against_fp64(systems, xs, 1e-6)


@pytest.mark.parametrize("update", ["FP32", "BF16X2"])
def test_mixed_solve_refines_to_tight_tolerance(handle, update):
    """The refinement keeps x in FP64: it reaches 1e-12 (reported, on the FP64 iterate); the returned FP32 X then
    carries only its own rounding (backward error at the FP32 level, recomputed on the host)."""
    systems = W.conditioned_systems(64, seed=11)
    xs, iters, berr, info, fb = _run(handle, systems, 1e-12, 10, update)
    assert (info == 0).all()
    assert (berr <= 1e-12).all(), berr.max()
    assert (iters >= 1).all() and (iters <= 10).all(), iters
    _check_against_fp64(systems, xs, 1e-6)


@pytest.mark.parametrize("update", ["FP32", "BF16", "BF16X2"])
def test_mixed_solve_singular_member(handle, update):
    systems = W.conditioned_systems(9, seed=5, nmin=16, nmax=80)
    bad = 4
    n = systems[bad]["n"]
    A = systems[bad]["A"].copy()
    A[3 * n:4 * n] = 0.0          # column 3 is zero: exact zero pivot at column 3
    systems[bad]["A"] = A
    xs, iters, berr, info, fb = _run(handle, systems, 1e-6, 5, update)
    assert info[bad] == 4, info
    assert np.isnan(xs[bad]).all()
    assert fb[bad] == (0 if update == "FP32" else 1)
    others = [i for i in range(len(systems)) if i != bad]
    assert (info[others] == 0).all() and (berr[others] <= 1e-6).all()


def test_mixed_solve_orders_across_panel_edges(handle):
    """Orders around the 16-row, 32-column and 64-row block edges, one and sixteen right-hand sides."""
    systems = []
    for i, n in enumerate([16, 17, 31, 32, 33, 47, 48, 63, 64, 65, 95, 96, 97, 127, 128, 129, 223, 224, 225, 255,
                           256]):
        for r in (1, 16):
            s = W.conditioned_systems(1, seed=100 + 2 * i + (r > 1), nmin=n, nmax=n, rmin=r, rmax=r,
                                      log_cond=(3.0, 3.0))[0]
            systems.append(s)
    for update in ("FP32", "BF16X2"):
        xs, iters, berr, info, fb = _run(handle, systems, 1e-7, 10, update)
        assert (info == 0).all() and (berr <= 1e-7).all(), (update, berr.max())
        _check_against_fp64(systems, xs, 1e-7)


def test_launch_events_bracket_the_kernels(handle):
    """aclblasSetLaunchEvents: the factor launch is bracketed by events 0 and 1, the solve by 2 and 3."""
    systems = W.conditioned_systems(32, seed=8)
    stream = torch.cuda.current_stream()
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
    for e in evs:
        e.record(stream)
    torch.cuda.synchronize()
    import ctypes
    arr = (ctypes.c_void_p * 8)(*[e.cuda_event for e in evs])
    assert B.set_launch_events(handle.value, arr, 8) == 0
    try:
        _run(handle, systems, 1e-6, 10, "BF16X2")
        torch.cuda.synchronize()
        f = evs[0].elapsed_time(evs[1])
        s = evs[2].elapsed_time(evs[3])
        assert f > 0 and s > 0, (f, s)
        assert evs[0].elapsed_time(evs[3]) >= f + s - 1e-3
    finally:
        assert B.set_launch_events(handle.value, None, 0) == 0


@pytest.mark.parametrize("update", ["FP32", "BF16X2"])
def test_mixed_solve_exact_ties(handle, update):
    """Tie-heavy integer systems A = P L U (L unit lower with one +-1 below the diagonal per row, U integer
    upper, P a row permutation) with an integer solution: the pivot candidates of most columns tie in |a|,
    every multiplier is 0 or +-1 and every intermediate of the factorisation and the solves a small integer
    (exact in FP32, and in two BF16 planes), so the returned X must be the integer solution exactly - a wrong
    tie choice or row map would break it."""
    rng = np.random.default_rng(71)
    systems = []
    for n in (20, 40, 64, 100, 129, 200, 256):
        L = np.eye(n)
        for i in range(1, n):
            L[i, rng.integers(0, i)] = rng.choice([-1.0, 1.0])
        U = np.triu(rng.integers(-4, 5, size=(n, n)).astype(np.float64))
        U[np.arange(n), np.arange(n)] = rng.integers(1, 5, size=n) * rng.choice([-1, 1], size=n)
        A = (L @ U)[rng.permutation(n)]
        r = 3
        X = rng.integers(-3, 4, size=(n, r)).astype(np.float64)
        Bm = A @ X
        systems.append({"n": n, "nrhs": r, "cond": float(np.linalg.cond(A)),
                        "A": np.ascontiguousarray(A.T, dtype=np.float32).reshape(-1),
                        "B": np.ascontiguousarray(Bm.T, dtype=np.float32).reshape(-1),
                        "X": np.ascontiguousarray(X.T, dtype=np.float32).reshape(-1)})
    xs, iters, berr, info, fb = _run(handle, systems, 1e-6, 10, update)
    assert (info == 0).all() and not fb.any()
    for s, x in zip(systems, xs):
        assert np.array_equal(x, s["X"]), (s["n"], np.abs(x - s["X"]).max())

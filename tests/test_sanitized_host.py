"""Host-side AddressSanitizer + UndefinedBehaviorSanitizer run of the C ABI and of the oracle (CPU only).

`make -C ops-blas_amd sanitize` builds libops_blas_asan.so: the same sources with the host code instrumented
(device code as in the product library); `make -C oracle sanitize` builds liboracle_asan.so with the same clang
runtime.  Both builds are incremental here.  The CPU suites that drive the library's host code (every argument
row of the reference's tables on a handle without a device, handle / stream / workspace / logger calls, tuning
buckets and overrides, the symbol table) and the oracle's own tests then run in a child Python with the ASan
runtime preloaded and both sanitizers set to stop at the first report, so any report fails this test.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"
SUITES = ["tests/test_abi.py", "tests/test_tuning.py", "tests/test_oracle.py", "tests/test_adjoint.py",
          "tests/test_gemm_criteria.py"]


def _runtime():
    if not os.path.exists(CLANG):
        return None
    out = subprocess.run([CLANG, "-print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    path = out.stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_host_code_under_asan_and_ubsan(tmp_path):
    rt = _runtime()
    if rt is None:
        pytest.skip("clang ASan runtime not present")
    jobs = str(min(8, os.cpu_count() or 1))
    for d in ("ops-blas_amd", "oracle"):
        subprocess.run(["make", "-s", "-j", jobs, "-C", os.path.join(ROOT, d), "sanitize"], check=True,
                       timeout=1500)
    env = dict(os.environ)
    env.update({
        "LD_PRELOAD": rt,
        "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0:log_path=%s" % (tmp_path / "asan"),
        "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1:log_path=%s" % (tmp_path / "ubsan"),
        "PYTHONMALLOC": "malloc",
        "ACLBLAS_LIBRARY": os.path.join(ROOT, "ops-blas_amd", "_build", "asan", "libops_blas_asan.so"),
        "ORACLE_LIBRARY": os.path.join(ROOT, "oracle", "_build", "liboracle_asan.so"),
    })
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"]
                       + SUITES, cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    reports = sorted(p.name for p in tmp_path.iterdir())
    text = "".join((tmp_path / p).read_text() for p in reports)
    assert r.returncode == 0 and not reports, "sanitized run failed (rc %d)\n%s\n%s" % (
        r.returncode, text[:4000], r.stdout[-3000:] + r.stderr[-2000:])
    assert " passed" in r.stdout, r.stdout[-2000:]

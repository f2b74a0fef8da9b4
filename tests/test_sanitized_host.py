"""The host code of libspx.so under AddressSanitizer and
UndefinedBehaviorSanitizer: the CPU tests of the C ABI, the Matrix Market
loader, the test_spmv driver, the persistent handle and bench's helpers run
in a child process against sparse-blas_amd/_asan/libspx.so (host code
instrumented, device code as usual; build.py builds it) with the clang ASan
runtime preloaded.  Any sanitizer report fails the test."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "sparse-blas_amd"))
import build  # noqa: E402  (sparse-blas_amd/build.py)


def test_host_code_clean_under_asan_and_ubsan(tmp_path):
    if not os.path.exists(build.ASAN_LIB):
        pytest.skip("sanitized build missing: python sparse-blas_amd/build.py --asan")
    rt = build.asan_runtime()
    assert rt, "clang ASan runtime not found"
    log = str(tmp_path / "san")
    env = dict(os.environ, SPX_LIBRARY=build.ASAN_LIB, LD_PRELOAD=rt,
               ASAN_OPTIONS=f"detect_leaks=0:log_path={log}", UBSAN_OPTIONS=f"log_path={log}:print_stacktrace=1")
    files = ["tests/test_capi.py", "tests/test_mmio.py", "tests/test_harness.py", "tests/test_handle.py",
             "tests/test_bench_cpu.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-m", "not gpu", "-p", "no:cacheprovider", *files],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    reports = glob.glob(log + "*")
    text = "".join(open(p).read() for p in reports)
    assert not reports, text[:4000]
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]

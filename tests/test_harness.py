"""The test_spmv driver (the reference's test workflow, SB/test/dspmv_test.cu:
38-471) run as a subprocess: argument errors on any machine; on the GPU the
'g' matrix and Matrix Market files through baseline / v1 / v2 with the
reference's agreement flags and the FP64 host-loop error it prints."""
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "sparse-blas_amd", "test_spmv")


def run(*args, timeout=300):
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout


@pytest.fixture(autouse=True)
def need_exe():
    if not os.path.exists(EXE):
        pytest.skip("test_spmv not built (python sparse-blas_amd/build.py)")


def test_usage_and_gpu_count_errors():
    rc, out = run()
    assert rc == 255 and "Usage" in out
    rc, out = run("g", 200, 0, 1, 1)
    assert rc == 255 and "greater than 0" in out


def table_rows(out):
    rows = [l.split() for l in out.splitlines() if re.match(r"\s+\d+\s+[\d.e+-]+", l)]
    err = float(re.search(r"host loop \(worst of all runs\): (\S+)", out).group(1))
    return rows, err


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_g_matrix_all_kernels(kernel):
    rc, out = run("g", 200, 1, 3, kernel)
    assert rc == 0, out
    assert "m: 200 n: 200 nnz: 4850" in out
    rows, err = table_rows(out)
    assert len(rows) == 3
    for r in rows:
        assert r[3] == "Y" and r[5] == "Y", out
    assert err <= 1e-13


@pytest.mark.gpu
def test_g_matrix_order_not_multiple_of_8_is_reported():
    rc, out = run("g", 100, 1, 1, 1)
    assert rc == 255 and "multiple of 8" in out


@pytest.mark.gpu
def test_kernel_out_of_range():
    rc, out = run("g", 200, 1, 1, 4)
    assert rc == 255 and "1, 2, or 3" in out


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", ["f", "b"])
def test_matrix_market_file(tmp_path, dtype):
    rng = np.random.default_rng(3)
    n, L = 3000, 20000
    r, c = rng.integers(0, n, L), rng.integers(0, n, L)
    v = rng.uniform(-1, 1, L)
    keep = r >= c  # lower triangle of a symmetric matrix
    lines = [f"%%MatrixMarket matrix coordinate real symmetric\n{n} {n} {int(keep.sum())}\n"]
    lines += [f"{a + 1} {b + 1} {float(x)!r}\n" for a, b, x in zip(r[keep], c[keep], v[keep])]
    p = tmp_path / "sym.mtx"
    p.write_text("".join(lines))
    rc, out = run("f", p, 1, 2, 2, dtype)
    assert rc == 0, out
    rows, err = table_rows(out)
    assert len(rows) == 2 and all(x[3] == "Y" and x[5] == "Y" for x in rows)
    assert err <= 1e-13

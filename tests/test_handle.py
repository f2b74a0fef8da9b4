"""The persistent handle of include/anonymouslib_hip.h (the reference's CSR5
handle class, SB/include/anonymouslib_cuda.h:11-327) driven from C++ by
tests/handle_check: its host logic on any machine, and on the GPU the
reference wrapper's call order (SB/src/csr5_kernel.cu:17-30) followed by a
second x through the same conversion, a reconversion with another sigma and
the segmented engine, each result checked against the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "handle_check")


@pytest.fixture(autouse=True)
def need_exe():
    if not os.path.exists(EXE):
        pytest.skip("handle_check not built (python sparse-blas_amd/build.py)")


def run(*args):
    r = subprocess.run([EXE, *map(str, args)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return [json.loads(l) for l in r.stdout.splitlines() if l.startswith("{")]


def test_host_logic():
    out = run("host")
    sig = [o for o in out if o["check"] == "auto_sigma"]
    assert len(sig) == 13
    for o in sig:  # the auto rule on nnz/m, as the oracle restates it
        assert o["sigma"] == oracle.tile_sigma(o["m"], o["nnz"]), o
    # the rule's branch points (anonymouslib_cuda.h:297-313) by hand at m = 10
    by_nnz = {o["nnz"]: o["sigma"] for o in sig if o["m"] == 10}
    assert [by_nnz[z] for z in (40, 41, 50, 320, 330, 2569, 2570)] == [4, 4, 5, 32, 32, 32, 6]
    codes = next(o for o in out if o["check"] == "codes")
    assert codes["explicit_sigma"] == 12
    assert codes["spmv_in_csr"] == -4  # ANONYMOUSLIB_UNSUPPORTED_CSR_SPMV
    assert codes["sigma2"] == -2 and codes["sigma33"] == -2  # UNSUPPORTED_CSR5_OMEGA
    assert codes["format"] == 0
    fl = next(o for o in out if o["check"] == "float")
    assert fl["setX"] == -5 and fl["asCSR5"] == -5  # UNSUPPORTED_VALUE_TYPE


def ragged(m, n, seed, integer):
    """Rows of 0..40 entries, every 7th empty, one row of 3000 entries."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, 41, m)
    lens[::7] = 0
    lens[m // 3] = min(3000, n)
    rp = np.zeros(m + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    if integer:
        val = rng.integers(-8, 9, int(rp[-1])).astype(np.float64)
    else:
        val = rng.standard_normal(int(rp[-1]))
    return rp, col, val


@pytest.mark.gpu
@pytest.mark.parametrize("integer", [True, False])
@pytest.mark.parametrize("m,n,sigma", [(1000, 1200, 16), (4096, 4096, 4), (37, 5000, 32)])
def test_handle_spmv(tmp_path, m, n, sigma, integer):
    rp, col, val = ragged(m, n, 7 + m, integer)
    rng = np.random.default_rng(m)
    if integer:
        x1 = rng.integers(-4, 5, n).astype(np.float64)
        x2 = rng.integers(-4, 5, n).astype(np.float64)
        a1, a2 = 2.0, -0.5
    else:
        x1, x2 = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        a1, a2 = 1.25, -0.75
    d = str(tmp_path)
    rp.astype(np.int32).tofile(os.path.join(d, "rowptr.bin"))
    col.tofile(os.path.join(d, "col.bin"))
    val.tofile(os.path.join(d, "val.bin"))
    x1.tofile(os.path.join(d, "x1.bin"))
    x2.tofile(os.path.join(d, "x2.bin"))
    with open(os.path.join(d, "meta.txt"), "w") as f:
        f.write(f"{m} {n} {len(col)} {a1!r} {a2!r} {sigma}\n")
    print(f"handle m={m} n={n} nnz={len(col)} sigma={sigma} integer={integer}", flush=True)
    (o,) = run("spmv", d)
    assert all(c == 0 for c in o["rc"]), o
    assert o["unchanged"] and o["spmv_after_asCSR"] == -4 and o["destroy"] == 0 and o["format"] == 0
    assert o["sigma_auto"] == oracle.tile_sigma(m, len(col))
    want = {
        "y1.bin": oracle.spmv(rp, col, val, x1, a1, 0.0),
        "y2.bin": oracle.spmv(rp, col, val, x2, a2, 0.0),
        "y3.bin": oracle.spmv(rp, col, val, x1, a1, 0.0),
        "y4.bin": oracle.spmv(rp, col, val, x2, a2, 0.0),
    }
    for name, ref in want.items():
        y = np.fromfile(os.path.join(d, name), np.float64)
        assert y.shape == ref.shape
        if integer:  # small integers: every summation order is exact
            assert np.array_equal(y, ref), name
        else:  # FP64 within 1e-13 normwise of the sequential host loop
            assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref), name

"""GPU parity: libspx.so on the MI355X against the CPU oracle.

Bar (FP64 path): normwise relative error <= 1e-13 against the oracle AND every
row within 2*gamma_k*|alpha|(|A||x|)_i + 2u|beta y0_i| (k = row length,
gamma_k = k u / (1 - k u), u = 2^-53: the oracle and the kernel each err by at
most gamma_k); bit-exact on small-integer inputs, where every partial sum is
an exact integer so any summation order gives the oracle's bits.
Precision form: normwise error <= 1e-7 (stated per matrix).
"""
import numpy as np
import pytest

import oracle
import spx
from spx import workloads as W

pytestmark = pytest.mark.gpu
U = 2.0 ** -53


def run_plan(A, x, alpha=1.0, beta=0.0, y0=None, **kw):
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **kw)
    y = P(x, alpha, beta, y0)
    return y, P


def check_fp64(A, x, y, alpha, beta, y0, tol=1e-13):
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, alpha, beta, y0)
    _, mag = oracle.spmv_dot2(A["rowptr"], A["col"], A["val"], x, alpha, beta, y0)
    k = np.diff(A["rowptr"]).astype(float)
    gam = k * U / (1 - k * U)
    yb = np.abs(beta * y0) if (y0 is not None and beta != 0) else 0.0
    bound = 2 * gam * mag + 2 * U * yb + 1e-300
    err = np.abs(y - ref)
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, f"{bad.size} rows over the bound, first {bad[:5]} err {err[bad[:5]]} bound {bound[bad[:5]]}"
    nrm = np.linalg.norm(ref)
    rel = np.linalg.norm(y - ref) / nrm if nrm > 0 else np.linalg.norm(y)
    assert rel <= tol, rel
    return rel


KERNEL_CASES = [
    dict(kernel=1), dict(kernel=1, lanes=64), dict(kernel=1, lanes=2),
    dict(kernel=2), dict(kernel=2, items=2), dict(kernel=2, items=16),
    dict(kernel=3), dict(kernel=3, sigma=4), dict(kernel=3, sigma=32), dict(kernel=3, sigma=7),
    dict(kernel=4), dict(kernel=4, items=1), dict(kernel=4, items=8),
]
MODE_CASES = [
    dict(mode=spx.MODE_NNZ, parts=1), dict(mode=spx.MODE_NNZ, parts=3), dict(mode=spx.MODE_NNZ, parts=8),
    dict(mode=spx.MODE_ROWS, parts=1), dict(mode=spx.MODE_ROWS, parts=4),
    dict(mode=spx.MODE_QUEUE, nb=5000, queues=1), dict(mode=spx.MODE_QUEUE, nb=777, queues=3),
]


@pytest.fixture(scope="module")
def c1():
    A = W.zipf(8192, 1.5, 512, seed=1)
    x = W.vector(A["n"], 101)
    return A, x


@pytest.mark.parametrize("kc", KERNEL_CASES, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
@pytest.mark.parametrize("mc", MODE_CASES, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_c1_all_kernels_and_modes(c1, kc, mc):
    A, x = c1
    print("case", kc, mc, flush=True)
    y, _ = run_plan(A, x, 1.0, 0.0, None, **kc, **mc)
    check_fp64(A, x, y, 1.0, 0.0, None)
    y0 = W.vector(A["m"], 7)
    y, _ = run_plan(A, x, 0.7, -0.3, y0, **kc, **mc)
    check_fp64(A, x, y, 0.7, -0.3, y0)


@pytest.mark.parametrize("kernel", [1, 2, 3, 4])
@pytest.mark.parametrize("mc", MODE_CASES, ids=lambda d: "-".join(f"{k}{v}" for k, v in d.items()))
def test_small_integers_bit_exact(kernel, mc):
    A = W.small_int(3000, 2000, 12, seed=5)
    x = W.vector(A["n"], 9, W.VAL_SMALLINT)
    y0 = W.vector(A["m"], 10, W.VAL_SMALLINT)
    y, _ = run_plan(A, x, 2.0, -3.0, y0, kernel=kernel, **mc)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 2.0, -3.0, y0)
    assert np.array_equal(y, ref)


def edge_matrices():
    rng = np.random.default_rng(3)
    out = {}

    def mk(lens, n, name):
        lens = np.asarray(lens, np.int64)
        rp = np.zeros(len(lens) + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        nnz = int(rp[-1])
        col = np.concatenate([np.sort(rng.choice(n, int(k), replace=False)) for k in lens]).astype(np.int32) \
            if nnz else np.zeros(0, np.int32)
        val = rng.uniform(-1, 1, nnz)
        out[name] = W.finish(rp, col, val, len(lens), n)

    mk([0, 0, 0], 5, "all_empty")
    mk([3], 3, "one_row")
    mk([1], 1, "one_by_one")
    mk([0, 0, 5, 0, 0, 0, 7, 0, 0], 10, "empty_runs")
    lens = np.zeros(500, np.int64)
    lens[250] = 20000
    mk(lens, 30000, "one_long_row_among_empty")
    lens = rng.integers(0, 3, 5000)
    lens[::97] = 0
    mk(lens, 4000, "short_with_empties")
    mk(np.full(4096, 16), 4096, "aligned_tiles")  # row ends on tile boundaries
    lens = rng.integers(1, 5, 2000)
    lens[[5, 900, 1999]] = [9000, 15000, 12000]
    mk(lens, 20000, "long_rows_at_edges")
    mk(np.full(3000, 1), 3000, "single_nonzero_rows")  # more rows than one unit holds
    mk(np.full(1024, 8), 2048, "eight_per_row")  # rows on 8-aligned boundaries
    mk([7, 2049, 1, 8, 4093, 2], 5000, "chunk_straddling_rows")
    return out


EDGE = edge_matrices()


@pytest.mark.parametrize("name", list(EDGE))
@pytest.mark.parametrize("kernel", [1, 2, 3, 4])
@pytest.mark.parametrize("parts", [1, 2, 5])
def test_edge_cases(name, kernel, parts):
    A = EDGE[name]
    x = W.vector(max(A["n"], 1), 11)[: A["n"]]
    y0 = np.full(A["m"], np.nan)  # beta = 0 must not read y
    y, _ = run_plan(A, x, 1.25, 0.0, y0, kernel=kernel, parts=parts)
    assert np.all(np.isfinite(y))
    check_fp64(A, x, y, 1.25, 0.0, None)
    y0 = W.vector(A["m"], 12)
    y, _ = run_plan(A, x, -0.5, 2.0, y0, kernel=kernel, parts=parts, mode=spx.MODE_QUEUE, nb=997, queues=2)
    check_fp64(A, x, y, -0.5, 2.0, y0)


def test_partitions_bit_exact_against_host_restatement():
    A = W.zipf(5000, 1.3, 900, seed=8)
    for P in (1, 2, 3, 7, 16):
        _, plan = run_plan(A, W.vector(A["n"], 1), kernel=2, parts=P)
        got = plan.export(0).reshape(-1, 6)
        ref = oracle.v1_partition(A["rowptr"], P, rule=1)
        exp = np.stack([ref["start_idx"], ref["end_idx"], ref["start_row"], ref["end_row"],
                        ref["start_flag"], ref["end_flag"]], 1).astype(np.int64)
        assert np.array_equal(got, exp), P
    for nb in (1000, 4321, 99999):
        _, plan = run_plan(A, W.vector(A["n"], 1), kernel=2, mode=spx.MODE_QUEUE, nb=nb)
        got = plan.export(0).reshape(-1, 6)
        ref = oracle.v2_tasks(A["rowptr"], nb, rule=1)
        exp = np.stack([ref["start_idx"], ref["end_idx"], ref["start_row"], ref["end_row"],
                        ref["start_flag"], ref["end_flag"]], 1).astype(np.int64)
        assert np.array_equal(got, exp), nb


def local_csr(A, part_rows, p):
    """Partition p as its own CSR (SB/src/dspmv_mgpu_v1.cu:125-133, split rows clamped)."""
    s, e, r0, r1, sf, ef = part_rows[p]
    P = len(part_rows)
    row_lo = 0 if p == 0 else r0
    if p == P - 1:
        row_hi = A["m"]
    elif ef:
        row_hi = r1 + 1
    else:
        row_hi = part_rows[p + 1][2]
    nz = e - s + 1
    lrp = np.clip(A["rowptr"][row_lo:row_hi + 1] - s, 0, nz)
    return lrp, A["col"][s:e + 1], A["val"][s:e + 1]


@pytest.mark.parametrize("sigma", [0, 4, 9, 32])
@pytest.mark.parametrize("parts", [1, 3])
def test_tile_descriptors_bit_exact(sigma, parts):
    A = EDGE["short_with_empties"] if sigma == 9 else W.zipf(6000, 1.4, 700, seed=4)
    _, plan = run_plan(A, W.vector(A["n"], 1), kernel=3, sigma=sigma, parts=parts)
    info = plan.info()
    sg = info["sigma"]
    if sigma == 0:
        assert sg == oracle.tile_sigma(A["m"], A["nnz"])
    rows = plan.export(0).reshape(-1, 6)
    tp, fl, ct, vt = plan.export(1), plan.export(2), plan.export(3), plan.export(4)
    otp, ofl, oct_, ovt = [], [], [], []
    for p in range(len(rows)):
        lrp, c, v = local_csr(A, rows, p)
        a, b, cc, vv = oracle.tile_convert(lrp, c, v, sg)
        otp.append(a); ofl.append(b); oct_.append(cc); ovt.append(vv)
    assert np.array_equal(tp, np.concatenate(otp))
    assert np.array_equal(fl, np.concatenate(ofl))
    assert np.array_equal(ct, np.concatenate(oct_))
    assert np.array_equal(vt, np.concatenate(ovt))


def test_merge_coordinates_on_the_merge_path():
    A = W.zipf(4000, 1.5, 512, seed=6)
    _, plan = run_plan(A, W.vector(A["n"], 1), kernel=2, items=4)
    co = plan.export(5).reshape(-1, 2)
    rp = A["rowptr"]
    tw = 256 * 4
    L = A["m"] + A["nnz"]
    for u, (i, j) in enumerate(co):
        d = min(u * tw, L)
        assert i + j == d
        # consumed row ends precede nonzero j; unconsumed ones follow nonzero j-1
        assert i == 0 or rp[i] <= j
        assert i == A["m"] or j == 0 or rp[i + 1] > j - 1


def test_reproducible_bits():
    A = W.zipf(8192, 1.2, 2000, seed=2)
    x = W.vector(A["n"], 3)
    for kc in ({"kernel": 2}, {"kernel": 3}, {"kernel": 1, "parts": 4}, {"kernel": 4}):
        y1, _ = run_plan(A, x, **kc)
        y2, _ = run_plan(A, x, **kc)
        assert np.array_equal(y1, y2)


# ---- the reference entry points -----------------------------------------------------

def g_inputs(n):
    A = W.g_matrix(n)
    alpha, beta = oracle.harness_alpha_beta(1, A["nnz"])  # draws right after the values
    return A, np.ones(n), alpha, beta


@pytest.mark.parametrize("n", [200, 1000])
def test_dropin_g_matrix_all_entry_points(n):
    A, x, alpha, beta = g_inputs(n)
    m, nnz = A["m"], A["nnz"]
    y0 = np.zeros(m)
    args = (m, n, nnz, alpha, A["val"], A["rowptr"], A["col"], x, beta, y0)
    rc, y = spx.spMV_mgpu_baseline(*args)
    assert rc == 0
    check_fp64(A, x, y, alpha, beta, y0)
    for k in (1, 2, 3):
        rc, y = spx.spMV_mgpu_v1(*args, ngpu=1, kernel=k)
        assert rc == 0
        check_fp64(A, x, y, alpha, beta, y0)
        # the harness's v2 search: nb = nnz/(d*c), q = c (dspmv_test.cu:314-332)
        for c in (1, 2, 4, 8):
            rc, y = spx.spMV_mgpu_v2(*args, ngpu=1, kernel=k, nb=nnz // c, q=c)
            assert rc == 0
            check_fp64(A, x, y, alpha, beta, y0)


def test_dropin_nonzero_beta_and_v1_streams():
    A = W.zipf(3000, 1.5, 512, seed=11)
    x = W.vector(A["n"], 2)
    y0 = W.vector(A["m"], 3)
    old = spx.set_v1_streams(4)
    try:
        for k in (1, 2, 3):
            rc, y = spx.spMV_mgpu_v1(A["m"], A["n"], A["nnz"], 0.9, A["val"], A["rowptr"], A["col"], x, 0.4, y0,
                                     kernel=k)
            assert rc == 0
            check_fp64(A, x, y, 0.9, 0.4, y0)
    finally:
        spx.set_v1_streams(old)


@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_dropin_v2_streams_tasks_beyond_the_memory_guard(kernel):
    """A matrix whose footprint exceeds 0.8 x free memory: spMV_mgpu_v2
    runs its nb-sized tasks one after another (SB/src/dspmv_mgpu_v2.cu:43,
    108-115) instead of refusing, and the result matches the oracle; v1 and
    the baseline still refuse (-1).  Free memory is set through the seam."""
    A = W.zipf(20000, 1.5, 512, seed=13)
    x = W.vector(A["n"], 2)
    y0 = W.vector(A["m"], 3)
    L = spx.lib()
    try:
        L.spx_debug_set_free_mem(2_000_000)  # the footprint is about 4.6 MB
        args = (A["m"], A["n"], A["nnz"], 0.7, A["val"], A["rowptr"], A["col"], x, -0.3, y0)
        assert spx.spMV_mgpu_v1(*args, kernel=kernel)[0] == -1
        for q in (1, 2):  # nb is capped at 0.8 * 2e6 / 16 / q: about 14 and 28 tasks
            rc, y = spx.spMV_mgpu_v2(*args, ngpu=1, kernel=kernel, nb=A["nnz"], q=q)
            assert rc == 0
            check_fp64(A, x, y, 0.7, -0.3, y0)
    finally:
        L.spx_debug_set_free_mem(-1)


def test_csr5_kernel_device_pointers():
    A = W.zipf(4000, 1.5, 512, seed=12)
    x = W.vector(A["n"], 4)
    y0 = W.vector(A["m"], 5)
    rp32 = A["rowptr"].astype(np.int32)
    bufs = [spx.DeviceBuffer.from_array(a) for a in (A["val"], rp32, A["col"], x, y0)]
    rc = spx.csr5_kernel(A["m"], A["n"], A["nnz"], 1.5, *bufs[:4], -0.5, bufs[4])
    assert rc == 0
    y = bufs[4].download(np.float64, A["m"])
    check_fp64(A, x, y, 1.5, -0.5, y0)


# ---- full-size workloads --------------------------------------------------------------------

@pytest.fixture(scope="module")
def c2():
    A = W.circuit()
    x = W.vector(A["n"], 22)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
    return A, x, ref


def mixed_rows(m, n, wide_frac, seed, n_long=3, long_len=6000):
    """Rows of 0..12 nonzeros: a fraction `wide_frac` at uniform columns over
    [0, n), the rest within +-1024 of the scaled diagonal; plus `n_long` rows
    of `long_len` uniform columns and some empty rows."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 13, m)
    longs = rng.choice(m, n_long, replace=False)
    k[longs] = long_len
    wide = rng.random(m) < wide_frac
    cols = []
    for r in range(m):
        if k[r] == 0:
            cols.append(np.zeros(0, np.int64))
        elif wide[r] or k[r] == long_len:
            cols.append(np.sort(rng.choice(n, k[r], replace=False)))
        else:
            c = int(r * (n / m))
            lo, hi = max(0, c - 1024), min(n, c + 1025)
            cols.append(np.sort(rng.choice(np.arange(lo, hi), min(k[r], hi - lo), replace=False)))
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum([len(c) for c in cols])
    col = np.concatenate(cols).astype(np.int32)
    val = rng.standard_normal(col.size)
    return W.finish(rp, col, val, m, n)


def check_any(A, x, y, alpha, beta, y0, precision):
    if precision == spx.PREC_F64:
        check_fp64(A, x, y, alpha, beta, y0 if beta != 0 else None)
    else:  # FP32 values exactly, x as a (hi, lo) pair: within 1e-12 of the FP32-value loop
        ref = oracle.spmv_f32val(A["rowptr"], A["col"], A["val"].astype(np.float32), x, alpha, beta,
                                 y0 if beta != 0 else None)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


@pytest.mark.parametrize("precision", [spx.PREC_F64, spx.PREC_F32])
@pytest.mark.parametrize("items", [1, 4])
def test_kernel4_split_into_two_concurrent_kernels(items, precision):
    """30 % of the regular rows at uniform columns over 600,000: their units
    cannot stage x and gather over 4.8 MB, so kernel 4 runs them as a second
    kernel beside the staged units (too few for column bands)."""
    A = mixed_rows(200_000, 600_000, 0.3, seed=5)
    x = W.vector(A["n"], 23)
    y0 = W.vector(A["m"], 24)
    print(f"split case items={items} precision={precision} nnz={A['nnz']}", flush=True)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=4, items=items, precision=precision)
    info = P.info()
    assert info["split"] == 1 and info["bands"] == 0 and 0 < info["staged_units"] < info["units"], info
    for alpha, beta in [(1.0, 0.0), (0.5, -2.0)]:
        check_any(A, x, P(x, alpha, beta, y0), alpha, beta, y0, precision)
    # the same plan twice gives the same bits (no order depends on the two kernels' timing)
    assert np.array_equal(P(x, 0.5, -2.0, y0), P(x, 0.5, -2.0, y0))
    # and a banded matrix of the same size does not split
    Ab = W.circuit(m=600_000, total=4_000_000, n_dense=20, window=1024, seed=5)
    assert spx.Plan(Ab["rowptr"], Ab["col"], Ab["val"], n=Ab["n"], kernel=4, items=items).info()["split"] == 0


@pytest.mark.parametrize("precision", [spx.PREC_F64, spx.PREC_F32])
@pytest.mark.parametrize("case", ["circuit-uniform", "mixed-empty-long"])
def test_kernel4_column_bands(case, precision):
    """Regular rows mostly at uniform columns over more than two bands of x:
    kernel 4 cuts them at band edges (262,144 columns), runs each band's
    slices as units and adds a row's slices in band order."""
    if case == "circuit-uniform":
        A = W.circuit(m=600_000, total=4_000_000, n_dense=20, window=0, seed=5)
    else:
        A = mixed_rows(120_000, 1_000_000, 0.9, seed=17)
    x = W.vector(A["n"], 25)
    y0 = W.vector(A["m"], 26)
    print(f"bands case {case} precision={precision} nnz={A['nnz']}", flush=True)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=4, precision=precision)
    info = P.info()
    assert info["bands"] == -(-A["n"] // 262_144) and info["nlong"] > 0, info
    for alpha, beta in [(1.0, 0.0), (0.5, -2.0), (-1.25, 0.75)]:
        check_any(A, x, P(x, alpha, beta, y0), alpha, beta, y0, precision)
    assert np.array_equal(P(x, 0.5, -2.0, y0), P(x, 0.5, -2.0, y0))
    if precision == spx.PREC_F64:
        # multi-RHS on a banded plan: each column as a separate execute
        k = 3
        X = W.vector(A["n"] * k, 27).reshape(k, A["n"])
        xd, yd = spx.DeviceBuffer.from_array(X.ravel()), spx.DeviceBuffer.from_array(np.tile(y0, k))
        P.execute_multi(k, xd, A["n"], yd, A["m"], 0.5, -2.0)
        spx.synchronize()
        Y = yd.download(np.float64, A["m"] * k).reshape(k, A["m"])
        for j in range(k):
            assert np.array_equal(Y[j], P(X[j], 0.5, -2.0, y0)), j


def test_kernel4_bands_small_integers_bit_exact():
    A = mixed_rows(60_000, 800_000, 0.8, seed=29, long_len=5000)
    rng = np.random.default_rng(3)
    A["val"] = rng.integers(-8, 9, A["nnz"]).astype(np.float64)
    x = rng.integers(-8, 9, A["n"]).astype(np.float64)
    y0 = rng.integers(-8, 9, A["m"]).astype(np.float64)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=4)
    assert P.info()["bands"] == 4
    assert np.array_equal(P(x, 2.0, -1.0, y0), oracle.spmv(A["rowptr"], A["col"], A["val"], x, 2.0, -1.0, y0))


@pytest.mark.parametrize("kc", [dict(kernel=2), dict(kernel=3), dict(kernel=2, parts=8),
                                dict(kernel=2, mode=spx.MODE_QUEUE, nb=59_500_000 // 8, queues=8),
                                dict(kernel=4), dict(kernel=4, items=8)])
def test_c2_circuit_full_size(c2, kc):
    A, x, ref = c2
    print("c2 case", kc, flush=True)
    y, _ = run_plan(A, x, **kc)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    assert rel <= 1e-13, rel
    # per-row bound on a sample of rows (the compensated reference is slower)
    k = np.diff(A["rowptr"]).astype(float)
    big = np.argsort(k)[-50:]
    for r in big:
        lo, hi = A["rowptr"][r], A["rowptr"][r + 1]
        prods = A["val"][lo:hi] * x[A["col"][lo:hi]]
        mag = np.abs(prods).sum()
        assert abs(y[r] - ref[r]) <= 2 * k[r] * U / (1 - k[r] * U) * mag


@pytest.mark.parametrize("name", ["rail", "asic"])
@pytest.mark.parametrize("kernel", [1, 2, 3, 4])
def test_precision_form(name, kernel):
    A = W.rail() if name == "rail" else W.asic()
    x = W.vector(A["n"], 31)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
    y, _ = run_plan(A, x, kernel=kernel, precision=spx.PREC_F32)
    rel = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"precision form {name} k{kernel}: normwise rel err {rel:.3e}", flush=True)
    assert rel <= 1e-7
    if name == "rail":
        # values are exact in FP32; x as a (hi, lo) pair keeps ~48 bits
        assert rel <= 1e-13


# ---- long rows (column-blocked slices, x staged in LDS) --------------------------------

@pytest.mark.parametrize("long_row", [16, 300, -1])
@pytest.mark.parametrize("kernel", [1, 2, 3, 4])
@pytest.mark.parametrize("mc", [dict(parts=1), dict(parts=4), dict(mode=spx.MODE_QUEUE, nb=3000, queues=2)],
                         ids=["p1", "p4", "queue"])
def test_long_row_path(long_row, kernel, mc):
    A = W.zipf(9000, 1.1, 3000, seed=21)
    x = W.vector(A["n"], 5)
    y0 = W.vector(A["m"], 6)
    y, P = run_plan(A, x, 0.8, 0.6, y0, kernel=kernel, long_row=long_row, **mc)
    info = P.info()
    lens = np.diff(A["rowptr"])
    if long_row > 0:
        # at most 1024 long rows per plan: the threshold rises to keep the longest
        assert info["nlong"] == min(int((lens >= long_row).sum()), 1024) > 0
        assert int((lens >= info["long_row"]).sum()) >= info["nlong"]
    else:
        assert info["nlong"] == 0
    check_fp64(A, x, y, 0.8, 0.6, y0)


def test_long_rows_unsorted_columns_and_bit_exact_integers():
    A = W.small_int(2000, 9000, 40, seed=3)
    rng = np.random.default_rng(1)
    rp = A["rowptr"]
    col = A["col"].copy()
    val = A["val"].copy()
    for r in range(0, A["m"], 7):  # shuffle some rows: the long-row split must sort them
        idx = rng.permutation(np.arange(rp[r], rp[r + 1]))
        col[rp[r]:rp[r + 1]] = col[idx]
        val[rp[r]:rp[r + 1]] = val[idx]
    B = W.finish(rp, col, val, A["m"], A["n"])
    x = W.vector(B["n"], 9, W.VAL_SMALLINT)
    for kernel in (1, 2, 3, 4):
        y, P = run_plan(B, x, 1.0, 0.0, None, kernel=kernel, long_row=50)
        assert P.info()["nlong"] > 0
        assert np.array_equal(y, oracle.spmv(rp, col, val, x, 1.0, 0.0))


def test_long_row_path_precision_form():
    A = W.zipf(6000, 1.1, 3000, seed=22)
    x = W.vector(A["n"], 7)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
    for kernel in (1, 2, 4):
        y, P = run_plan(A, x, kernel=kernel, long_row=100, precision=spx.PREC_F32)
        assert P.info()["nlong"] > 0
        assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= 1e-7


class _At:
    """A device address inside a DeviceBuffer (what Plan.execute reads)."""

    def __init__(self, buf, offset_bytes):
        self.ptr = buf.ptr + offset_bytes


@pytest.mark.parametrize("precision", [spx.PREC_F64, spx.PREC_F32])
@pytest.mark.parametrize("offset", [0, 1])
def test_kernel4_x_at_either_16_byte_phase(offset, precision):
    """Kernel 4 copies each staged x window into LDS 16 bytes per lane from
    the window's first 16-byte aligned entry, plus single leading and trailing
    entries.  Which of those a window has depends on the address of x: here x
    starts on a 16-byte boundary (offset 0) or 8 bytes past one (offset 1),
    with windows starting at even and odd columns alike, and y must match the
    oracle either way."""
    A = W.circuit(m=200_000, total=2_000_000, n_dense=10, window=700, seed=8)
    x = W.vector(A["n"], 31)
    y0 = W.vector(A["m"], 32)
    print(f"x phase case offset={offset} precision={precision} nnz={A['nnz']}", flush=True)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=4, items=1, precision=precision)
    assert P.info()["staged_units"] > 0
    xb = spx.DeviceBuffer((A["n"] + 2) * 8)
    xb.upload(np.concatenate([np.full(offset, np.nan), x, np.full(2 - offset, np.nan)]))
    yd = spx.DeviceBuffer.from_array(y0)
    P.execute(_At(xb, 8 * offset), yd, 0.5, -2.0)
    y = yd.download(np.float64, A["m"])
    if precision == spx.PREC_F64:
        check_fp64(A, x, y, 0.5, -2.0, y0)
    else:
        ref = oracle.spmv_f32val(A["rowptr"], A["col"], A["val"].astype(np.float32), x, 0.5, -2.0, y0)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_plan_timing_with_and_without_events():
    """spx_plan_time_each records HIP events around every execute and its SpMV
    kernel, or, with both per-execute arrays NULL, none (wall clock only);
    one array without the other is an argument error.  The executes still
    produce y either way."""
    A = W.circuit(m=100_000, total=1_000_000, n_dense=4, window=500, seed=9)
    x = W.vector(A["n"], 41)
    print(f"timing case nnz={A['nnz']}", flush=True)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], kernel=4, items=2)
    xd, yd = spx.DeviceBuffer.from_array(x), spx.DeviceBuffer(A["m"] * 8)
    w = P.time_wall(xd, yd, 1.0, 0.0, warmup=1, reps=5)
    assert np.isfinite(w) and w > 0
    check_fp64(A, x, yd.download(np.float64, A["m"]), 1.0, 0.0, None)
    e, k, w2 = P.time_each(xd, yd, 1.0, 0.0, warmup=1, reps=5)
    assert np.all(e > 0) and np.all(k > 0) and np.all(k <= e) and w2 > 0
    only = np.zeros(5)
    rc = spx.lib().spx_plan_time_each(P.h, 1.0, xd.ptr, 0.0, yd.ptr, 0, 5, 0, only.ctypes.data, None, None)
    assert rc == spx.EARG


def _kernel1_shapes():
    """Rows around kernel 1's unit capacity (2,048 nonzeros), a run of empty
    rows longer than its row cap (1,024), and short rows between them."""
    rng = np.random.default_rng(17)
    lens = np.concatenate([[3000], np.zeros(1500, np.int64), rng.integers(1, 5, 3000), [2500, 2049, 2048, 2047],
                           np.zeros(100, np.int64), np.full(500, 7), [4000], np.zeros(1030, np.int64), [1]])
    n = 9000
    rp = np.zeros(len(lens) + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, int(k), replace=False)) for k in lens if k > 0]).astype(np.int32)
    val = rng.uniform(-1, 1, int(rp[-1]))
    return W.finish(rp, col, val, len(lens), n)


K1_SHAPES = _kernel1_shapes()


@pytest.mark.parametrize("precision", [spx.PREC_F64, spx.PREC_F32])
@pytest.mark.parametrize("lanes", [0, 2, 64])
@pytest.mark.parametrize("mc", [dict(parts=1), dict(parts=3), dict(mode=spx.MODE_QUEUE, nb=2600, queues=2)],
                         ids=["p1", "p3", "queue"])
def test_kernel1_unit_shapes(mc, lanes, precision):
    """Kernel 1 cuts a partition's rows into units of at most 2,048 nonzeros
    and at most 8 passes of its row groups (1,024 rows): a longer row is a
    unit alone, summed by the whole workgroup; a run of empty rows fills
    units up to the row cap; partition bounds cut rows of either kind."""
    A = K1_SHAPES
    x = W.vector(A["n"], 71)
    y0 = W.vector(A["m"], 72)
    print(f"kernel-1 shapes lanes={lanes} precision={precision} {mc}", flush=True)
    y, P = run_plan(A, x, 0.75, -1.5, y0, kernel=1, lanes=lanes, precision=precision, **mc)
    assert P.info()["nlong"] == 0
    if precision == spx.PREC_F64:
        check_fp64(A, x, y, 0.75, -1.5, y0)
    else:
        ref = oracle.spmv_f32val(A["rowptr"], A["col"], A["val"].astype(np.float32), x, 0.75, -1.5, y0)
        assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def _random_matrix(seed):
    """A seeded random CSR matrix: one of four row-length laws (short uniform,
    heavy-tailed, mostly empty, a few rows far beyond the rest), columns
    distinct and sorted per row, values U[-1, 1)."""
    rng = np.random.default_rng(1000 + seed)
    m = int(rng.integers(1, 20000))
    n = int(rng.integers(1, 30000))
    law = seed % 4
    if law == 0:
        lens = rng.integers(0, 9, m)
    elif law == 1:
        lens = np.minimum(rng.zipf(1.4, m), 6000)
    elif law == 2:
        lens = (rng.random(m) < 0.05) * rng.integers(1, 40, m)
    else:
        lens = rng.integers(0, 4, m)
        big = rng.choice(m, min(m, int(rng.integers(1, 6))), replace=False)
        lens[big] = rng.integers(2000, 9000, len(big))
    lens = np.minimum(lens, n).astype(np.int64)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, int(k), replace=False)) for k in lens if k > 0]).astype(np.int32) \
        if rp[-1] else np.zeros(0, np.int32)
    return W.finish(rp, col, rng.uniform(-1, 1, int(rp[-1])), m, n)


@pytest.mark.parametrize("seed", range(24))
def test_random_matrices_all_kernels(seed):
    """Every kernel, at its default schedule and at a seeded random one
    (partition count or task size), on seeded random matrices against the
    oracle's per-row bound."""
    A = _random_matrix(seed)
    rng = np.random.default_rng(seed)
    x = W.vector(max(A["n"], 1), 200 + seed)[: A["n"]]
    y0 = W.vector(A["m"], 300 + seed)
    print(f"random matrix seed={seed} m={A['m']} n={A['n']} nnz={A['nnz']}", flush=True)
    for kernel in (1, 2, 3, 4):
        y, _ = run_plan(A, x, 1.5, 0.0, None, kernel=kernel)
        check_fp64(A, x, y, 1.5, 0.0, None)
        if kernel == 4:
            continue
        if seed % 2:
            kw = dict(mode=spx.MODE_QUEUE, nb=int(rng.integers(500, 20000)), queues=int(rng.integers(1, 4)))
        else:
            kw = dict(mode=spx.MODE_NNZ, parts=int(rng.integers(2, 9)))
        y, _ = run_plan(A, x, -0.5, 0.25, y0, kernel=kernel, **kw)
        check_fp64(A, x, y, -0.5, 0.25, y0)

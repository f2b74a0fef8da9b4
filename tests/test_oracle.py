"""Pins the CPU oracle (oracle/oracle.c) against what the mathematics and the
reference's own code fix: dense brute force, exact integer arithmetic, closed
forms, hand-traced tables and brute-force recomputation.  Each check is
chosen so that a plausible slip in the oracle (dropped beta term, swapped
operand, off-by-one row search, wrong flag bit) fails it."""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

rng = np.random.default_rng(12345)


def rand_csr(m, n, density, rng, empty_frac=0.0, ints=False):
    A = (rng.random((m, n)) < density)
    if empty_frac:
        A[rng.random(m) < empty_frac] = False
    vals = rng.integers(-9, 10, size=(m, n)).astype(float) if ints else rng.uniform(-1, 1, (m, n))
    D = np.where(A, vals, 0.0)
    if ints:
        D[A & (D == 0)] = 1.0
    rowptr = np.zeros(m + 1, np.int64)
    rowptr[1:] = np.cumsum(A.sum(1))
    rows, cols = np.nonzero(A)
    return rowptr, cols.astype(np.int32), D[rows, cols], D


@pytest.mark.parametrize("m,n,density", [(1, 1, 1.0), (7, 5, 0.5), (33, 64, 0.2), (64, 17, 0.9)])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.75, 0.0), (-1.5, 0.5), (2.0, -1.0)])
def test_spmv_matches_dense_bruteforce(m, n, density, alpha, beta):
    rowptr, col, val, D = rand_csr(m, n, density, rng, empty_frac=0.2)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    y = oracle.spmv(rowptr, col, val, x, alpha, beta, y0)
    ref = alpha * (D @ x) + beta * y0
    mag = abs(alpha) * (np.abs(D) @ np.abs(x)) + abs(beta * y0)
    assert np.all(np.abs(y - ref) <= 4 * (n + 2) * 2.2e-16 * mag + 1e-300)


def test_spmv_exact_on_small_integers():
    # main.cu:324-333 recipe: every product and partial sum is an exact integer
    rowptr, col, val, D = rand_csr(40, 30, 0.4, rng, ints=True)
    x = rng.integers(0, 10, 30).astype(float)
    y0 = rng.integers(-5, 5, 40).astype(float)
    y = oracle.spmv(rowptr, col, val, x, 3.0, -2.0, y0)
    exact = [3 * sum(int(D[i, j]) * int(x[j]) for j in range(30)) - 2 * int(y0[i]) for i in range(40)]
    assert np.array_equal(y, np.array(exact, dtype=float))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_openmp_rows_equal_the_sequential_loop(threads):
    # the CPU baseline's parallel form sums every row exactly as the
    # sequential loop does, so the two agree bit for bit
    rowptr, col, val, D = rand_csr(301, 120, 0.3, rng, empty_frac=0.1)
    x = rng.uniform(-1, 1, 120)
    y0 = rng.uniform(-1, 1, 301)
    for alpha, beta in [(1.0, 0.0), (-0.5, 1.5)]:
        ref = oracle.spmv(rowptr, col, val, x, alpha, beta, y0)
        y = oracle.spmv_omp(rowptr, col, val, x, alpha, beta, y0, threads=threads)
        assert np.array_equal(y, ref)
    assert np.allclose(oracle.spmv_omp(rowptr, col, val, x, threads=threads), D @ x, rtol=1e-13, atol=1e-14)


def test_beta_zero_never_reads_y():
    rowptr, col, val, D = rand_csr(10, 10, 0.5, rng)
    x = rng.uniform(-1, 1, 10)
    y = oracle.spmv(rowptr, col, val, x, 1.0, 0.0, np.full(10, np.nan))
    assert np.all(np.isfinite(y))
    assert np.allclose(y, D @ x, rtol=1e-14, atol=1e-15)


def test_identity_and_diagonal():
    n = 50
    d = rng.uniform(-2, 2, n)
    rowptr = np.arange(n + 1, dtype=np.int64)
    col = np.arange(n, dtype=np.int32)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, n)
    y = oracle.spmv(rowptr, col, d, x, 1.0, 0.5, y0)
    assert np.array_equal(y, d * x * 1.0 + 0.5 * y0)


def test_g_closed_form_row_sums():
    # x = 1 (dspmv_test.cu:270): y_i = alpha * sum_j v_ij + beta*y0_i
    r, c, v = oracle.gen_g(64)
    rowptr = oracle.coo_to_csr_rowptr(64, r)
    x = np.ones(64)
    alpha, beta = 0.5, 0.25
    y0 = rng.uniform(-1, 1, 64)
    y = oracle.spmv(rowptr, c, v, x, alpha, beta, y0)
    for i in range(64):
        s = math.fsum(v[rowptr[i]:rowptr[i + 1]])
        k = rowptr[i + 1] - rowptr[i]
        assert abs(y[i] - (alpha * s + beta * y0[i])) <= (k + 3) * 2.3e-16 * (abs(alpha) * k + abs(beta * y0[i]))


def test_dot2_is_correctly_rounded_on_hard_sums():
    # cancellation-heavy rows: Dot2 must be within an ulp of the exact rational sum
    for trial in range(20):
        k = 50
        a = rng.uniform(-1, 1, k) * 10.0 ** rng.integers(-8, 8, k)
        b = rng.uniform(-1, 1, k)
        rowptr = np.array([0, k], np.int64)
        col = np.arange(k, dtype=np.int32)
        y, _ = oracle.spmv_dot2(rowptr, col, a, b, 1.0, 0.0)
        exact = sum(Fraction(float(p)) * Fraction(float(q)) for p, q in zip(a, b))
        assert abs(Fraction(y[0]) - exact) <= abs(exact) * Fraction(2, 2 ** 52) + Fraction(1, 2 ** 1000)


def test_plain_oracle_error_within_classical_bound():
    rowptr, col, val, D = rand_csr(200, 300, 0.3, rng)
    x = rng.uniform(-1, 1, 300)
    y = oracle.spmv(rowptr, col, val, x, 1.0, 0.0)
    yx, mag = oracle.spmv_dot2(rowptr, col, val, x, 1.0, 0.0)
    k = np.diff(rowptr)
    gamma = k * 1.11e-16 / (1 - k * 1.11e-16)
    assert np.all(np.abs(y - yx) <= gamma * mag * 1.0001 + 1e-300)


# ---- row search and partitions ----------------------------------------------------------

def test_get_row_from_index_hand_traced():
    # spmv_helper.cu:16-39 on rowptr [0,5,5,10]: idx 5 probes mid=1 first and
    # returns the empty row 1; the row holding nonzero 5 is row 2.
    rp = np.array([0, 5, 5, 10], np.int64)
    assert oracle.get_row_from_index(rp, 5) == 1
    assert oracle.row_of(rp, 5) == 2
    assert [oracle.get_row_from_index(rp, i) for i in (0, 4, 6, 9)] == [0, 0, 2, 2]


def test_row_search_bruteforce():
    for trial in range(30):
        m = int(rng.integers(1, 40))
        lens = rng.integers(0, 5, m)
        rp = np.zeros(m + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        for k in range(int(rp[-1])):
            brute = max(i for i in range(m + 1) if rp[i] <= k)
            assert oracle.row_of(rp, k) == brute
            if np.all(lens > 0):
                assert oracle.get_row_from_index(rp, k) == brute


@pytest.mark.parametrize("P", [1, 2, 3, 5, 8, 13])
def test_v1_partition_closed_form_and_invariants(P):
    m = 300
    lens = rng.integers(0, 30, m)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    p = oracle.v1_partition(rp, P)
    s, e = p["start_idx"], p["end_idx"]
    assert list(s) == [i * nnz // P for i in range(P)]
    assert list(e) == [(i + 1) * nnz // P - 1 for i in range(P)]
    assert s[0] == 0 and e[-1] == nnz - 1 and all(s[1:] == e[:-1] + 1)
    sizes = e - s + 1
    assert sizes.max() - sizes.min() <= 1 and sizes.sum() == nnz
    for i in range(P):
        r0, r1 = p["start_row"][i], p["end_row"][i]
        assert rp[r0] <= s[i] < rp[r0 + 1] and rp[r1] <= e[i] < rp[r1 + 1]
        assert p["start_flag"][i] == (s[i] not in set(rp.tolist()))
        assert p["end_flag"][i] == (e[i] + 1 not in set(rp.tolist()))
    assert np.all(np.diff(p["start_row"]) >= 0)


@pytest.mark.parametrize("nb", [1, 7, 100, 999, 5000])
def test_v2_tasks_closed_form(nb):
    m = 200
    lens = rng.integers(1, 20, m)
    rp = np.zeros(m + 1, np.int64)
    rp[1:] = np.cumsum(lens)
    nnz = int(rp[-1])
    t = oracle.v2_tasks(rp, nb)
    T = -(-nnz // nb)
    assert len(t["start_idx"]) == T
    assert list(t["start_idx"]) == [i * nnz // T for i in range(T)]
    assert list(t["end_idx"]) == [(i + 1) * nnz // T - 1 for i in range(T)]


@pytest.mark.parametrize("P", [1, 2, 4, 7])
def test_v1_split_and_merge_equals_unsplit(P):
    rowptr, col, val, D = rand_csr(60, 40, 0.5, rng, ints=True)
    x = rng.integers(0, 10, 40).astype(float)
    y0 = rng.integers(-5, 5, 60).astype(float)
    full = oracle.spmv(rowptr, col, val, x, 2.0, 3.0, y0)
    split = oracle.v1_spmv(rowptr, col, val, x, 2.0, 3.0, y0, P)
    assert np.array_equal(full, split)  # exact integers: the fix-up algebra is exact


# ---- CSR5-style tiles ---------------------------------------------------------------------

@pytest.mark.parametrize("q,sigma", [(0, 4), (3, 4), (4, 4), (5, 5), (17, 17), (32, 32), (33, 32), (256, 32), (257, 6), (1000, 6)])
def test_tile_sigma_table(q, sigma):
    # anonymouslib_cuda.h:297-313
    assert oracle.tile_sigma(100, 100 * q) == sigma


def brute_tiles(rp, nnz, sigma):
    w = 64 * sigma
    p = -(-nnz // w)
    tp = []
    for t in range(p + 1):
        b = min(t * w, nnz)
        tp.append(max(i for i in range(len(rp)) if rp[i] <= b))
    tpf = list(tp)
    for t in range(p):
        if tp[t] != tp[t + 1] and any(rp[r] == rp[r + 1] for r in range(tp[t], tp[t + 1])):
            tpf[t] = tp[t] | 0x80000000
    starts = {int(rp[r]) for r in range(len(rp) - 1) if rp[r] < rp[r + 1]}
    flags = np.zeros(p * 64, np.uint32)
    for pos in starts:
        flags[pos // sigma] |= np.uint32(1 << (pos % sigma))
    return np.array(tpf, np.uint32), flags


@pytest.mark.parametrize("sigma", [4, 5, 16, 32])
def test_tile_conversion_bruteforce_and_roundtrip(sigma):
    for trial in range(6):
        m = int(rng.integers(1, 400))
        lens = rng.integers(0, 12, m)
        if trial == 0:
            lens[:] = 0
            lens[m // 2] = 3000  # one long row spanning many tiles
        rp = np.zeros(m + 1, np.int64)
        rp[1:] = np.cumsum(lens)
        nnz = int(rp[-1])
        if nnz == 0:
            continue
        col = rng.integers(0, 1000, nnz).astype(np.int32)
        val = rng.uniform(-1, 1, nnz)
        tp, fl, ct, vt = oracle.tile_convert(rp, col, val, sigma)
        btp, bfl = brute_tiles(rp, nnz, sigma)
        assert np.array_equal(tp, btp)
        assert np.array_equal(fl, bfl)
        # decoding the flags gives back every nonempty row start
        pos = [w * sigma + b for w in range(len(fl)) for b in range(32) if (int(fl[w]) >> b) & 1]
        assert pos == sorted(int(rp[r]) for r in range(m) if rp[r] < rp[r + 1])
        # transpose round trip, zero padding past nnz
        p = len(tp) - 1
        back_c = ct.reshape(p, sigma, 64).transpose(0, 2, 1).reshape(-1)
        back_v = vt.reshape(p, sigma, 64).transpose(0, 2, 1).reshape(-1)
        assert np.array_equal(back_c[:nnz], col) and np.array_equal(back_v[:nnz], val)
        assert not back_c[nnz:].any() and not back_v[nnz:].any()


def test_tile_hand_built_example():
    # sigma = 4, one tile of 256 slots; rows of lengths 3, 0, 5, 1 -> starts at 0, 3, 8
    rp = np.array([0, 3, 3, 8, 9], np.int64)
    col = np.arange(9, dtype=np.int32)
    tp, fl, ct, vt = oracle.tile_convert(rp, col, col.astype(float), 4)
    # tile_ptr[p] = upper_bound(rp, nnz) - 1 = m; row 1 is empty -> bit 31 on tile 0
    assert list(tp) == [0 | 0x80000000, 4]
    assert fl[0] == 0b1001 and fl[1] == 0 and fl[2] == 0b0001 and not fl[3:].any()
    assert ct[0] == 0 and ct[64] == 1 and ct[1] == 4 and ct[2] == 8


# ---- generator, harness scalars, metrics --------------------------------------------

def test_g_generator_nnz_closed_form():
    # nb*ceil(0.9 n) + (m - nb)*ceil(0.01 n)
    for n, expect in [(200, 4850), (16, 2 * 15 + 14 * 1), (800, 100 * 720 + 700 * 8)]:
        r, c, v = oracle.gen_g(n)
        assert len(r) == expect
    r, c, v = oracle.gen_g(200)
    assert np.all(np.diff(r) >= 0) and c.max() == 179
    # first glibc rand() values after srand(1)
    assert v[0] == 1804289383 / 2147483647 and v[1] == 846930886 / 2147483647


def test_harness_alpha_beta_follow_the_values():
    # dspmv_test.cu:281-282: alpha, beta are the draws right after the g values
    a, b = oracle.harness_alpha_beta(1, 0)
    assert a == 1804289383 / 2147483647 and b == 846930886 / 2147483647


def test_metrics():
    assert oracle.getB(10, 100) == (10 + 1 + 100) * 4 + (200 + 10) * 8
    assert oracle.getFLOP(100) == 200


# ---- the precision form's checker: orc_spmv_f32val ------------------------------------


def test_f32val_matches_dense_bruteforce_on_float32_values():
    # the values the checker sees are the FP32 ones, widened exactly to FP64
    m, n = 37, 29
    rowptr, col, val, D = rand_csr(m, n, 0.3, rng, empty_frac=0.15)
    v32 = val.astype(np.float32)
    D32 = np.zeros_like(D)
    r = np.repeat(np.arange(m), np.diff(rowptr))
    D32[r, col] = v32.astype(np.float64)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    for alpha, beta in [(1.0, 0.0), (-0.75, 1.25)]:
        y = oracle.spmv_f32val(rowptr, col, v32, x, alpha, beta, y0)
        ref = alpha * (D32 @ x) + beta * y0
        mag = abs(alpha) * (np.abs(D32) @ np.abs(x)) + abs(beta * y0)
        assert np.all(np.abs(y - ref) <= 4 * (n + 2) * 2.2e-16 * mag + 1e-300)
        # and it is not the FP64-valued product: FP32 rounding of the values shows
        assert np.abs(y - (alpha * (D @ x) + beta * y0)).max() > 1e-9


def test_f32val_exact_fraction_sums_keep_x_in_fp64():
    # FP32-representable values (small dyadic rationals) times x entries that
    # need all 53 bits: every product and partial sum is exact in FP64, so the
    # checker must reproduce the rational result exactly.  Rounding x (or the
    # products) to FP32 anywhere would lose the 2^-40 terms.
    m, n = 12, 9
    rr = np.random.default_rng(7)
    rowptr, col, _, A = rand_csr(m, n, 0.5, rr)
    val = rr.integers(-16, 17, len(col)) / 8.0
    val[val == 0] = 0.125
    x = 1.0 + rr.integers(1, 8, n) * 2.0 ** -40
    y0 = rr.integers(-4, 5, m) / 4.0
    y = oracle.spmv_f32val(rowptr, col, val.astype(np.float32), x, 2.0, -0.5, y0)
    for i in range(m):
        s = sum(Fraction(float(val[j])) * Fraction(float(x[col[j]])) for j in range(rowptr[i], rowptr[i + 1]))
        exact = 2 * s - Fraction(1, 2) * Fraction(float(y0[i]))
        assert Fraction(float(y[i])) == exact, (i, y[i], float(exact))


def test_f32val_beta_zero_ignores_nan_y():
    rowptr = np.array([0, 2, 2, 3], np.int64)
    col = np.array([0, 2, 1], np.int32)
    v32 = np.array([0.5, -2.0, 3.0], np.float32)
    x = np.array([2.0, 1.0, 0.25])
    y = oracle.spmv_f32val(rowptr, col, v32, x, 1.0, 0.0, np.full(3, np.nan))
    assert y.tolist() == [0.5, 0.0, 3.0]

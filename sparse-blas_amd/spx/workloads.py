"""Synthetic CSR workloads, generated by libspxgen.so (deterministic, seeded).

Every generator returns a dict with rowptr (int64, m+1), col (int32), val
(float64), m, n, nnz and the realized row-length statistics, which are printed
with every result because several stated statistics of the workloads conflict
with their laws (noted per generator).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import gen

ZIPF, CIRCUIT, ASIC, LOGNORMAL, BANDED, BLOCKDIAG = 1, 2, 3, 4, 5, 6
VAL_UNIFORM, VAL_NORMAL, VAL_ONES, VAL_SIGNED_LOGNORMAL, VAL_SMALLINT = 0, 1, 2, 3, 4


def _params(vals, value_law=None):
    p = np.zeros(16, dtype=np.float64)
    p[: len(vals)] = vals
    p[15] = -1 if value_law is None else value_law
    return p


def build(kind, m, n, seed, prm, value_law=None):
    G = gen()
    p = _params(prm, value_law)
    lens = np.zeros(max(m, 1), dtype=np.int32)
    tot = G.spx_gen_rowlens(kind, m, n, seed, p.ctypes.data, 16, lens.ctypes.data)
    if tot < 0:
        raise ValueError("unknown generator kind")
    rowptr = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens[:m], out=rowptr[1:])
    nnz = int(rowptr[-1])
    col = np.zeros(max(nnz, 1), dtype=np.int32)
    val = np.zeros(max(nnz, 1), dtype=np.float64)
    G.spx_gen_fill(kind, m, n, seed, p.ctypes.data, 16, rowptr.ctypes.data, col.ctypes.data, val.ctypes.data)
    return finish(rowptr, col[:nnz], val[:nnz], m, n)


def finish(rowptr, col, val, m, n):
    lens = np.diff(rowptr)
    return dict(rowptr=rowptr, col=col, val=val, m=m, n=n, nnz=int(rowptr[-1]),
                mean_row=float(lens.mean()) if m else 0.0, max_row=int(lens.max()) if m else 0,
                empty_rows=int((lens == 0).sum()))


def vector(n, seed, law=VAL_UNIFORM):
    x = np.zeros(max(n, 1), dtype=np.float64)
    gen().spx_gen_vector(n, seed, law, x.ctypes.data)
    return x[:n]


def zipf(n=8192, s=1.5, kmax=512, seed=1, value_law=None):
    """n-by-n, k_i ~ Zipf(s) truncated to [1, kmax], uniform distinct columns,
    values U[-1,1).  The stated "about 3 nnz/row" conflicts with Zipf(1.5) on
    [1, 512], whose mean is about 17.4; the law is kept."""
    return build(ZIPF, n, n, seed, [s, 1, kmax], value_law)


def circuit(m=5_558_326, seed=2, total=59_500_000, n_dense=200, window=1024, value_law=None):
    """Circuit-like skewed matrix.  Regular rows: k ~ Geometric(0.1) truncated
    to [1, 12] (realized mean ~5.3, not the stated 10), columns distinct in a
    +-window band around the diagonal (window=0: uniform over [0, n)).
    n_dense rows at seeded positions hold log-uniform [1e4, 1.3e6] nonzeros
    rescaled so the matrix totals `total` (one pinned at 1.3e6), at uniform
    random columns.  Values N(0,1)."""
    base = [0.1, 12, 0, 1e4, 1.3e6, 0, 1.3e6, window]
    reg = build_lens_total(CIRCUIT, m, m, seed, base)
    dense_total = max(0.0, total - reg)
    prm = [0.1, 12, n_dense, 1e4, 1.3e6, dense_total, 1.3e6, window]
    return build(CIRCUIT, m, m, seed, prm, value_law)


def build_lens_total(kind, m, n, seed, prm):
    p = _params(prm)
    lens = np.zeros(max(m, 1), dtype=np.int32)
    return gen().spx_gen_rowlens(kind, m, n, seed, p.ctypes.data, 16, lens.ctypes.data)


def rail(m=4284, n=1_092_610, total=11_300_000, sigma=1.2, seed=3, value_law=VAL_ONES):
    """LP-constraint-like wide binary matrix.  k ~ round(lognormal(mu', sigma))
    clipped to [1, n] with mu' shifted from the stated 7.5 so the total is
    about `total` (lognormal(7.5, 1.2) would give ~15.9M).  Values 1.0."""
    mu = math.log(total / m) - sigma * sigma / 2
    return build(LOGNORMAL, m, n, seed, [mu, sigma, 1, n], value_law)


def asic(m=682_862, seed=4, window=1024, value_law=None):
    """ASIC-like: regular rows geometric (p=0.31, [1, 64], mean ~3.2), columns
    banded like `circuit`; 8 heavy rows log-uniform [1e4, 3.95e5], one pinned
    at 395k, uniform columns; values +-lognormal(0, 2).  About 3.9M nnz."""
    prm = [0.31, 64, 8, 1e4, 3.95e5, 0, 395_000, window]
    return build(ASIC, m, m, seed, prm, value_law)


def lognormal_rows(m, n, mu, sigma, seed, kmax=None, value_law=None):
    return build(LOGNORMAL, m, n, seed, [mu, sigma, 1, kmax or n], value_law)


def banded(n, bw, seed, value_law=None):
    return build(BANDED, n, n, seed, [bw], value_law)


def blockdiag(n, block, k_mean, n_dense, dense_len, seed, value_law=None):
    return build(BLOCKDIAG, n, n, seed, [block, k_mean, n_dense, dense_len], value_law)


def small_int(m, n, k_mean, seed):
    """Integer values 0..9 with integer x (the recipe of main.cu:324-333):
    every product and partial sum is an exact integer in FP64, so any
    summation order gives the same bits."""
    return build(ZIPF, m, n, seed, [1.2, 1, max(1, min(n, 4 * k_mean))], VAL_SMALLINT)


def g_matrix(n, seed=1):
    """The reference harness's 'g' matrix (SB/test/dspmv_test.cu:137-199) as
    CSR: rows [0, n/8) hold ceil(0.9 n) entries, the rest ceil(0.01 n),
    columns 0..k-1, values rand()/RAND_MAX (glibc, srand(seed))."""
    G = gen()
    nnz = G.spx_gen_g(n, seed, None, None, None)
    if nnz < 0:
        raise ValueError("g generator needs n % 8 == 0 and n >= 8")
    r = np.zeros(nnz, np.int32)
    c = np.zeros(nnz, np.int32)
    v = np.zeros(nnz, np.float64)
    G.spx_gen_g(n, seed, r.ctypes.data, c.ctypes.data, v.ctypes.data)
    counts = np.bincount(r, minlength=n).astype(np.int64)
    rowptr = np.zeros(n + 1, np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return finish(rowptr, c, v, n, n)


def batch_sizes(count=512, seed=5, nmin=1000, nmax=200_000, mu=math.log(14142), sigma=1.32):
    """Square sizes n_i ~ round(lognormal(mu, sigma)) clipped to [nmin, nmax]
    (numpy PCG64 seeded; the law's parameters are not stated, these give a
    median of ~14k and a total of ~1.2e9 nonzeros for 512 matrices)."""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(mu, sigma, count)), nmin, nmax).astype(np.int64)


def batch(count=512, seed=5, s=1.3, kmax=2048, value_law=None, **size_kw):
    """Independent matrices for the batched form: A_i is n_i-by-n_i with
    k ~ Zipf(s) truncated to [1, min(kmax, n_i)] nonzeros per row at uniform
    distinct columns, values U[-1,1).  Member i uses seed 1000*seed + i."""
    ns = batch_sizes(count, seed, **size_kw)
    return [zipf(int(n), s, min(kmax, int(n)), seed=1000 * seed + i, value_law=value_law) for i, n in enumerate(ns)]


def bytes_model(m, n, nnz, value_bytes=8, beta_nonzero=False, k=1):
    """Algorithmic HBM bytes of one SpMV (SpMM when k > 1): int32 row pointer,
    int32 columns, values, x read once, y written (and read when beta != 0)."""
    yb = 8 * m * k * (2 if beta_nonzero else 1)
    return 4 * (m + 1) + (4 + value_bytes) * nnz + 8 * n * k + yb


def ref_getB(m, nnz):
    """SB/include/detail/utils.h:10-14 (x counted per nonzero)."""
    return (m + 1 + nnz) * 4 + (2 * nnz + m) * 8

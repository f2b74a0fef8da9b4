"""Multi-RHS SpMM (spx_plan_execute_multi): Y = alpha*A*X + beta*Y with X, Y
column-major, against the CPU oracle column by column.  Kernel-4 plans with
FLAG_FUSED_SPMM run the fused passes (up to 64 columns per read of A, the
matrix held in registers while the columns go by); the others loop the
single-vector path.  Fused sums are formed in the single-vector engine's
order, so every fused column is bitwise the single-vector result, and
small-integer inputs give the oracle's bits."""
import numpy as np
import pytest

import oracle
import spx
from spx import workloads as W

pytestmark = pytest.mark.gpu


def run_multi(A, k, alpha, beta, pad=(0, 0), seed=1, vals=W.VAL_UNIFORM, **kw):
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **kw)
    m, n = A["m"], A["n"]
    ldx, ldy = n + pad[0], m + pad[1]
    X = W.vector(ldx * k, seed, vals).reshape(k, ldx)  # row j = column j of X
    Y0 = W.vector(ldy * k, seed + 1, vals).reshape(k, ldy)
    xd, yd = spx.DeviceBuffer.from_array(X.ravel()), spx.DeviceBuffer.from_array(Y0.ravel())
    P.execute_multi(k, xd, ldx, yd, ldy, alpha, beta)
    spx.synchronize()
    Y = yd.download(np.float64, ldy * k).reshape(k, ldy)
    refs = [oracle.spmv(A["rowptr"], A["col"], A["val"], X[j, :n], alpha, beta, Y0[j, :m]) for j in range(k)]
    # the padding between columns must be untouched
    assert np.array_equal(Y[:, m:], Y0[:, m:])
    return Y[:, :m], np.stack(refs), P


@pytest.mark.parametrize("k", [1, 2, 3, 8, 17, 32, 37, 70])
@pytest.mark.parametrize("kc", [dict(kernel=4, flags=spx.FLAG_FUSED_SPMM), dict(kernel=4, items=8, flags=spx.FLAG_FUSED_SPMM),
                                dict(kernel=4), dict(kernel=2)],
                         ids=["k4-fused", "k4-fused-8chunk-units", "k4-loop", "k2-loop"])
def test_spmm_matches_oracle(k, kc):
    A = W.zipf(6000, 1.2, 3000, seed=31)
    Y, R, P = run_multi(A, k, 0.75, -0.5, pad=(5, 3), long_row=200, **kc)
    assert P.info()["nlong"] > 0
    for j in range(k):
        rel = np.linalg.norm(Y[j] - R[j]) / np.linalg.norm(R[j])
        assert rel <= 1e-13, (j, rel)


@pytest.mark.parametrize("k", [8, 32])
def test_spmm_small_integers_bit_exact(k):
    A = W.small_int(3000, 4000, 30, seed=8)
    Y, R, _ = run_multi(A, k, 2.0, -1.0, seed=3, vals=W.VAL_SMALLINT, kernel=4, long_row=60,
                        flags=spx.FLAG_FUSED_SPMM)
    assert np.array_equal(Y, R)


def test_spmm_beta_zero_never_reads_y_and_empty_rows():
    rp = np.array([0, 0, 3, 3, 5, 5], np.int64)
    col = np.array([0, 2, 4, 1, 3], np.int32)
    val = np.array([1.0, -2.0, 3.0, 0.5, 4.0])
    A = W.finish(rp, col, val, 5, 5)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=5, kernel=4, flags=spx.FLAG_FUSED_SPMM)
    k = 8
    X = W.vector(5 * k, 4).reshape(k, 5)
    xd = spx.DeviceBuffer.from_array(X.ravel())
    yd = spx.DeviceBuffer.from_array(np.full(5 * k, np.nan))
    P.execute_multi(k, xd, 5, yd, 5, 1.5, 0.0)
    spx.synchronize()
    Y = yd.download(np.float64, 5 * k).reshape(k, 5)
    assert np.all(np.isfinite(Y))
    for j in range(k):
        ref = oracle.spmv(rp, col, val, X[j], 1.5, 0.0)
        # the oracle applies alpha to every product (main.cu:354): equal to rounding
        assert np.allclose(Y[j], ref, rtol=1e-14, atol=0), (j, Y[j], ref)
        assert np.array_equal(Y[j][[0, 2, 4]], np.zeros(3))


def test_spmm_reproducible_and_equal_to_repeated_calls():
    A = W.zipf(5000, 1.3, 2000, seed=9)
    Y1, _, P = run_multi(A, 8, 1.0, 0.0, kernel=4, flags=spx.FLAG_FUSED_SPMM)
    Y2, _, _ = run_multi(A, 8, 1.0, 0.0, kernel=4, flags=spx.FLAG_FUSED_SPMM)
    assert np.array_equal(Y1, Y2)


@pytest.mark.parametrize("items", [2, 8])
def test_spmm_fused_columns_bitwise_equal_single_vector_execute(items):
    # the fused pass adds in the single-vector engine's order: same bits per
    # column, long rows included
    A = W.zipf(8000, 1.2, 4000, seed=12)
    m, n, k = A["m"], A["n"], 9
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=n, kernel=4, items=items, long_row=150,
                 flags=spx.FLAG_FUSED_SPMM)
    assert P.info()["nlong"] > 0
    X = W.vector(n * k, 21).reshape(k, n)
    Y0 = W.vector(m * k, 22).reshape(k, m)
    xd, yd = spx.DeviceBuffer.from_array(X.ravel()), spx.DeviceBuffer.from_array(Y0.ravel())
    P.execute_multi(k, xd, n, yd, m, 1.25, 0.5)
    spx.synchronize()
    Yf = yd.download(np.float64, m * k).reshape(k, m)
    for j in range(k):
        assert np.array_equal(Yf[j], P(X[j], 1.25, 0.5, y=Y0[j])), j

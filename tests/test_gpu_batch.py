"""Batched form (block-diagonal packing, one persistent launch) and hipGraph
replay, against the CPU oracle per matrix.  Graph replays are compared
bitwise with the eager result: the kernels are deterministic, so any
difference is a replay defect."""
import numpy as np
import pytest

import oracle
import spx
from spx import workloads as W

pytestmark = pytest.mark.gpu


def members(count=24, seed=9):
    return W.batch(count, seed=seed, nmin=200, nmax=6000, mu=np.log(1500), sigma=1.0, kmax=400)


def check_members(mats, y, x, x_off, y_off, alpha, beta, y0, tol=1e-13):
    for i, A in enumerate(mats):
        xi = x[x_off[i]:x_off[i + 1]]
        y0i = None if y0 is None else y0[y_off[i]:y_off[i + 1]]
        ref = oracle.spmv(A["rowptr"], A["col"], A["val"], xi, alpha, beta, y0i)
        got = y[y_off[i]:y_off[i + 1]]
        rel = np.linalg.norm(got - ref) / max(np.linalg.norm(ref), 1e-300)
        assert rel <= tol, (i, rel)


@pytest.mark.parametrize("kc", [dict(kernel=1), dict(kernel=2), dict(kernel=3),
                                dict(kernel=2, mode=spx.MODE_NNZ, parts=3),
                                dict(kernel=2, precision=spx.PREC_F32), dict(kernel=4),
                                dict(kernel=4, precision=spx.PREC_F32)],
                         ids=["k1", "k2", "k3", "k2-parts3", "k2-f32", "k4", "k4-f32"])
def test_batch_matches_oracle_per_matrix(kc):
    mats = members()
    P = spx.BatchPlan(mats, nb=20000, **kc)
    x = W.vector(P.n, 4)
    y0 = W.vector(P.m, 5)
    y = P(x, 0.9, -0.4, y0)
    tol = 1e-7 if kc.get("precision") == spx.PREC_F32 else 1e-13
    check_members(mats, y, x, P.x_off, P.y_off, 0.9, -0.4, y0, tol)


@pytest.mark.parametrize("kc", [dict(kernel=2, mode=spx.MODE_QUEUE, nb=20000),
                                dict(kernel=3, mode=spx.MODE_QUEUE, nb=7000, queues=2),
                                dict(kernel=2, mode=spx.MODE_NNZ, parts=4, long_row=100),
                                dict(kernel=4, long_row=100)],
                         ids=["queue", "queue-q2", "parts4-long", "k4-long"])
def test_graph_replays_match_eager_bitwise(kc):
    mats = members(16, seed=3)
    P = spx.BatchPlan(mats, **kc)
    x = W.vector(P.n, 6)
    y0 = W.vector(P.m, 7)
    xd = spx.DeviceBuffer.from_array(x)
    yd = spx.DeviceBuffer(P.m * 8)
    yd.upload(y0)
    P.execute(xd, yd, 1.5, 0.5)
    spx.synchronize()
    eager = yd.download(np.float64, P.m)
    check_members(mats, eager, x, P.x_off, P.y_off, 1.5, 0.5, y0)
    G = spx.Graph(P, xd, yd, 1.5, 0.5)
    for rep in range(3):
        yd.upload(y0)
        # eager work between replays
        P.execute(xd, yd, 1.5, 0.5)
        spx.synchronize()
        assert np.array_equal(yd.download(np.float64, P.m), eager)
        yd.upload(y0)
        G.launch()
        spx.synchronize()
        assert np.array_equal(yd.download(np.float64, P.m), eager), rep
    G.close()

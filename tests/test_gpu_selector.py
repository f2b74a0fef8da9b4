"""The selector's HIP kernels against the numpy restatement (oracle/mlp.py):
forward scores, the loss gradient, an Adam step, the device features and the
masked argmax; then the trained model against exhaustive search on held-out
matrices (regret and the gain over fixed schedules)."""
import os

import numpy as np
import pytest

import oracle
from oracle import mlp
import spx
from spx import selector as S
from spx import workloads as W

pytestmark = pytest.mark.gpu


def rand_case(seed, B=37):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.25, mlp.NPARAMS).astype(np.float32)
    X = rng.uniform(0, 1, (B, 32)).astype(np.float32)
    T = rng.dirichlet(np.full(96, 0.3), B).astype(np.float32)
    return p, X, T


def test_forward_matches_numpy():
    p, X, _ = rand_case(1)
    Z = S.MLP(p).forward(X)
    _, _, Zr = mlp.forward(p, X)
    assert np.allclose(Z, Zr, rtol=1e-4, atol=1e-4 * np.abs(Zr).max())


@pytest.mark.parametrize("B", [1, 37, 256])
def test_gradient_matches_numpy(B):
    p, X, T = rand_case(2 + B, B)
    loss, g = S.MLP(p).grad(X, T)
    lr, gr = mlp.loss_grad(p, X, T)
    assert abs(loss - lr) <= 1e-4 * max(1.0, abs(lr))
    scale = np.abs(gr).max()
    assert np.allclose(g, gr, rtol=2e-3, atol=2e-4 * scale), np.abs(g - gr).max() / scale


def test_adam_step_matches_numpy():
    p, X, T = rand_case(7, 64)
    net = S.MLP(p)
    for _ in range(3):
        net._step(X, T, 1e-2, 1e-3)
    q, m1, m2 = p.astype(np.float64), np.zeros(mlp.NPARAMS), np.zeros(mlp.NPARAMS)
    for step in range(1, 4):
        _, g = mlp.loss_grad(q, X, T)
        q, m1, m2 = mlp.adam(q, m1, m2, g, step, 1e-2, 1e-3)
    got = net.params()
    assert np.allclose(got, q, rtol=1e-3, atol=1e-4), np.abs(got - q).max()


MATS = {
    "zipf": lambda: W.zipf(20000, 1.4, 3000, seed=3),
    "banded": lambda: W.banded(50000, 21, 4),
    "blockdiag": lambda: W.blockdiag(40000, 64, 12, 5, 30000, 5),
    "circuit_small": lambda: W.circuit(m=300_000, total=3_000_000, n_dense=20, seed=6),
    "empty_rows": lambda: W.finish(np.array([0, 0, 3, 3, 3, 7, 7]), np.array([1, 2, 5, 0, 1, 3, 6], np.int32),
                                   np.ones(7), 6, 8),
}


@pytest.mark.parametrize("name", list(MATS))
def test_device_features_match_counting(name):
    A = MATS[name]()
    f = S.features(A)
    ref = mlp.features(A["rowptr"], A["col"], A["m"], A["n"], S.long_row_threshold(A["n"]))
    # counts are exact; the log-distance mean may differ in a float32 rounding
    assert np.allclose(f[:28], ref[:28], rtol=1e-6, atol=1e-7), np.abs(f - ref)[:28].max()
    assert np.allclose(f[28:], ref[28:], rtol=1e-3, atol=1e-6)


def test_argmax_with_mask():
    rng = np.random.default_rng(3)
    p = rng.normal(0, 0.5, mlp.NPARAMS).astype(np.float32)
    X = rng.uniform(0, 1, (50, 32)).astype(np.float32)
    net = S.MLP(p)
    Z = net.forward(X)
    assert np.array_equal(net.choose(X), Z.argmax(1))
    mask = np.zeros(96, np.int32)
    mask[72:] = 1  # kernel 4 only
    ch = net.choose(X, mask)
    assert np.array_equal(ch, 72 + Z[:, 72:].argmax(1))


MODEL = S.MODEL_PATH
HELDOUT = os.path.join(os.path.dirname(S.MODEL_PATH), "selector_data", "heldout.npz")
TRAIN = os.path.join(os.path.dirname(S.MODEL_PATH), "selector_data", "train.npz")


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained selector model")
def test_selected_plan_is_correct():
    import oracle
    A = W.zipf(30000, 1.3, 2000, seed=12)
    s, kw = S.select(A)
    assert 0 <= s < 96 and kw == S.schedule(s)
    x = W.vector(A["n"], 2)
    P = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **kw)
    y = P(x)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x, 1.0, 0.0)
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) <= 1e-13


@pytest.mark.skipif(not os.path.exists(MODEL), reason="no trained selector model")
def test_shortlist_and_autotune():
    """The shortlist starts with the selector's pick and holds distinct
    schedules; the autotuner returns the fastest of the ones it timed, and
    the schedule it returns computes the oracle's result."""
    A, _ = S.training_matrix(10_005)
    f = S.features(A)
    sl = S.shortlist(f, top=4)
    assert len(sl) == 4 and len(set(sl)) == 4
    assert sl[0] == S.representative(S.select(A)[0])
    x = W.vector(A["n"], 3)
    xd, yd = spx.DeviceBuffer.from_array(x), spx.DeviceBuffer(A["m"] * 8)
    best, t = S.autotune(lambda kw: spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **kw), xd, yd, sl)
    assert set(t) <= set(sl) and best in t and t[best] == min(t.values()) and t[best] > 0
    y = spx.Plan(A["rowptr"], A["col"], A["val"], n=A["n"], **S.schedule(best))(x)
    ref = oracle.spmv(A["rowptr"], A["col"], A["val"], x)
    assert np.linalg.norm(y - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.skipif(not (os.path.exists(MODEL) and os.path.exists(HELDOUT) and os.path.exists(TRAIN)),
                    reason="no trained selector model or held-out timings")
def test_selector_ablation_on_heldout():
    """The trained network against the method without it: the default
    schedule and the best single fixed schedule (chosen on the training
    timings), judged by regret against exhaustive search on held-out
    matrices timed on the MI355X.  The features are recomputed on the device
    for the first held-out matrices and must match the recorded ones."""
    with np.load(HELDOUT, allow_pickle=False) as f:
        X, Tm, seeds, m, gen = f["features"], f["times"], f["seeds"], f["m"], S.gen_params(f)
    with np.load(TRAIN, allow_pickle=False) as f:
        Ttr = f["times"]
    for i in range(3):
        A, _ = S.training_matrix(int(seeds[i]), *gen)
        assert np.allclose(S.features(A), X[i], rtol=1e-5, atol=1e-6)
    rep = S.report(S.MLP.load(), X, Tm, train_times=Ttr)
    print(rep, flush=True)
    # the held-out set: 1,000 matrices spanning n in [1e4, 6e6], every
    # distinct schedule timed (stream partitions included)
    assert rep["matrices"] >= 1000 and m.min() >= 10_000 and m.max() >= 3_000_000
    assert not np.isnan(Tm).any()
    assert rep["selector_regret_mean"] < rep["default_regret_mean"]
    # the network beats the best single schedule by at least one point of regret
    assert rep["selector_regret_mean"] <= rep["fixed_regret_mean"] - 0.01

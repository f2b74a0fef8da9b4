"""The selector's CPU-side pieces: the numpy restatement of the network is
pinned by finite differences, the schedule table by enumeration, the soft
targets and the regret by their definitions."""
import numpy as np

from oracle import mlp
from spx import selector as S


def test_numpy_backward_matches_central_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, mlp.NPARAMS)
    X = rng.uniform(0, 1, (5, 32))
    T = rng.dirichlet(np.ones(96), 5)
    _, g = mlp.loss_grad(p, X, T)
    idx = rng.choice(mlp.NPARAMS, 200, replace=False)
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        num = (mlp.loss(p + e, X, T) - mlp.loss(p - e, X, T)) / (2 * h)
        assert abs(num - g[i]) <= 1e-7 + 1e-5 * abs(num), (i, num, g[i])


def test_parameter_layout():
    assert mlp.NPARAMS == 64 * 32 + 64 + 64 * 64 + 64 + 96 * 64 + 96 == 12512


def test_schedule_table():
    seen = set()
    for s in range(S.NSCHED):
        kw = S.schedule(s)
        assert kw["kernel"] == 1 + s // 24
        seen.add(tuple(sorted(kw.items())))
    assert len(seen) == len(S.distinct()) == 76
    assert S.schedule(S.DEFAULT) == dict(kernel=2, items=8, parts=1)
    assert all(S.schedule(S.representative(s)) == S.schedule(s) for s in range(S.NSCHED))


def test_targets_and_regret():
    t = np.array([[2.0, 1.0, 4.0, np.nan], [1.0, 1.0, 3.0, 1.5]])
    T = S.targets(t, tau=0.05)
    assert np.allclose(T.sum(1), 1.0)
    assert T[0, 3] == 0 and T[0].argmax() == 1
    assert np.isclose(T[1, 0], T[1, 1])
    r = S.regret(t, np.array([0, 2]))
    assert np.allclose(r, [1.0, 2.0])


# ---- Adam: closed forms of the first steps ------------------------------------------


def test_adam_first_step_closed_form():
    # with zero moments, step 1 and no weight decay the bias-corrected moments
    # are g and g^2, so p1 = p0 - lr * g / (|g| + eps)
    rng = np.random.default_rng(3)
    p0 = rng.normal(0, 1, 50)
    g = rng.normal(0, 1, 50) * 10.0 ** rng.integers(-3, 3, 50)
    lr, eps = 0.01, 1e-8
    p1, m1, m2 = mlp.adam(p0, np.zeros(50), np.zeros(50), g, 1, lr, 0.0)
    assert np.allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=1e-12, atol=0)
    assert np.allclose(m1, 0.1 * g, rtol=1e-15) and np.allclose(m2, 0.001 * g * g, rtol=1e-15)


def test_adam_constant_gradient_geometric_moments():
    # a constant gradient for t steps: m1 = (1 - b1^t) g and m2 = (1 - b2^t) g^2
    # (geometric series), and the bias correction makes every step move each
    # parameter by lr * g / (|g| + eps) whatever t is
    g = np.array([0.5, -3.0, 1e-3, 7.0])
    p = np.array([1.0, 2.0, -1.0, 0.0])
    m1 = np.zeros(4)
    m2 = np.zeros(4)
    lr = 0.05
    for t in range(1, 8):
        pn, m1, m2 = mlp.adam(p, m1, m2, g, t, lr, 0.0)
        assert np.allclose(m1, (1 - 0.9 ** t) * g, rtol=1e-13)
        assert np.allclose(m2, (1 - 0.999 ** t) * g * g, rtol=1e-12)
        assert np.allclose(p - pn, lr * g / (np.abs(g) + 1e-8), rtol=1e-9)
        p = pn


def test_adam_weight_decay_enters_the_gradient():
    # wd * p is added to the gradient: with g = 0 the first step moves each
    # parameter by lr * sign(p) (|wd p| >> eps)
    p0 = np.array([2.0, -0.5, 4.0])
    p1, _, _ = mlp.adam(p0, np.zeros(3), np.zeros(3), np.zeros(3), 1, 0.1, 0.01)
    assert np.allclose(p1, p0 - 0.1 * np.sign(p0), rtol=1e-6)


# ---- features: a hand-computed example ----------------------------------------------


def test_features_hand_computed_six_rows():
    import math

    # m = 6, n = 8: row i's diagonal column is i * 8 // 6 = 0, 1, 2, 4, 5, 6.
    # Row lengths 2, 0, 4, 1, 0, 8 (two empty rows; with the long-row
    # threshold at 4, rows 2 and 5 are long).  Columns are chosen so every
    # |col - diag| is 0, 1 or 3, whose log2(d + 1) = 0, 1, 2 are exact:
    #   row 0 (diag 0): cols 0 1          -> d 0 1
    #   row 2 (diag 2): cols 1 2 3 5      -> d 1 0 1 3
    #   row 3 (diag 4): col 7             -> d 3
    #   row 5 (diag 6): cols 3 5 6 7 3 5 6 7 -> d 3 1 0 1 3 1 0 1
    # so d = 0 four times, 1 seven times, 3 four times (15 nonzeros).
    rowptr = np.array([0, 2, 2, 6, 7, 7, 15], np.int64)
    col = np.array([0, 1, 1, 2, 3, 5, 7, 3, 5, 6, 7, 3, 5, 6, 7], np.int32)
    f = mlp.features(rowptr, col, 6, 8, 4)
    sd = math.sqrt(47.5 / 6)  # E[L^2] - mean^2 = 85/6 - 2.5^2
    exp = np.zeros(32)
    exp[0] = math.log2(7) / 24
    exp[1] = math.log2(9) / 24
    exp[2] = 4 / 32  # log2(16) / 32
    exp[3] = math.log2(3.5) / 12  # mean 2.5
    exp[4] = sd / 2.5 / 8
    exp[5] = math.log2(9) / 24  # longest row 8
    exp[6] = 2 / 6  # empty rows
    exp[7] = 12 / 15  # nonzeros in rows of at least 4
    exp[8:12] = 1 / 6  # floor(log2 L) of 2, 4, 1, 8 = 1, 2, 0, 3
    exp[24] = 3 / 15  # rows under 4 nonzeros hold 2 + 1
    exp[25] = 12 / 15  # rows of 4..31 hold 4 + 8
    exp[28] = (7 * 1024 + 4 * 2048) / 1024 / 15 / 24  # = 1/24
    exp[29] = 1.0  # every d < 1024
    exp[30] = 1.0  # every d < 32768
    exp[31] = math.log2(sd + 1) / 12
    np.testing.assert_allclose(f, exp, rtol=1e-12, atol=1e-15)
    # a transposed diagonal (row * m // n) or a dropped empty-row count would
    # move f[28] and f[6]
    assert abs(exp[28] - 1 / 24) < 1e-15

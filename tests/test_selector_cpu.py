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

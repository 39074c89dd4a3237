"""Pins the MLP oracle (oracle/mlp.py): its back-propagated gradient equals
central finite differences of its loss, and its Adam step equals the
textbook update written out for a single parameter."""
import numpy as np

from oracle import mlp


def _problem(seed=0, n=40):
    rng = np.random.default_rng(seed)
    w = rng.normal(0, 0.3, mlp.NPARAMS)
    x = rng.normal(0, 1, (n, mlp.NF))
    t = rng.normal(2, 1, n)
    return w, x, t


def test_parameter_count():
    assert mlp.NPARAMS == 6337


def test_gradient_matches_finite_differences():
    w, x, t = _problem()
    L, g = mlp.gradient(w, x, t)
    assert abs(L - mlp.loss(w, x, t)) < 1e-12
    rng = np.random.default_rng(1)
    # every block of the parameter vector is probed
    idx = np.concatenate([rng.choice(2048, 20, replace=False), 2048 + rng.choice(64, 5, replace=False),
                          2112 + rng.choice(4096, 20, replace=False), 6208 + rng.choice(64, 5, replace=False),
                          6272 + rng.choice(64, 5, replace=False), [6336]])
    fd = mlp.finite_difference(w, x, t, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-7)


def test_gradient_flags_a_planted_error():
    """A backward pass that forgets the second ReLU mask disagrees with the
    finite differences (so the check above can fail)."""
    w, x, t = _problem(2)
    p = mlp.unpack(w)
    s1 = x @ p["W1"].T + p["b1"]
    h1 = np.maximum(s1, 0)
    s2 = h1 @ p["W2"].T + p["b2"]
    h2 = np.maximum(s2, 0)
    y = h2 @ p["W3"] + p["b3"]
    dy = 2 * (y - t) / len(x)
    wrong_gW2 = (np.outer(dy, p["W3"])).T @ h1  # missing (s2 > 0)
    idx = 2112 + np.arange(0, 4096, 97)
    fd = mlp.finite_difference(w, x, t, idx)
    assert not np.allclose(wrong_gW2.ravel()[idx - 2112], fd, rtol=1e-3, atol=1e-6)


def test_adam_step_textbook():
    w = np.array([1.0])
    g = np.array([0.5])
    m = np.zeros(1)
    v = np.zeros(1)
    w1, m1, v1 = mlp.adam_step(w, g, m, v, lr=0.1, t=1)
    # first step: m_hat = g, v_hat = g^2, so the move is lr * sign(g)
    assert np.allclose(w1, 1.0 - 0.1 * 0.5 / (0.5 + 1e-8))
    assert np.allclose(m1, 0.05) and np.allclose(v1, 0.001 * 0.25)

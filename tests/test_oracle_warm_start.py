"""The FP64 oracle of the learned warm-start network, pinned by central
finite differences of its own loss (every parameter group), and its BF16
emulation against the plain FP64 network."""
import numpy as np
import pytest

from oracle import warm_start as W


def _case(seed, n=40):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=W.N_PARAMS) * 0.3
    X = np.zeros((n, 16))
    X[:, :8] = rng.uniform(-1, 1, (n, 8))
    t = rng.normal(size=n)
    return p, X, t


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    p, X, t = _case(seed)
    _, g, _ = W.loss_and_grad(p, X, t)
    rng = np.random.default_rng(100 + seed)
    groups = [(W.W1, W.B1), (W.B1, W.W2), (W.W2, W.B2), (W.B2, W.W3), (W.W3, W.B3),
              (W.B3, W.N_PARAMS)]
    eps = 1e-6
    for lo, hi in groups:
        for i in rng.choice(np.arange(lo, hi), size=min(12, hi - lo), replace=False):
            q = p.copy()
            q[i] += eps
            lp = W.loss_and_grad(q, X, t)[0]
            q[i] -= 2 * eps
            lm = W.loss_and_grad(q, X, t)[0]
            assert abs((lp - lm) / (2 * eps) - g[i]) <= 1e-7 * max(1.0, abs(g[i])), (lo, i)


def test_bf16_rounding_and_emulated_network():
    # bf16(): nearest even on the 16 dropped bits (the ulp at 1 is 2^-7).
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -9, -2.5, 0.0])
    np.testing.assert_array_equal(W.bf16(v), [1.0, 1.0, 1.0, 1.0 + 2 ** -7, -2.5, 0.0])
    # The emulated network differs from the FP64 one by BF16-sized amounts.
    p, X, t = _case(3, n=200)
    y, _ = W.forward(p, X)
    ye, _ = W.forward(p, X, emulate=True)
    err = np.abs(y - ye).max() / np.abs(y).max()
    assert 1e-5 < err < 3e-2

"""CPU pins of the correction-network oracle: backward formulas against
central finite differences (float64), BF16 rounding against exact cases."""
import numpy as np

from oracle import mlp as M


def _init(n_in, n_out, seed):
    in_p, out_p, n = M.sizes(n_in, n_out)
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(n) * 0.2).astype(np.float32)
    W1, b1, W2, b2, W3, b3 = M.unpack(p.copy(), in_p, out_p)
    W1[:, n_in:] = 0
    W3[n_out:, :] = 0
    b3[n_out:] = 0
    return M.pack(W1, b1, W2, b2, W3, b3), in_p, out_p


def test_bf16_rounding_exact_cases():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 0.0, 65280.0], np.float32)
    r = M.bf16(x)
    assert r[0] == 1.0
    assert r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2 ** -7  # above half -> up
    assert r[3] == -2.5 and r[4] == 0.0 and r[5] == 65280.0


def test_backward_matches_finite_differences():
    n_in, n_out = 26, 20
    p, in_p, out_p = _init(n_in, n_out, 0)
    rng = np.random.default_rng(1)
    X = np.zeros((24, in_p))
    X[:, :n_in] = rng.standard_normal((24, n_in))
    T = np.zeros((24, out_p))
    T[:, :n_out] = rng.standard_normal((24, n_out))
    g = M.grad64(p, X, T, n_out, in_p, out_p)
    idx = rng.choice(p.size, 60, replace=False)
    eps = 1e-6
    for i in idx:
        e = np.zeros(p.size)
        e[i] = eps
        fd = (M.loss64(p + e, X, T, n_out, in_p, out_p) - M.loss64(p - e, X, T, n_out, in_p, out_p)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd)


def test_fp32_gradient_matches_float64_formulas():
    n_in, n_out = 16, 10
    p, in_p, out_p = _init(n_in, n_out, 2)
    rng = np.random.default_rng(3)
    X = np.zeros((64, in_p), np.float32)
    X[:, :n_in] = rng.standard_normal((64, n_in))
    T = np.zeros((64, out_p), np.float32)
    T[:, :n_out] = rng.standard_normal((64, n_out))
    l32, g32 = M.loss_grad(p, X, T, n_out, in_p, out_p, round_weights=False)
    g64 = M.grad64(p, X.astype(np.float64), T.astype(np.float64), n_out, in_p, out_p)
    assert np.linalg.norm(g32 - g64) <= 1e-5 * np.linalg.norm(g64)
    assert abs(l32 - M.loss64(p, X.astype(np.float64), T.astype(np.float64), n_out, in_p, out_p)) < 1e-6


def test_upstream_backward_matches_finite_differences():
    """The back-propagation of a given upstream gradient dY is the gradient of the
    linear functional sum(dY * y): central differences of the float64 forward."""
    n_in, n_out = 22, 10
    p, in_p, out_p = _init(n_in, n_out, 4)
    rng = np.random.default_rng(5)
    X = np.zeros((20, in_p))
    X[:, :n_in] = rng.standard_normal((20, n_in))
    dY = np.zeros((20, out_p))
    dY[:, :n_out] = rng.standard_normal((20, n_out))
    g = M.backward_upstream(p, X, dY, n_out, in_p, out_p, dtype=np.float64, round_weights=False)
    idx = rng.choice(p.size, 60, replace=False)
    eps = 1e-6
    for i in idx:
        e = np.zeros(p.size)
        e[i] = eps
        fd = (M.linear_loss64(p + e, X, dY, n_out, in_p, out_p) - M.linear_loss64(p - e, X, dY, n_out, in_p, out_p)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd)
    # with the MSE's upstream gradient 2 (y - T) / (B n_out) it is the MSE gradient
    T = np.zeros((20, out_p))
    T[:, :n_out] = rng.standard_normal((20, n_out))
    W1, b1, W2, b2, W3, b3 = M.unpack(np.asarray(p, np.float64), in_p, out_p)
    y = np.maximum(np.maximum(X @ W1.T + b1, 0) @ W2.T + b2, 0) @ W3.T + b3
    dmse = np.zeros_like(y)
    dmse[:, :n_out] = 2 * (y[:, :n_out] - T[:, :n_out]) / (20 * n_out)
    g2 = M.backward_upstream(p, X, dmse, n_out, in_p, out_p, dtype=np.float64, round_weights=False)
    assert np.allclose(g2, M.grad64(p, X, T, n_out, in_p, out_p), rtol=1e-6, atol=1e-9)   # pack() stores FP32

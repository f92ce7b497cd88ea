"""CPU oracle of the per-cell correction network (TEST INFRASTRUCTURE).

numpy restatement of the device network: in -> 128 -> 128 -> out, ReLU.
Parameter layout (flat FP32): W1[H][in_p], b1[H], W2[H][H], b2[H],
W3[out_p][H], b3[out_p]; rows are output units.
  - inference emulates the MFMA path: inputs, weights and hidden activations
    rounded to BF16 (round-to-nearest-even), FP32 accumulation;
  - training emulates the FP32 kernels: BF16-rounded weights, FP32 activations,
    mean-squared error over the first n_out outputs.
The backward formulas are pinned against central finite differences of the
float64 forward (tests/test_mlp_oracle.py).
"""
import numpy as np

H = 128


def bf16(a):
    """round float32 values to bfloat16 (nearest-even), returned as float32."""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    nonfinite = ~np.isfinite(a)
    out[nonfinite] = a[nonfinite]
    return out


def sizes(n_in, n_out):
    in_p = (n_in + 15) // 16 * 16
    out_p = (n_out + 15) // 16 * 16
    return in_p, out_p, H * in_p + H + H * H + H + out_p * H + out_p


def unpack(p, in_p, out_p):
    o = 0
    W1 = p[o:o + H * in_p].reshape(H, in_p); o += H * in_p
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + out_p * H].reshape(out_p, H); o += out_p * H
    b3 = p[o:o + out_p]
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3]).astype(np.float32)


def forward_infer(p, X, in_p, out_p):
    W1, b1, W2, b2, W3, b3 = unpack(np.asarray(p, np.float32), in_p, out_p)
    x = bf16(X).astype(np.float64)
    h1 = bf16(np.maximum(x @ bf16(W1).T.astype(np.float64) + b1, 0).astype(np.float32)).astype(np.float64)
    h2 = bf16(np.maximum(h1 @ bf16(W2).T.astype(np.float64) + b2, 0).astype(np.float32)).astype(np.float64)
    return (h2 @ bf16(W3).T.astype(np.float64) + b3).astype(np.float32)


def loss_grad(p, X, T, n_out, in_p, out_p, dtype=np.float32, round_weights=True):
    """MSE loss over the first n_out outputs and its gradient w.r.t. p."""
    p = np.asarray(p, np.float32)
    W1, b1, W2, b2, W3, b3 = [a.astype(dtype) for a in unpack(p, in_p, out_p)]
    if round_weights:
        W1, W2, W3 = (bf16(W).astype(dtype) for W in (W1, W2, W3))
    X = np.asarray(X, dtype)
    T = np.asarray(T, dtype)
    B = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3.T + b3
    diff = np.zeros_like(y)
    diff[:, :n_out] = y[:, :n_out] - T[:, :n_out]
    loss = float(np.sum(diff.astype(np.float64) ** 2) / (B * n_out))
    dy = 2 * diff / (B * n_out)
    gW3 = dy.T @ h2
    gb3 = dy.sum(0)
    dh2 = (dy @ W3) * (a2 > 0)
    gW2 = dh2.T @ h1
    gb2 = dh2.sum(0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    gW1 = dh1.T @ X
    gb1 = dh1.sum(0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def loss64(p, X, T, n_out, in_p, out_p):
    """float64 loss with unrounded weights (finite-difference reference)."""
    W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in unpack(np.asarray(p, np.float64), in_p, out_p)]
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ W3.T + b3
    d = y[:, :n_out] - T[:, :n_out]
    return float(np.sum(d ** 2) / (X.shape[0] * n_out))


def grad64(p, X, T, n_out, in_p, out_p):
    """float64 analytic gradient (same formulas as loss_grad, no rounding)."""
    p64 = np.asarray(p, np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p64, in_p, out_p)
    B = X.shape[0]
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    y = h2 @ W3.T + b3
    diff = np.zeros_like(y)
    diff[:, :n_out] = y[:, :n_out] - T[:, :n_out]
    dy = 2 * diff / (B * n_out)
    dh2 = (dy @ W3) * (a2 > 0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    return np.concatenate([(dh1.T @ X).ravel(), dh1.sum(0), (dh2.T @ h1).ravel(), dh2.sum(0),
                           (dy.T @ h2).ravel(), dy.sum(0)])


def backward_upstream(p, X, dY, n_out, in_p, out_p, dtype=np.float32, round_weights=True):
    """Parameter gradient for a given upstream gradient dY = dL/dy (first n_out
    columns), as the device's bh_mlp_backward: FP32 forward with BF16-rounded
    weights, the same back-propagation as loss_grad with dy = dY."""
    W1, b1, W2, b2, W3, b3 = [a.astype(dtype) for a in unpack(np.asarray(p, np.float32), in_p, out_p)]
    if round_weights:
        W1, W2, W3 = (bf16(W).astype(dtype) for W in (W1, W2, W3))
    X = np.asarray(X, dtype)
    a1 = X @ W1.T + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2.T + b2
    h2 = np.maximum(a2, 0)
    dy = np.zeros((X.shape[0], out_p), dtype)
    dy[:, :n_out] = np.asarray(dY, dtype)[:, :n_out]
    dh2 = (dy @ W3) * (a2 > 0)
    dh1 = (dh2 @ W2) * (a1 > 0)
    return pack(dh1.T @ X, dh1.sum(0), dh2.T @ h1, dh2.sum(0), dy.T @ h2, dy.sum(0))


def linear_loss64(p, X, dY, n_out, in_p, out_p):
    """float64 L = sum(dY * y) with unrounded weights: its gradient is the
    back-propagation of the upstream gradient dY (finite-difference reference)."""
    W1, b1, W2, b2, W3, b3 = [a.astype(np.float64) for a in unpack(np.asarray(p, np.float64), in_p, out_p)]
    h1 = np.maximum(X @ W1.T + b1, 0)
    h2 = np.maximum(h1 @ W2.T + b2, 0)
    y = h2 @ W3.T + b3
    return float(np.sum(np.asarray(dY, np.float64)[:, :n_out] * y[:, :n_out]))

"""Training of the learned per-cell correction (host orchestration; all
arithmetic runs in the HIP library: multigrid iterations, operator applies,
feature maps, forward/backward and Adam are device kernels called through the
C ABI).

The network is used after a multigrid cycle: it maps the residual r' that the
cycle leaves (per cell: its own modes, its face neighbours' mean residuals and
diffusion ratios) to a correction e, which the preconditioner adds in the
residual-minimising combination with the cycle's correction.  It is trained
for exactly that use, online, on the post-cycle residuals of multigrid
iterations x <- x + MG(f - A x) over fresh right-hand sides f: each step takes
one such residual (a (residual, error) pair with error A^-1 r') and minimises
the residual left by the optimally scaled correction,
    L = min_a |r' - a A e|^2 / |r'|^2 = 1 - <A e, r'>^2 / (|A e|^2 |r'|^2),
i.e. the error of the prediction in the A^T A norm up to its scale (the step in
the preconditioner chooses the scale), without solving for A^-1 r'.  With
q = A e, a = <q, r'>, b = <q, q>, c = |r'|^2 and alpha = a / b, the upstream
gradient is dL/de = -(2 a / (b c)) A (r' - alpha q) (A symmetric); it is mapped
through e = s h^2 / nu * y to the network outputs and back-propagated by the
device kernels.
"""
import time

import numpy as np

from . import CellMLP, DeviceArray, dot, mlp_correction, mlp_features, mlp_upstream


def network_for(op, seed=0):
    """A network of the widths the learned step of op's multigrid expects:
    Np + 4D inputs (own residual modes, neighbour residuals, diffusion ratios),
    Np outputs."""
    return CellMLP(op.ctx, op.Np + 4 * op.grid.D, op.Np, seed=seed)


def train_resmin(net, op, mg, rhs_fn, steps, iters_per_rhs=8, lr=3e-3, lr_final=3e-4, log_every=0, cells=32768,
                 seed=0):
    """Online training on post-cycle residuals.

    The correction and the residual involve every cell; the parameter gradient is
    back-propagated from a contiguous slice of `cells` cells at a random offset (an
    unbiased estimate up to a constant factor, which Adam's per-parameter
    normalisation removes).

    rhs_fn(i) -> the i-th fresh right-hand side (N FP64 values); every
    `iters_per_rhs` steps a new one starts from x = 0, so the samples cover the
    early and the late residuals of the iteration.  Returns (loss history, seconds,
    number of right-hand sides used)."""
    ctx = op.ctx
    N, J = op.N, op.grid.J
    in_p, out_p = net.in_p, net.out_p
    X = DeviceArray(ctx, J * in_p, np.float32)
    s = DeviceArray(ctx, J, np.float32)
    Y = DeviceArray(ctx, J * out_p, np.float32)
    dY = DeviceArray(ctx, J * out_p, np.float32)
    f = DeviceArray(ctx, N)
    x = DeviceArray(ctx, N)
    r = DeviceArray(ctx, N)
    e = DeviceArray(ctx, N)
    res = DeviceArray(ctx, N)
    g = DeviceArray(ctx, N)
    hist = []
    nrhs = 0
    rng = np.random.default_rng(seed)
    nb = min(J, cells)
    t0 = time.perf_counter()
    for it in range(steps):
        if it % iters_per_rhs == 0:
            f.upload(rhs_fn(nrhs))
            nrhs += 1
            x.upload(np.zeros(N))
        mg.Solve(x, f)                                    # x <- x + MG(f - A x)
        r.copy_from(f)
        op.Evaluate(-1.0, 1.0, r, x, affine=False)        # r' = f - A x
        rr = dot(ctx, r, r)
        if not rr > 0:
            continue
        mlp_features(op, r, X, s, in_p)
        net.forward(X, Y)
        mlp_correction(op, s, Y, e, out_p)                # e = s h^2 / nu * y
        op.Evaluate(1.0, 0.0, g, e, affine=False)         # q = A e (in g)
        a = dot(ctx, g, r)
        b = dot(ctx, g, g)
        if not b > 0:
            continue
        res.copy_from(r)
        op.Evaluate(-a / b, 1.0, res, e, affine=False)    # r' - alpha A e
        loss = 1.0 - a * a / (b * rr)
        op.Evaluate(1.0, 0.0, g, res, affine=False)       # A (r' - alpha A e)
        mlp_upstream(op, s, g, -2.0 * a / (b * rr), dY, out_p)   # dL/dy
        off = int(rng.integers(0, J - nb + 1))
        net.backward(X.view(off * in_p, nb * in_p), dY.view(off * out_p, nb * out_p))
        lr_t = lr * (lr_final / lr) ** (it / max(1, steps - 1))
        net.adam_step(lr_t)
        hist.append(loss)
        if log_every and it % log_every == 0:
            print(f"  train step {it} residual fraction left {loss:.4f}", flush=True)
    ctx.sync()
    for a in (X, s, Y, dY, f, x, r, e, res, g):
        a.free()
    return hist, time.perf_counter() - t0, nrhs


def train(mlp, X, T, nrows, steps=1500, batch=8192, lr=1e-3, seed=0, log_every=0):
    """Supervised Adam on (features, targets) rows with the mean-squared error,
    contiguous mini-batches at random offsets; returns (loss history, seconds)."""
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        off = int(rng.integers(0, max(1, nrows - batch)))
        b = min(batch, nrows)
        loss, _ = mlp.gradient(X.view(off * mlp.in_p, b * mlp.in_p), T.view(off * mlp.out_p, b * mlp.out_p),
                               want_grad=False)
        lr_t = lr * (0.1 ** (it / max(1, steps)))
        mlp.adam_step(lr_t)
        hist.append(loss)
        if log_every and it % log_every == 0:
            print(f"  train step {it} loss {loss:.4e}", flush=True)
    mlp.ctx.sync()
    return hist, time.perf_counter() - t0

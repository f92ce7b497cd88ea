"""Training of the learned per-cell coarse correction (host orchestration; all
arithmetic runs in the HIP library: solves, feature maps, forward/backward and
Adam are device kernels called through the C ABI).

Data: (residual, error) pairs harvested from plain multigrid iterations
x_{k+1} = x_k + MG(f - A x_k) on fresh right-hand sides, with the error
e_k = u - x_k taken against an FGMRES + multigrid solution u converged to
1e-10: A e_k = r_k exactly, so every pair is a true residual/error pair.
"""
import time

import numpy as np

from . import (CellMLP, DeviceArray, FlexGMRES, mlp_features, mlp_targets)


def harvest_pairs(op, mg, rhs_list, iters_per_rhs=6, skip=1):
    """-> (X, T) device arrays with one row per (sample, cell)."""
    ctx = op.ctx
    N, J = op.N, op.grid.J
    D = op.grid.D
    in_p = (op.Np + 2 * D + 15) // 16 * 16
    out_p = (op.Np + 15) // 16 * 16
    nsamp = len(rhs_list) * (iters_per_rhs - skip)
    X = DeviceArray(ctx, nsamp * J * in_p, np.float32)
    T = DeviceArray(ctx, nsamp * J * out_p, np.float32)
    s = DeviceArray(ctx, J, np.float32)
    u = DeviceArray(ctx, N)
    x = DeviceArray(ctx, N)
    r = DeviceArray(ctx, N)
    e = DeviceArray(ctx, N)
    solver = FlexGMRES(op=op, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-10, MaxSolverIterations=400)
    k = 0
    for f in rhs_list:
        fd = DeviceArray.from_host(ctx, f)
        u.upload(np.zeros(N))
        solver.Solve(u, fd)
        ud = u.download()
        x.upload(np.zeros(N))
        for it in range(iters_per_rhs):
            mg.Solve(x, fd)                       # x <- x + V(f - A x)
            if it < skip:
                continue
            op.Evaluate(-1.0, 0.0, r, x, affine=False)   # r = -A x
            rr = r.download() + f
            r.upload(rr)
            e.upload(ud - x.download())
            mlp_features(op, r, X.view(k * J * in_p, J * in_p), s, in_p)
            mlp_targets(op, s, e, T.view(k * J * out_p, J * out_p), out_p)
            k += 1
    ctx.sync()
    return X, T, nsamp, in_p, out_p


def train(mlp, X, T, nrows, steps=1500, batch=8192, lr=1e-3, seed=0, log_every=0):
    """Adam on contiguous mini-batches of rows (random offsets); returns the loss history."""
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

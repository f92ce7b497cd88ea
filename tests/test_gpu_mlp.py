"""GPU checks of the learned per-cell correction against the numpy oracle.

forward (BF16 MFMA inference) vs oracle emulating the same rounding points;
gradient (FP32 device backward) vs oracle; feature/target maps vs numpy;
training reduces the loss; the multigrid preconditioner with the trained
network attached still converges to the FP64 tolerance (fallback rule), and
the test reports what the network gains in iterations.
"""
import numpy as np
import pytest

import bosss_amd as B
import oracle as O
from bosss_amd import learned
from oracle import mlp as M
from tests.helpers import log_uniform_nu

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _data(n_in, n_out, rows, seed):
    in_p, out_p, _ = M.sizes(n_in, n_out)
    rng = np.random.default_rng(seed)
    X = np.zeros((rows, in_p), np.float32)
    X[:, :n_in] = rng.standard_normal((rows, n_in))
    T = np.zeros((rows, out_p), np.float32)
    T[:, :n_out] = rng.standard_normal((rows, n_out))
    return X, T


@pytest.mark.parametrize("n_in,n_out", [(16, 10), (26, 20)])
def test_forward_matches_oracle(ctx, n_in, n_out):
    _say(f"case mlp forward {n_in}->{n_out}")
    net = B.CellMLP(ctx, n_in, n_out, seed=3)
    p = net.get_params()
    X, _ = _data(n_in, n_out, 1000, 1)
    Y = net.forward(X)
    ref = M.forward_infer(p, X, net.in_p, net.out_p)
    assert Y.shape == ref.shape
    err = np.linalg.norm(Y[:, :n_out] - ref[:, :n_out]) / np.linalg.norm(ref[:, :n_out])
    assert err < 5e-3, err


@pytest.mark.parametrize("n_in,n_out", [(16, 10), (26, 20)])
def test_gradient_matches_oracle(ctx, n_in, n_out):
    _say(f"case mlp gradient {n_in}->{n_out}")
    net = B.CellMLP(ctx, n_in, n_out, seed=4)
    p = net.get_params()
    X, T = _data(n_in, n_out, 777, 2)
    loss, g = net.gradient(X, T)
    lref, gref = M.loss_grad(p, X, T, n_out, net.in_p, net.out_p)
    assert abs(loss - lref) <= 1e-5 * abs(lref)
    assert np.linalg.norm(g - gref) <= 1e-4 * np.linalg.norm(gref)


def test_training_reduces_loss(ctx):
    _say("case mlp training")
    n_in, n_out = 16, 10
    net = B.CellMLP(ctx, n_in, n_out, seed=5)
    rng = np.random.default_rng(0)
    X = np.zeros((20000, 16), np.float32)
    X[:, :n_in] = rng.standard_normal((20000, n_in))
    Wt = rng.standard_normal((n_in, n_out)).astype(np.float32) * 0.5
    T = np.zeros((20000, 16), np.float32)
    T[:, :n_out] = np.tanh(X[:, :n_in] @ Wt)
    dX = B.DeviceArray.from_host(ctx, X.ravel())
    dT = B.DeviceArray.from_host(ctx, T.ravel())
    hist, _ = learned.train(net, dX, dT, 20000, steps=400, batch=4096, lr=3e-3)
    assert np.mean(hist[-20:]) < 0.2 * np.mean(hist[:5])


def test_features_and_targets(ctx):
    _say("case mlp features")
    n = 6
    x = np.linspace(0, 1, n + 1)
    g = B.Grid(ctx, x, x, x)
    nu = log_uniform_nu(g.J, 9)
    op = B.SipOperator(g, 2, nu)
    in_p, out_p = 16, 16
    rng = np.random.default_rng(3)
    r = rng.standard_normal(op.N)
    e = rng.standard_normal(op.N)
    X = B.DeviceArray(ctx, g.J * in_p, np.float32)
    s = B.DeviceArray(ctx, g.J, np.float32)
    T = B.DeviceArray(ctx, g.J * out_p, np.float32)
    B.mlp_features(op, B.DeviceArray.from_host(ctx, r), X, s, in_p)
    B.mlp_targets(op, s, B.DeviceArray.from_host(ctx, e), T, out_p)
    Xh, sh, Th = X.download().reshape(g.J, in_p), s.download(), T.download().reshape(g.J, out_p)
    R = r.reshape(g.J, 10)
    sref = np.linalg.norm(R, axis=1)
    assert np.allclose(sh, sref, rtol=1e-6)
    assert np.allclose(Xh[:, :10], R / sref[:, None], rtol=1e-5, atol=1e-6)
    # neighbour features, cell c = i + j*n + k*n*n
    c = 2 + 3 * n + 1 * n * n
    nb = [c - 1, c + 1, c - n, c + n, c - n * n, c + n * n]
    assert np.allclose(Xh[c, 10:16], np.log10(nu[nb] / nu[c]), atol=1e-5)
    assert Xh[0, 10] == 2.0  # -x boundary
    h2 = (1.0 / n) ** 2
    assert np.allclose(Th[:, :10], e.reshape(g.J, 10) / (sref[:, None] * h2 / nu[:, None]), rtol=1e-5, atol=1e-6)


def test_learned_correction_in_solver(ctx):
    """Train on harvested (residual, error) pairs of a 16^3 k=2 log-uniform problem,
    then solve fresh right-hand sides with and without the network."""
    _say("case learned correction gain")
    n = 16
    x = np.linspace(0, 1, n + 1)
    g = B.Grid(ctx, x, x, x)
    nu = log_uniform_nu(g.J, 21)
    op = B.SipOperator(g, 2, nu)
    mg = B.PMultigrid(op, degrees=[2, 1], min_coarse=64, kcycle_inner=2, prec=B.BH_FP32)
    rng = np.random.default_rng(7)
    rhs = [rng.standard_normal(op.N) for _ in range(4)]
    X, T, ns, in_p, out_p = learned.harvest_pairs(op, mg, rhs, iters_per_rhs=5)
    net = B.CellMLP(ctx, op.Np + 6, op.Np, seed=1)
    hist, t_train = learned.train(net, X, T, ns * g.J, steps=600, batch=8192, lr=1e-3)
    _say(f"  train loss {hist[0]:.3e} -> {np.mean(hist[-20:]):.3e} in {t_train:.2f} s")
    its = {}
    for use in (False, True):
        B.pmg_set_learned(mg, net if use else None)
        fg = B.FlexGMRES(op=op, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=300)
        tot = 0
        for s in range(3):
            b = np.random.default_rng(100 + s).standard_normal(op.N)
            xs = fg.Solve(np.zeros(op.N), b)
            st = fg.stats()
            assert st["converged"]
            r = b - op.Evaluate(1.0, 0.0, np.zeros(op.N), xs, affine=False)
            assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(b)
            tot += st["iterations"]
        its[use] = tot
        if use:
            acc, rej = B.pmg_learned_stats(mg)
            _say(f"  learned correction accepted {acc}, rejected {rej}")
    _say(f"  FGMRES iterations over 3 fresh RHS: plain {its[False]}, with network {its[True]}")
    assert its[True] <= its[False] + 3

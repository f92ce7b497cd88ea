"""GPU check of the allocation-free entry point bh_mlp_gradient_noalloc.

It must give exactly what bh_mlp_gradient gives (same kernels, same block
count, same summation order): the gradient bitwise and the loss bitwise, on two
calls in a row, the second with a smaller batch (fewer blocks, so stale partials
of the first call sit in the shared workspace beyond the ones it reads).
"""
import numpy as np
import pytest

import bosss_amd as B

pytestmark = pytest.mark.gpu


def _batch(in_p, out_p, n_in, n_out, rows, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((rows, in_p), np.float32)
    X[:, :n_in] = rng.standard_normal((rows, n_in))
    T = np.zeros((rows, out_p), np.float32)
    T[:, :n_out] = rng.standard_normal((rows, n_out))
    return X, T


@pytest.mark.parametrize("n_in,n_out", [(16, 10), (26, 20)])
def test_mlp_gradient_noalloc_equals_gradient(ctx, n_in, n_out):
    print(f"case mlp gradient_noalloc {n_in}->{n_out}", flush=True)
    net = B.CellMLP(ctx, n_in, n_out, seed=11)
    g_dev = B.DeviceArray(ctx, net.n_params, np.float32)
    l_dev = B.DeviceArray(ctx, 1, np.float64)
    for rows, seed in ((3000, 1), (97, 2)):
        X, T = _batch(net.in_p, net.out_p, n_in, n_out, rows, seed)
        Xd = B.DeviceArray.from_host(ctx, X.ravel())
        Td = B.DeviceArray.from_host(ctx, T.ravel())
        loss_ref, g_ref = net.gradient(Xd, Td)
        g_dev.upload(np.full(net.n_params, np.nan, np.float32))
        l_dev.upload(np.array([np.nan]))
        net.gradient_noalloc(Xd, Td, g_dev, l_dev)
        ctx.sync()
        g = g_dev.download()
        loss = l_dev.download()[0]
        print(f"  rows {rows}: loss {loss:.9e} (reference {loss_ref:.9e})", flush=True)
        assert np.all(np.isfinite(g))
        assert np.array_equal(g, g_ref), np.max(np.abs(g - g_ref))
        assert loss == loss_ref
        # the network's own gradient buffer feeds Adam exactly as after bh_mlp_gradient
        net.gradient_noalloc(Xd, Td)
        Xd.free()
        Td.free()


def test_mlp_gradient_noalloc_rejects_bad_input(ctx):
    net = B.CellMLP(ctx, 16, 10, seed=1)
    X = B.DeviceArray(ctx, net.in_p, np.float32)
    with pytest.raises(B.BoSSSHipError):
        B._check(B.lib().bh_mlp_gradient_noalloc(net.h, 0, X.ptr, X.ptr, None, None))
    with pytest.raises(B.BoSSSHipError):
        B._check(B.lib().bh_mlp_gradient_noalloc(net.h, 1, None, X.ptr, None, None))

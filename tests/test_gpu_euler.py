"""GPU parity of the explicit DG Euler / HLLC path against the oracle."""
import numpy as np
import pytest

import bosss_amd as B
import oracle as O

pytestmark = pytest.mark.gpu
G = 1.4


def _say(*a):
    print(*a, flush=True)


def _nodes(D, n, seed, jitter=0.3):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(D):
        a = np.linspace(0, 1, n + 1)
        a[1:-1] += jitter * (rng.random(n - 1) - 0.5) / n
        out.append(a)
    return out


def _state(og, Np, nv, seed, amp=0.02):
    """smooth-ish random state: constant mean + small higher modes (density and
    pressure stay positive at every quadrature node for amp <= 0.02)."""
    rng = np.random.default_rng(seed)
    lo, h = og.cells()
    vol = np.prod(h, axis=1)
    D = og.D
    U = np.zeros((og.J, nv, Np))
    u0 = rng.uniform(-0.3, 0.3, D)
    base = np.concatenate([[1.0], u0, [1.0 / (G - 1) + 0.5 * u0 @ u0]])
    U[:, :, 0] = base[None, :] * np.sqrt(vol)[:, None] * (1 + amp * rng.standard_normal((og.J, 1)))
    U[:, :, 1:] = amp * rng.standard_normal((og.J, nv, Np - 1)) * np.sqrt(vol)[:, None, None]
    return U.ravel()


@pytest.mark.parametrize("D,k,nq1", [(3, 2, 4), (3, 1, 3), (2, 2, 4), (2, 3, 5)])
def test_rhs_matches_oracle(ctx, D, k, nq1):
    _say(f"case euler rhs D={D} k={k} nq1={nq1}")
    nodes = _nodes(D, 5 if D == 3 else 7, 1)
    og = O.Grid(*nodes)
    oe = O.Euler(og, k=k, nq1=nq1, gamma=G)
    ge = B.EulerOperator(B.Grid(ctx, *nodes), k=k, nq1=nq1, gamma=G)
    U = _state(og, oe.Np, oe.nv, 2)
    ref = oe.rhs(U)
    assert np.isfinite(ref).all()
    out = ge.rhs(U, np.zeros(ge.N))
    assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)


def test_tvd3_matches_oracle(ctx):
    _say("case euler tvd3")
    nodes = _nodes(3, 4, 3)
    og = O.Grid(*nodes)
    oe = O.Euler(og, k=2, nq1=4, gamma=G)
    ge = B.EulerOperator(B.Grid(ctx, *nodes), k=2, nq1=4, gamma=G)
    U = _state(og, oe.Np, oe.nv, 4)
    ref = oe.tvd3(U, 2e-3, 5)
    assert np.isfinite(ref).all()
    out = ge.tvd3(U.copy(), 2e-3, 5)
    assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)


def test_fp32_rhs(ctx):
    _say("case euler fp32")
    nodes = _nodes(3, 6, 5)
    og = O.Grid(*nodes)
    oe = O.Euler(og, k=2, nq1=4, gamma=G)
    ge = B.EulerOperator(B.Grid(ctx, *nodes), k=2, nq1=4, gamma=G)
    U = _state(og, oe.Np, oe.nv, 6)
    ref = oe.rhs(U)
    assert np.isfinite(ref).all()
    dU = B.DeviceArray.from_host(ctx, U.astype(np.float32))
    dR = B.DeviceArray(ctx, ge.N, np.float32)
    ge.rhs(dU, dR, prec=B.BH_FP32)
    ctx.sync()
    out = dR.download().astype(np.float64)
    assert np.linalg.norm(out - ref) <= 2e-5 * np.linalg.norm(ref)


def test_conservation_full_size(ctx):
    """48^3 periodic, k=2: the cell-mean totals of all conserved variables
    stay constant over 5 TVD3 steps (to round-off), and free-stream is kept."""
    _say("case euler conservation 48^3")
    n = 48
    x = np.linspace(0, 1, n + 1)
    g = B.Grid(ctx, x, x, x)
    ge = B.EulerOperator(g, k=2, nq1=4, gamma=G)
    og = O.Grid(x, x, x)
    U0 = _state(og, ge.Np, 5, 7, amp=0.02)
    dU = B.DeviceArray.from_host(ctx, U0)
    ge.tvd3(dU, 1e-3, 5)
    U1 = dU.download().reshape(-1, 5, ge.Np)
    vol = (1.0 / n) ** 3
    t0 = U0.reshape(-1, 5, ge.Np)[:, :, 0].sum(0) * np.sqrt(vol)
    t1 = U1[:, :, 0].sum(0) * np.sqrt(vol)
    assert np.all(np.abs(t1 - t0) <= 1e-11 * np.abs(t0).max())

"""Shared problem builders for the tests (seeded, reference-style inputs)."""
import numpy as np


def nodes_2d(nx, ny, stretch=False, x0=0.0, x1=1.0, y0=0.0, y1=1.0):
    x = np.linspace(x0, x1, nx + 1)
    y = np.linspace(y0, y1, ny + 1)
    if stretch:  # geometric stretching as in SipHardcodedControl (y-stretch 1.01)
        dy = 1.01 ** np.arange(ny)
        y = y0 + (y1 - y0) * np.concatenate([[0], np.cumsum(dy)]) / dy.sum()
    return x, y


def nodes_3d(nx, ny, nz, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for n in (nx, ny, nz):
        a = np.linspace(0.0, 1.0, n + 1)
        if jitter:
            a[1:-1] += jitter * (rng.random(n - 1) - 0.5) / n
        out.append(a)
    return out


def log_uniform_nu(J, seed, lo=-2.0, hi=2.0):
    """per-cell diffusion 10^U(lo,hi) (log-uniform in [1e-2, 1e2])."""
    return 10.0 ** np.random.default_rng(seed).uniform(lo, hi, J)


def sine_rhs(X, seed, nmodes=8, kmax=6):
    """sum of nmodes products of sines with N(0,1) amplitudes, integer wavenumbers 1..kmax."""
    rng = np.random.default_rng(seed)
    D = X.shape[-1]
    f = np.zeros(X.shape[:-1])
    for _ in range(nmodes):
        a = rng.standard_normal()
        w = rng.integers(1, kmax + 1, D)
        t = np.ones(X.shape[:-1])
        for d in range(D):
            t = t * np.sin(np.pi * w[d] * X[..., d])
        f += a * t
    return f

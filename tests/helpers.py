"""Shared problem builders and checks for the tests (seeded, reference-style inputs)."""
import math

import numpy as np

from oracle import xdg as _OX


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


# ---------------------------------------------------------------- XDG mass matrix
def _moment_coeffs(G, k, j, d, c):
    """Coefficients of x_d - c_d in the orthonormal basis of cell j (least squares on
    the Gauss k+1 tensor nodes, exact since the function lies in the space)."""
    lo, hi = G.box(j)
    X_, _ = _OX._tensor_nodes(G.D, k, lo, hi - lo, 0)
    v, _ = _OX._phys_basis(G.D, k, lo, hi - lo, X_)
    return np.linalg.lstsq(v, X_[:, d] - c[d], rcond=None)[0]


def check_mass_blocks(M, G, asm, k, c, r, tol):
    """Properties the XDG mass matrix must have (M: (N, Np), rows of the full layout):
    the identity on uncut cells; M_A + M_B = I on cut cells; M_s[0, 0] = species volume /
    cell volume (closed forms); symmetric positive semi-definite blocks; and the second
    moments of the ball and of its complement in the unit cube,
    sum_j U_j^T M_s,j U_j with U_j the coefficients of x_d - c_d:
    4 pi r^5 / 15 for the ball (pi r^4 / 4 for the disk in 2D), ((1 - c_d)^3 + c_d^3) / 3
    minus that for the rest."""
    Np = asm.Np
    mom = np.zeros((2, G.D))
    for j in range(G.J):
        blocks = {}
        for s in (0, 1):
            if (G.spmask[j] >> s) & 1:
                B_ = M[asm.dofs(j, s)]
                blocks[s] = B_
                assert np.abs(B_ - B_.T).max() <= 1e-14
                assert np.linalg.eigvalsh(B_).min() >= -1e-13
                assert abs(B_[0, 0] - G.vol[s, j] / G.cellvol[j]) <= tol
                for d in range(G.D):
                    U = _moment_coeffs(G, k, j, d, c)
                    mom[s, d] += U @ B_ @ U
        if len(blocks) == 1:
            assert np.abs(next(iter(blocks.values())) - np.eye(Np)).max() <= 1e-13
        else:
            assert np.abs(blocks[0] + blocks[1] - np.eye(Np)).max() <= tol
    ball = 4 * math.pi * r ** 5 / 15 if G.D == 3 else math.pi * r ** 4 / 4
    cube = np.array([((1 - c[d]) ** 3 + c[d] ** 3) / 3 for d in range(G.D)])
    return mom, ball, cube

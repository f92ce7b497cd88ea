"""CPU pins of the Euler / HLLC oracle (OptimizedHLLC*Flux restated):
consistency, exact resolution of a stationary contact, supersonic upwinding,
rotational invariance, free-stream preservation, discrete conservation and the
convergence of an advected density wave."""
import numpy as np
import pytest

import oracle as O

G = 1.4
MACH = 1.0 / np.sqrt(G)   # MachScaling = 1: standard non-dimensional Euler


def cons(rho, u, p):
    u = np.asarray(u, float)
    return np.concatenate([[rho], rho * u, [p / (G - 1) + 0.5 * rho * u @ u]])


@pytest.mark.parametrize("D", [2, 3])
def test_hllc_consistency(D):
    rng = np.random.default_rng(0)
    for _ in range(20):
        U = cons(rng.uniform(0.5, 2), rng.uniform(-1, 1, D), rng.uniform(0.5, 2))
        n = rng.standard_normal(D)
        n /= np.linalg.norm(n)
        F = O.hllc(D, G, MACH, U, U, n)
        ref = O.euler_flux(D, G, MACH, U) @ n
        assert np.allclose(F, ref, rtol=1e-13, atol=1e-13)


def test_hllc_stationary_contact_exact():
    D = 3
    p = 0.7
    UL = cons(1.0, [0, 0, 0], p)
    UR = cons(0.125, [0, 0, 0], p)
    n = np.array([1.0, 0, 0])
    F = O.hllc(D, G, MACH, UL, UR, n)
    assert np.allclose(F, [0, p, 0, 0, 0], atol=1e-14)


def test_hllc_supersonic_upwind():
    D = 2
    UL = cons(1.0, [3.0, 0.2], 1.0)   # |u| > a = sqrt(1.4) on both sides
    UR = cons(0.8, [2.9, -0.1], 0.9)
    n = np.array([1.0, 0.0])
    assert np.allclose(O.hllc(D, G, MACH, UL, UR, n), O.euler_flux(D, G, MACH, UL) @ n, rtol=1e-13)
    assert np.allclose(O.hllc(D, G, MACH, UR, UL, -n), O.euler_flux(D, G, MACH, UL) @ -n, rtol=1e-13)


def test_hllc_rotational_invariance():
    D = 3
    rng = np.random.default_rng(1)
    Q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    for _ in range(10):
        UL = cons(rng.uniform(0.5, 2), rng.uniform(-1, 1, D), rng.uniform(0.5, 2))
        UR = cons(rng.uniform(0.5, 2), rng.uniform(-1, 1, D), rng.uniform(0.5, 2))
        n = np.array([1.0, 0, 0])
        T = np.eye(5)
        T[1:4, 1:4] = Q
        F = O.hllc(D, G, MACH, T @ UL, T @ UR, Q @ n)
        assert np.allclose(F, T @ O.hllc(D, G, MACH, UL, UR, n), rtol=1e-12, atol=1e-12)


def _grid3(n, jitter=0.0, seed=0):
    rng = np.random.default_rng(seed)
    nodes = []
    for _ in range(3):
        a = np.linspace(0, 1, n + 1)
        a[1:-1] += jitter * (rng.random(n - 1) - 0.5) / n
        nodes.append(a)
    return O.Grid(*nodes)


def test_free_stream_preserved():
    g = _grid3(4, jitter=0.4)
    e = O.Euler(g, k=2, nq1=4, gamma=G, mach=MACH)
    U = np.zeros((g.J, e.nv, e.Np))
    lo, h = g.cells()
    vol = np.prod(h, axis=1)
    s = cons(1.3, [0.2, -0.1, 0.3], 0.9)
    U[:, :, 0] = s[None, :] * np.sqrt(vol)[:, None]   # constant state in the orthonormal basis
    R = e.rhs(U.ravel())
    assert np.abs(R).max() < 1e-12


def test_discrete_conservation():
    g = _grid3(4, jitter=0.4, seed=2)
    e = O.Euler(g, k=2, nq1=4, gamma=G, mach=MACH)
    rng = np.random.default_rng(3)
    U = np.zeros((g.J, e.nv, e.Np))
    lo, h = g.cells()
    vol = np.prod(h, axis=1)
    U[:, :, 0] = cons(1.0, [0.1, 0.0, 0.0], 1.0)[None, :] * np.sqrt(vol)[:, None]
    U[:, :, 1:] = 0.01 * rng.standard_normal((g.J, e.nv, e.Np - 1)) * np.sqrt(vol)[:, None, None]
    R = e.rhs(U.ravel()).reshape(g.J, e.nv, e.Np)
    total = (R[:, :, 0] * np.sqrt(vol)[:, None]).sum(0)   # d/dt of the integrals of the conserved variables
    assert np.abs(total).max() < 1e-12


def test_density_wave_convergence():
    """rho = 1 + 0.2 sin(2 pi (x - t)), u = (1, 0, 0), p = 1: exact translation."""
    errs = []
    for n in (4, 8):
        x = np.linspace(0, 1, n + 1)
        g = O.Grid(x, np.linspace(0, 1, 2), np.linspace(0, 1, 2))
        e = O.Euler(g, k=2, nq1=4, gamma=G, mach=MACH)
        sip = O.Sip(g, 2, 1.0, nq1=6)
        X, W = sip.vol_nodes()

        def state(t):
            rho = 1 + 0.2 * np.sin(2 * np.pi * (X[..., 0] - t))
            Uq = [rho, rho * 1.0, 0 * rho, 0 * rho, 1.0 / (G - 1) + 0.5 * rho]
            return np.stack([sip.project(u).reshape(g.J, -1) for u in Uq], axis=1).ravel()

        U0 = state(0.0)
        T, steps = 0.1, 40 * n // 4
        U = e.tvd3(U0, T / steps, steps)
        errs.append(np.linalg.norm(U - state(T)) * np.sqrt(1.0 / n))
    rate = np.log2(errs[0] / errs[1])
    assert errs[1] < 5e-4 and rate > 2.5, (errs, rate)

"""Newton-Krylov solve of the backward-Euler step of the DG Euler/HLLC operator
(bh_euler_implicit_step, after AS/Newton.cs with its matrix-free GMRES option and
line search) against an independent solve of the same nonlinear system:
scipy.optimize.newton_krylov on F(U) = U - Un - dt R_oracle(U), with the oracle's
right-hand side.  The root of F near Un is unique, so both solvers must meet:
periodic 6^3 grid, k = 2, the seeded smooth Mach-0.2 state, dt 10x the explicit
step of the benchmark.  Also: the device solution satisfies the oracle's F to the
Newton tolerance, and the cell-mean totals of U equal those of Un (R conserves
them, so U - Un = dt R(U) has zero totals).
"""
import numpy as np
import pytest
import scipy.optimize as sopt

import bench
import bosss_amd as B
import oracle as O

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def test_implicit_euler_step_matches_newton_krylov_of_the_oracle(ctx):
    _say("case newton-krylov backward euler 6^3 k=2")
    n, k = 6, 2
    x = np.linspace(0.0, 1.0, n + 1)
    oe = O.Euler(O.Grid(x, x, x), k=k, nq1=4, gamma=1.4)
    ge = B.EulerOperator(B.Grid(ctx, x, x, x), k=k, nq1=4, gamma=1.4)
    Un = bench.euler_initial_state(n, k, seed=4)
    dt = 0.5 / n
    U, info = ge.implicit_step(Un, dt, ConvCrit=1e-11, MaxIter=20)
    _say(f"  device: {info}")
    assert info["res"] <= 1e-11 * info["res0"] + 1e-11
    assert 1 <= info["newton"] <= 12

    def F(V):
        return V - Un - dt * oe.rhs(V)

    Uo = sopt.newton_krylov(F, Un.copy(), f_tol=1e-12, method="lgmres")
    _say(f"  |U_device - U_scipy| / |U| = {rel(U, Uo):.2e}, |F_oracle(U_device)| = {np.linalg.norm(F(U)):.2e}")
    assert rel(U, Uo) <= 1e-9
    assert np.linalg.norm(F(U)) <= 1e-9 * np.linalg.norm(Un)
    t0 = Un.reshape(-1, 5, ge.Np)[:, :, 0].sum(0)
    t1 = U.reshape(-1, 5, ge.Np)[:, :, 0].sum(0)
    assert np.abs(t1 - t0).max() <= 1e-10 * np.abs(t0).max()
    # the step moved the state by more than the tolerance (a real nonlinear solve)
    assert rel(U, Un) > 1e-4

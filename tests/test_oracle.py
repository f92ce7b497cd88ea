"""CPU checks of the oracle against what the reference holds and what the
mathematics fixes (no GPU).  Each test names the reference file it pins.
"""
import json
import os

import numpy as np
import pytest

import oracle as O
from tests.helpers import log_uniform_nu, nodes_2d, nodes_3d

FIX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ref_tables.json")))


def _eval_table(polys, xi):
    out = np.zeros((xi.shape[0], len(polys)))
    for n, terms in enumerate(polys):
        for c, e in terms:
            out[:, n] += c * np.prod(xi ** np.array(e), axis=1)
    return out


@pytest.mark.parametrize("D,key,tol", [(2, "square", 2e-15), (3, "cube", 1e-14)])
def test_basis_matches_reference_tables(D, key, tol):
    """Orthonormal polynomials and mode order == F/Square.cs:192-6495 / F/Cube_polynomials.cs."""
    polys = FIX[key]
    k = 4
    assert len(polys) == O.basis_count(D, k)
    xi = np.random.default_rng(0).uniform(-1, 1, (64, D))
    val, _ = O.basis_eval(D, k, xi)
    ref = _eval_table(polys, xi)
    # the 3D tables carry 15 significant digits
    assert np.abs(val - ref).max() < tol * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("D", [2, 3])
def test_basis_gradients_finite_difference(D):
    xi = np.random.default_rng(1).uniform(-0.9, 0.9, (16, D))
    k = 4
    _, g = O.basis_eval(D, k, xi)
    eps = 1e-6
    for d in range(D):
        e = np.zeros(D)
        e[d] = eps
        vp, _ = O.basis_eval(D, k, xi + e)
        vm, _ = O.basis_eval(D, k, xi - e)
        assert np.abs((vp - vm) / (2 * eps) - g[:, :, d]).max() < 1e-7


@pytest.mark.parametrize("D,k", [(2, 4), (3, 3)])
def test_basis_orthonormal(D, k):
    x, w = O.gauss(k + 2)
    grids = np.meshgrid(*([x] * D), indexing="ij")
    X = np.stack([a.ravel() for a in grids], -1)
    W = np.prod(np.stack(np.meshgrid(*([w] * D), indexing="ij"), -1).reshape(-1, D), axis=1)
    val, _ = O.basis_eval(D, k, X)
    M = (val * W[:, None]).T @ val
    assert np.abs(M - np.eye(M.shape[0])).max() < 1e-13


@pytest.mark.parametrize("n", range(1, 13))
def test_gauss_matches_reference_rules(n):
    """F/LineQuadRules/Line{n}.txt (15-digit tables) and exactness to degree 2n-1."""
    r = FIX["line_rules"][str(n)]
    x, w = O.gauss(n)
    assert r["order"] == 2 * n - 1
    assert np.abs(x - np.array(r["nodes"])).max() < 1e-14
    assert np.abs(w - np.array(r["weights"])).max() < 1e-14
    for p in range(2 * n):
        exact = 0.0 if p % 2 else 2.0 / (p + 1)
        assert abs(np.sum(w * x ** p) - exact) < 1e-14


def test_grid_numbering_2d():
    """GlobalID = i*nY + j (BoSSS.Foundation.Grid/Grid2D.cs:586)."""
    x, y = nodes_2d(5, 3)
    g = O.Grid(x, y)
    lo, h = g.cells()
    for i in range(5):
        for j in range(3):
            c = i * 3 + j
            assert lo[c, 0] == x[i] and lo[c, 1] == y[j]


def test_grid_numbering_3d():
    """GlobalID = i + j*nX + k*nX*nY (BoSSS.Foundation.Grid/Grid3D.cs:261)."""
    xs = nodes_3d(4, 3, 2)
    g = O.Grid(*xs)
    lo, h = g.cells()
    for k in range(2):
        for j in range(3):
            for i in range(4):
                c = i + j * 4 + k * 12
                assert (lo[c] == [xs[0][i], xs[1][j], xs[2][k]]).all()


def test_cj_closed_form():
    """cj = (surface + boundary area)/2/volume (F/GridData_CellData.cs:666-700)."""
    h = 0.25
    g = O.Grid(np.linspace(0, 1, 5), np.linspace(0, 1, 5))
    cj = g.cj()
    lo, _ = g.cells()
    for c in range(g.J):
        nb = sum([lo[c, 0] == 0, lo[c, 0] == 0.75, lo[c, 1] == 0, lo[c, 1] == 0.75])
        assert abs(cj[c] - (4 * h + nb * h) / 2 / h ** 2) < 1e-12


def test_penalty_closed_form():
    """mode-0 diagonal entry = sum over faces of -mu*nu*area/vol
    (penalty base*max((p+1)(p+D)/D,(p+1)^2)*max(cj), boundary 2x: ipLaplace.cs:100-144,224)."""
    k, base, nu = 2, 1.3, 0.7
    n = 6
    h = 1.0 / n
    g = O.Grid(np.linspace(0, 1, n + 1), np.linspace(0, 1, n + 1))
    op = O.Sip(g, k, nu, penalty_base=base)
    A, _ = op.assemble()
    rp, col, val = A.export()
    deg = (k + 1) ** 2
    cj_int, cj_bnd = 2 / h, 2.5 / h
    # centre cell (all neighbours interior): 4 faces with mu = base*deg*2/h
    c = 3 * n + 3
    p = [q for q in range(rp[c], rp[c + 1]) if col[q] == c][0]
    assert abs(val[p][0, 0] - (-4 * base * deg * cj_int * nu * h / h ** 2)) < 1e-9 * abs(val[p][0, 0])
    # cell (0, 3): one Dirichlet face (2*mu_b), 3 interior faces (one neighbour is a boundary cell)
    c = 0 * n + 3
    p = [q for q in range(rp[c], rp[c + 1]) if col[q] == c][0]
    mu_b = base * deg * cj_bnd
    expect = -(2 * mu_b + 3 * base * deg * cj_bnd) * nu * h / h ** 2
    assert abs(val[p][0, 0] - expect) < 1e-9 * abs(expect)


def _random_op(D, k, seed, bc=None):
    if D == 2:
        g = O.Grid(*nodes_2d(5, 4, stretch=True))
    else:
        g = O.Grid(*nodes_3d(3, 4, 3, jitter=0.3, seed=seed))
    nu = log_uniform_nu(g.J, seed)
    return g, O.Sip(g, k, nu, bc=bc)


@pytest.mark.parametrize("D,k", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_evaluate_equals_assembled(D, k):
    """matrix-free evaluate == ComputeMatrix * u + affine (two oracle paths)."""
    rng = np.random.default_rng(7)
    bc = [0, 1] * D
    g, op = _random_op(D, k, 3, bc)
    op.set_bnd_values(rng.standard_normal(op.bnd_nodes().shape[0]))
    op.set_rhs(rng.standard_normal(op.N))
    u = rng.standard_normal(op.N)
    y = op.evaluate(u)
    A, aff = op.assemble()
    y2 = A.spmv(1.0, u, 0.0, np.zeros(op.N)) + aff
    assert np.linalg.norm(y - y2) <= 1e-13 * np.linalg.norm(y)
    # alpha / beta semantics of Evaluate
    y0 = rng.standard_normal(op.N)
    y3 = op.evaluate(u, alpha=-2.0, beta=0.5, y=y0.copy())
    assert np.allclose(y3, 0.5 * y0 - 2.0 * y, rtol=0, atol=1e-12 * np.abs(y).max())


@pytest.mark.parametrize("D,k", [(2, 2), (3, 2)])
def test_operator_symmetric_negative_definite(D, k):
    g, op = _random_op(D, k, 5)
    A, _ = op.assemble()
    M = A.dense()
    assert np.abs(M - M.T).max() <= 1e-13 * np.abs(M).max()
    assert np.linalg.eigvalsh(0.5 * (M + M.T)).max() < 0


@pytest.mark.parametrize("D,k,bc", [(2, 2, [0, 0, 0, 0]), (2, 3, [0, 1, 1, 0]), (3, 2, [0, 0, 0, 0, 0, 0]),
                                    (3, 3, [1, 0, 0, 1, 0, 0])])
def test_polynomial_consistency(D, k, bc):
    """DerivativeTest principle (L4-application/DerivativeTest/DerivativeTestMain.cs:1060-1150):
    the SIP operator of a degree-<=k polynomial with matching Dirichlet/Neumann data
    equals the projection of nu*Laplace(u)."""
    nu = 2.5
    if D == 2:
        g = O.Grid(*nodes_2d(4, 3, stretch=True, y0=-1.0, y1=0.5))
        f = lambda X: X[..., 0] ** 2 + 3 * X[..., 0] * X[..., 1] - X[..., 1] + 2 - 0.5 * X[..., 1] ** 2 + 0.1 * X[..., 0] ** 2 * X[..., 1] * (k >= 3)
        grad = lambda X: np.stack([2 * X[..., 0] + 3 * X[..., 1] + 0.2 * X[..., 0] * X[..., 1] * (k >= 3),
                                   3 * X[..., 0] - 1 - X[..., 1] + 0.1 * X[..., 0] ** 2 * (k >= 3)], -1)
        lap = lambda X: 2 - 1 + 0.2 * X[..., 1] * (k >= 3) + 0 * X[..., 0]
    else:
        g = O.Grid(*nodes_3d(3, 2, 3, jitter=0.4, seed=1))
        f = lambda X: X[..., 0] ** 2 + 3 * X[..., 0] * X[..., 1] - X[..., 1] + 2 - X[..., 2] ** 2 + X[..., 1] * X[..., 2]
        grad = lambda X: np.stack([2 * X[..., 0] + 3 * X[..., 1], 3 * X[..., 0] - 1 + X[..., 2],
                                   -2 * X[..., 2] + X[..., 1]], -1)
        lap = lambda X: 0 * X[..., 0]
    op = O.Sip(g, k, nu, bc=bc)
    X, _ = op.vol_nodes()
    uh = op.project(f(X))
    Xb = op.bnd_nodes()
    # per boundary node: Dirichlet value or outward normal derivative
    gvals = np.zeros(Xb.shape[0])
    lo, hh = g.cells()
    ein, eout, edir, eside = g.edges()
    nqf = (k + 1) ** (D - 1)
    slot = 0
    for e in range(g.E):
        if eout[e] >= 0:
            continue
        side = eside[e]
        xs = Xb[slot:slot + nqf]
        if bc[side] == 0:
            gvals[slot:slot + nqf] = f(xs)
        else:
            sgn = 1.0 if side % 2 else -1.0
            gvals[slot:slot + nqf] = sgn * grad(xs)[:, side // 2]
        slot += nqf
    op.set_bnd_values(gvals)
    y = op.evaluate(uh)
    ref = op.project(nu * lap(X))
    assert np.abs(y - ref).max() < 1e-9 * max(1.0, np.abs(y).max())


def test_chunk_size_independence():
    """bulksize_limit sweep of DerivativeTest (DerivativeTestMain.cs:65-105)."""
    g, op = _random_op(3, 2, 11)
    u = np.random.default_rng(2).standard_normal(op.N)
    ys = []
    for ch in (1, 2, 7, 20, 10 ** 7):
        op.set_chunk(ch)
        ys.append(op.evaluate(u))
    for y in ys[1:]:
        assert np.abs(y - ys[0]).max() <= 1e-14 * np.abs(ys[0]).max()


def test_cartesian_poisson_exact():
    """ipPoisson TestCartesian (L4-application/ipPoisson/SipHardcodedControl.cs:206;
    Tests/TestProgram.cs): u = 0.5x^2 - 10x on [0,10]x[-1,1], 32x16 cells with
    y-stretch 1.01, k=2, Dirichlet at x=0, Neumann elsewhere; L2 error <= 5e-9."""
    k = 2
    x = np.linspace(0, 10, 33)
    _, y = nodes_2d(1, 16, stretch=True, y0=-1.0, y1=1.0)
    g = O.Grid(x, y)
    bc = [0, 1, 1, 1]
    op = O.Sip(g, k, 1.0, bc=bc)
    X, W = op.vol_nodes()
    uex = lambda P: 0.5 * P[..., 0] ** 2 - 10 * P[..., 0]
    op.set_rhs(op.project(np.ones(X.shape[:2])))
    Xb = op.bnd_nodes()
    gv = np.zeros(Xb.shape[0])
    ein, eout, edir, eside = g.edges()
    nqf = k + 1
    slot = 0
    for e in range(g.E):
        if eout[e] >= 0:
            continue
        s = eside[e]
        xs = Xb[slot:slot + nqf]
        gv[slot:slot + nqf] = uex(xs) if bc[s] == 0 else (xs[:, 0] - 10 if s == 1 else 0.0)
        slot += nqf
    op.set_bnd_values(gv)
    A, aff = op.assemble()
    u = np.linalg.solve(A.dense(), -aff)
    err = op.project(uex(X)) - u
    assert np.sqrt(np.sum(err ** 2)) < 5e-9


def test_bjacobi_and_gmres_against_dense():
    """BlockJacobi sweep rule (AS/BlockJacobi.cs:196-255) and SoftGMRES
    (AS/SoftGMRES.cs:227-455) against numpy."""
    g, op = _random_op(2, 2, 13)
    A, _ = op.assemble()
    M = A.dense()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(op.N)
    # block Jacobi: NoOfIterations=1 -> 2 sweeps of x <- x + w Dinv (b - A x)
    Np = op.Np
    Dinv = np.zeros_like(M)
    for j in range(g.J):
        s = slice(j * Np, (j + 1) * Np)
        Dinv[s, s] = np.linalg.inv(M[s, s])
    x = np.zeros(op.N)
    for _ in range(2):
        x = x + 0.7 * Dinv @ (b - M @ x)
    xo = A.bjacobi(np.zeros(op.N), b, omega=0.7, iters=1)
    assert np.abs(xo - x).max() < 1e-12 * np.abs(x).max()
    xg, st = A.gmres(b, precond=1, restart=30, tol=1e-12)
    xe = np.linalg.solve(M, b)
    assert st["converged"]
    # termination compares the left-preconditioned estimate |s_i| = ||Dinv r|| with
    # R0*tol + tol, R0 = ||b - A x0|| (AS/SoftGMRES.cs:300-330, AS/IterativeSolverConfig.cs:48-58)
    Dr = Dinv @ (b - M @ xg)
    assert np.linalg.norm(Dr) < 1e-12 * np.linalg.norm(b) + 1e-12 + 1e-14 * np.linalg.norm(Dinv @ b)
    assert np.linalg.norm(xg - xe) < 1e-6 * np.linalg.norm(xe)


def _agg_parent_3d(n):
    """parent aggregate of each cell of an n[0] x n[1] x n[2] grid (cell i + j nx + k nx ny)."""
    nc = [(m + 1) // 2 for m in n]
    par = []
    for k in range(n[2]):
        for j in range(n[1]):
            for i in range(n[0]):
                par.append(i // 2 + (j // 2) * nc[0] + (k // 2) * nc[0] * nc[1])
    return np.array(par)


def test_schwarz_smoother_against_dense():
    """Additive Schwarz over the multigrid aggregates, no overlap (AS/Schwarz.cs with
    MG-aggregate blocks): x = omega sum_I R_I^T (R_I A R_I^T)^-1 R_I r on the k = 1
    level, against numpy on the dense level matrix (odd grid: partial aggregates)."""
    nodes = nodes_3d(5, 4, 3, jitter=0.3, seed=8)
    g = O.Grid(*nodes)
    op = O.Sip(g, 1, log_uniform_nu(g.J, 8))
    A, _ = op.assemble()
    P = O.Pmg(A, g, [1], min_coarse=4, omega=0.7, pre=1, post=1)
    P.set_schwarz(True)
    M = A.dense()
    par = _agg_parent_3d([5, 4, 3])
    r = np.random.default_rng(4).standard_normal(op.N)
    x = np.zeros(op.N)
    Np = op.Np
    for I in np.unique(par):
        dofs = np.concatenate([np.arange(c * Np, (c + 1) * Np) for c in np.flatnonzero(par == I)])
        x[dofs] = 0.7 * np.linalg.solve(M[np.ix_(dofs, dofs)], r[dofs])
    xo = P.smooth(0, r)
    assert np.abs(xo - x).max() < 1e-12 * np.abs(x).max()
    P.set_schwarz(False)   # back to block-Jacobi: cell blocks
    xb = np.zeros(op.N)
    for c in range(g.J):
        s = slice(c * Np, (c + 1) * Np)
        xb[s] = 0.7 * np.linalg.solve(M[s, s], r[s])
    assert np.abs(P.smooth(0, r) - xb).max() < 1e-12 * np.abs(xb).max()


def _overlap_blocks(n, par, ov):
    """cells of each aggregate plus every cell within ov face steps of it (grid graph
    distance; cell ids as the grid numbers them), by breadth-first search."""
    D = len(n)
    if D == 2:
        ijk = [(c // n[1], c % n[1]) for c in range(n[0] * n[1])]
        cid = {t: c for c, t in enumerate(ijk)}
    else:
        ijk = [(c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1])) for c in range(n[0] * n[1] * n[2])]
        cid = {t: c for c, t in enumerate(ijk)}
    out = []
    for I in np.unique(par):
        blk = set(np.flatnonzero(par == I).tolist())
        front = set(blk)
        for _ in range(ov):
            nxt = set()
            for c in front:
                for d in range(D):
                    for sg in (-1, 1):
                        t = list(ijk[c])
                        t[d] += sg
                        if tuple(t) in cid and cid[tuple(t)] not in blk:
                            nxt.add(cid[tuple(t)])
            blk |= nxt
            front = nxt
        out.append(sorted(blk))
    return out


@pytest.mark.parametrize("scaling", [False, True])
@pytest.mark.parametrize("D,ov", [(3, 1), (2, 1), (2, 2)])
def test_schwarz_overlap_against_dense(D, ov, scaling):
    """Additive Schwarz with overlap (AS/Schwarz.cs:683-730 block growth, :1056-1085 the
    additive sum, :791-815 the optional overlap scaling): x = omega sum_I R_I^T
    (R_I A R_I^T)^-1 R_I r with each aggregate grown by ov layers of face neighbours
    (with scaling, divided per DOF by the number of blocks holding it), against numpy on
    the dense level matrix; blocks found by breadth-first search in numpy."""
    if D == 3:
        n = [5, 4, 3]
        g = O.Grid(*nodes_3d(*n, jitter=0.3, seed=8))
        par = _agg_parent_3d(n)
    else:
        n = [7, 6]
        g = O.Grid(*nodes_2d(*n))
        par = np.array([(i // 2) * 3 + j // 2 for i in range(7) for j in range(6)])
    op = O.Sip(g, 1, log_uniform_nu(g.J, 8))
    A, _ = op.assemble()
    P = O.Pmg(A, g, [1], min_coarse=4, omega=0.7, pre=1, post=1)
    P.set_schwarz(True, overlap=ov, scaling=scaling)
    M = A.dense()
    Np = op.Np
    r = np.random.default_rng(4).standard_normal(op.N)
    x = np.zeros(op.N)
    blocks = _overlap_blocks(n, par, ov)
    assert max(len(b) for b in blocks) > 2 ** D   # the blocks overlap
    mult = np.zeros(op.N)
    for cells in blocks:
        dofs = np.concatenate([np.arange(c * Np, (c + 1) * Np) for c in cells])
        x[dofs] += 0.7 * np.linalg.solve(M[np.ix_(dofs, dofs)], r[dofs])
        mult[dofs] += 1
    assert mult.min() >= 1 and mult.max() > 1
    if scaling:
        x /= mult
    xo = P.smooth(0, r)
    assert np.abs(xo - x).max() < 1e-12 * np.abs(x).max()


def test_schwarz_single_aggregate_is_exact_solve():
    """A 2x2x2 grid is one aggregate: with omega = 1 the Schwarz smoother is A^-1, so one
    pre-smoothing sweep already solves and the V-cycle returns A^-1 b."""
    x = np.linspace(0.0, 1.0, 3)
    g = O.Grid(x, x, x)
    op = O.Sip(g, 1, log_uniform_nu(g.J, 2))
    A, _ = op.assemble()
    P = O.Pmg(A, g, [1], min_coarse=1, omega=1.0, pre=1, post=0)
    P.set_schwarz(True)
    b = np.random.default_rng(1).standard_normal(op.N)
    xe = np.linalg.solve(A.dense(), b)
    assert np.abs(P.vcycle(b) - xe).max() < 1e-10 * np.abs(xe).max()


@pytest.mark.parametrize("D,k", [(2, 2), (3, 2), (3, 3)])
def test_level_pmg_against_dense(D, k):
    """LevelPmg with diagonal high-order blocks (AS/LevelPmg.cs:232-330, :397-470,
    :561-600): the low-order modes (degree <= 1) solved exactly, then every cell's
    high-order modes with the inverse of their part of its diagonal block, against
    numpy on the dense matrix; a two-level hierarchy, so the low-order solve is the
    oracle's dense coarsest solve."""
    if D == 2:
        g = O.Grid(*nodes_2d(5, 4))
    else:
        g = O.Grid(*nodes_3d(4, 3, 3, jitter=0.3, seed=2))
    op = O.Sip(g, k, log_uniform_nu(g.J, 3))
    A, _ = op.assemble()
    P = O.Pmg(A, g, [k, 1], min_coarse=10 ** 6, omega=0.7, pre=1, post=1)
    assert len(P.sizes) == 2
    P.set_level_pmg(True)
    M = A.dense()
    Np = op.Np
    Npl = D + 1
    lo = np.concatenate([np.arange(c * Np, c * Np + Npl) for c in range(g.J)])
    b = np.random.default_rng(11).standard_normal(op.N)
    x = np.zeros(op.N)
    x[lo] = np.linalg.solve(M[np.ix_(lo, lo)], b[lo])
    r = b - M @ x
    for c in range(g.J):
        hi = np.arange(c * Np + Npl, (c + 1) * Np)
        x[hi] += np.linalg.solve(M[np.ix_(hi, hi)], r[hi])
    xo = P.vcycle(b)
    assert np.abs(xo - x).max() < 1e-11 * np.abs(x).max()
    P.set_level_pmg(False)
    assert np.abs(P.vcycle(b) - xo).max() > 1e-3 * np.abs(x).max()   # the plain cycle differs

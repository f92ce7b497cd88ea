"""GPU parity of the XDG two-phase Poisson path against the oracle.

* full (cell, species) matrix assembled on the device vs the oracle's
  restatement of XLaplace_Bulk / XLaplace_Interface on the same element rules
  and length scales (rules and length scales are pinned by the host tests);
* agglomerated matrix Q^T A Q and right-hand side vs the oracle's;
* variable-block SpMV (FP64 / FP32) vs scipy;
* a piecewise parabola u = a_s |x - c|^2 + c_s with flux continuity across the
  sphere, which the k = 2 space holds exactly (the analogue of XdgPoisson3's
  ParabolaTest): the FGMRES + two-level FP32 preconditioner solution equals the
  direct sparse solve of the oracle's system and the exact solution;
* the solver converges at 1:1000 diffusion contrast.
"""
import math

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import bosss_amd as B
from oracle import xdg as OX
from tests.helpers import check_mass_blocks

pytestmark = pytest.mark.gpu

NU = (1.0, 1000.0)


def _say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _setup(ctx, D=3, n=10, c=(0.5, 0.47, 0.52), r=0.27, q=6, nu=NU, agg=0.1, k=2, graded=False):
    c = list(c)[:D]
    x = np.linspace(0, 1, n + 1)
    if graded:   # cells of growing width: neighbours across a face differ in size
        x = x ** 1.3
    g = B.Grid(ctx, *([x] * D))
    X = B.XdgPoisson(g, k, c, r, nu[0], nu[1], agg_threshold=agg, quad_order=q)
    G = OX.Geometry([x] * D, c, r)
    e = X.elements()
    L = X.cells()["len"]
    asm = OX.XdgAssembly(G, k, nu, 1.3, L, e["ints"], e["nx"], e["nw"], e["nn"])
    tgt = -np.ones((2, G.J), np.int64)
    for (s_, t_, sp) in X.agglomeration():
        tgt[sp, s_] = t_
    return X, G, asm, tgt, e, L


def _parabola(D, c, r, nu, f=1.0):
    a, b = f / (2 * D * nu[0]), f / (2 * D * nu[1])
    cA, cB = (b - a) * r * r, 0.0

    def u(xx, s):
        d2 = ((xx[:, :D] - np.asarray(c)[:D]) ** 2).sum(1)
        return a * d2 + cA if s == 0 else b * d2 + cB
    return u


@pytest.mark.parametrize("D,n", [(3, 8), (2, 16)])
def test_full_matrix_matches_oracle(ctx, D, n):
    _say(f"case xdg full matrix D={D} n={n}")
    X, G, asm, tgt, e, L = _setup(ctx, D=D, n=n, r=0.27 if D == 3 else 0.31)
    A = X.scipy_matrix(0)
    assert A.shape == asm.A.shape
    d = (A - asm.A)
    assert spla.norm(d) <= 1e-12 * spla.norm(asm.A)
    # symmetric (SIP) up to round-off
    assert spla.norm(A - A.T) <= 1e-12 * spla.norm(A)


@pytest.mark.parametrize("D,n,k,graded", [(3, 8, 2, False), (2, 16, 2, False), (3, 8, 1, False), (2, 16, 1, False),
                                          (2, 12, 3, False), (3, 8, 2, True), (2, 14, 3, True)])
def test_matrix_free_evaluate(ctx, D, n, k, graded):
    """The evaluate over the element list (uncut items in closed modal form, cut items
    from their kept element matrices, one launch per colour group) equals the oracle's
    assembled operator (quadrature on the same rules) applied to u, for degrees 1 to 3
    and on graded grids (neighbours of different widths), and the Evaluate(alpha, beta)
    semantics hold."""
    _say(f"case xdg evaluate D={D} n={n} k={k} graded={graded}")
    X, G, asm, tgt, e, L = _setup(ctx, D=D, n=n, r=0.27 if D == 3 else 0.31, k=k, graded=graded)
    rng = np.random.default_rng(21)
    u = rng.standard_normal(X.N_full)
    y0 = rng.standard_normal(X.N_full)
    out = X.Evaluate(1.0, 0.0, np.zeros(X.N_full), u)
    assert rel(out, asm.A @ u) <= 1e-12
    out2 = X.Evaluate(-0.7, 1.3, y0.copy(), u)
    assert rel(out2, 1.3 * y0 - 0.7 * (asm.A @ u)) <= 1e-12


def test_agglomerated_matrix_and_rhs(ctx):
    _say("case xdg agglomerated matrix + rhs")
    X, G, asm, tgt, e, L = _setup(ctx, D=3, n=10)
    assert X.npairs > 0
    Q, roff = asm.injection(tgt)
    Ar = (Q.T @ asm.A @ Q).tocsr()
    A = X.scipy_matrix(1)
    assert A.shape == Ar.shape
    assert spla.norm(A - Ar) <= 1e-12 * spla.norm(Ar)
    c = [0.5, 0.47, 0.52]
    u = _parabola(3, c, 0.27, NU)
    xb = X.bnd_nodes()
    X.set_rhs([1.0, 1.0], u(xb, 1))
    b_or = asm.rhs(e["ints"], e["nx"], e["nw"], [1.0, 1.0], NU, 1.3, L, lambda xx: u(xx, 1))
    assert rel(X.rhs(0), b_or) <= 1e-12
    assert rel(X.rhs(1), Q.T @ b_or) <= 1e-12


def test_spmv(ctx):
    _say("case xdg spmv fp64/fp32")
    X, G, asm, tgt, e, L = _setup(ctx, D=3, n=10)
    rng = np.random.default_rng(5)
    for which in (0, 1):
        A = X.scipy_matrix(which)
        xv = rng.standard_normal(A.shape[0])
        y0 = rng.standard_normal(A.shape[0])
        y = X.SpMV(1.7, xv, 0.3, y0.copy(), which=which)
        ref = 1.7 * (A @ xv) + 0.3 * y0
        assert rel(y, ref) <= 1e-13
        dx = B.DeviceArray.from_host(ctx, xv.astype(np.float32))
        dy = B.DeviceArray(ctx, A.shape[0], np.float32)
        X.SpMV(1.0, dx, 0.0, dy, which=which, prec=B.BH_FP32)
        ctx.sync()
        assert rel(dy.download().astype(np.float64), A @ xv) <= 1e-5


def _solve(X, tol, prec=B.BH_FP32, restart=30, maxit=2000):
    P = B.XdgPrecond(X, coarse_degree=1, prec=prec)
    S = B.XdgFlexGMRES(X, P, MaxKrylovDim=restart, ConvergenceCriterion=tol, MaxSolverIterations=maxit)
    b = X.rhs(1)
    u = np.zeros(X.N)
    S.Solve(u, b)
    return u, b, S.stats()


def test_parabola_exact(ctx):
    """k = 2 holds the piecewise parabola exactly: solution error at the rule
    nodes is set by the cut-cell quadrature only (q = 8 here)."""
    _say("case xdg parabola")
    c, r = [0.5, 0.47, 0.52], 0.27
    X, G, asm, tgt, e, L = _setup(ctx, D=3, n=10, c=c, r=r, q=8)
    u = _parabola(3, c, r, NU)
    X.set_rhs([1.0, 1.0], u(X.bnd_nodes(), 1))
    ur, b, st = _solve(X, 1e-11)
    _say("  stats", st)
    assert st["converged"]
    A = X.scipy_matrix(1)
    assert rel(A @ ur, b) <= 1e-10
    # direct solve of the oracle's agglomerated system
    Q, roff = asm.injection(tgt)
    Ar = (Q.T @ asm.A @ Q).tocsc()
    b_or = Q.T @ asm.rhs(e["ints"], e["nx"], e["nw"], [1.0, 1.0], NU, 1.3, L, lambda xx: u(xx, 1))
    ud = spla.spsolve(Ar, b_or)
    assert rel(ur, ud) <= 1e-8
    # exactness at the nodes of every element rule (volume elements)
    uf = X.extend(ur)
    err, scale = 0.0, 0.0
    for (t, P, Qc, sP, sQ, d, side, r0, rn) in e["ints"]:
        if t != 0:
            continue
        lo, hi = G.box(P)
        if rn > 0:
            xx = e["nx"][r0:r0 + rn, :3]
        else:
            xx, _ = OX._tensor_nodes(3, 2, lo, hi - lo, 0)
        v, _ = OX._phys_basis(3, 2, lo, hi - lo, xx)
        uh = v @ uf[asm.dofs(P, sP)]
        ex = u(xx, sP)
        err = max(err, np.abs(uh - ex).max())
        scale = max(scale, np.abs(ex).max())
    _say("  max nodal error", err, "scale", scale)
    assert err <= 1e-7 * scale


def test_precond_graph_replay(ctx):
    """XDG two-level preconditioner: graph replay of its coarse V-cycles equals eager launches."""
    _say("case xdg precond graph replay")
    x = np.linspace(0, 1, 13)
    g = B.Grid(ctx, x, x, x)
    X = B.XdgPoisson(g, 2, [0.5, 0.47, 0.52], 0.27, NU[0], NU[1])
    Pg = B.XdgPrecond(X, coarse_degree=1, prec=B.BH_FP32, min_coarse=8)
    Pe = B.XdgPrecond(X, coarse_degree=1, prec=B.BH_FP32, min_coarse=8)
    B.pmg_set_graphs(Pe, False)
    rng = np.random.default_rng(3)
    for rep in range(4):
        b = rng.standard_normal(X.N)
        xg = Pg.Solve(np.zeros(X.N), b)
        ctx.sync()
        xe = Pe.Solve(np.zeros(X.N), b)
        assert np.array_equal(xg, xe), f"replay {rep}"


def test_solver_contrast_1000(ctx):
    """FGMRES(30) + two-level FP32 preconditioner reaches 1e-8 on a 20^3 grid at
    1:1000 contrast; the FP64 preconditioner gives the same solution."""
    _say("case xdg solve 20^3")
    x = np.linspace(0, 1, 21)
    g = B.Grid(ctx, x, x, x)
    X = B.XdgPoisson(g, 2, [0.5, 0.47, 0.52], 0.27, NU[0], NU[1])
    X.set_rhs([1.0, 1.0], None)
    u32, b, st32 = _solve(X, 1e-8, prec=B.BH_FP32)
    _say("  fp32 precond", st32)
    u64, _, st64 = _solve(X, 1e-8, prec=B.BH_FP64)
    _say("  fp64 precond", st64)
    assert st32["converged"] and st64["converged"]
    A = X.scipy_matrix(1)
    assert rel(A @ u32, b) <= 1.5e-8
    assert rel(u32, u64) <= 1e-6


def test_assembly_and_evaluate_bitwise_reproducible(ctx):
    """No atomics in the element kernels (colour groups in a fixed order): two set-ups
    of the same sphere give bitwise identical matrices, and repeated matrix-free
    evaluates are bitwise identical."""
    _say("case xdg bitwise reproducibility")
    x = np.linspace(0, 1, 17)
    c, r = [0.51, 0.48, 0.53], 0.29
    Xs = [B.XdgPoisson(B.Grid(ctx, x, x, x), 2, c, r, NU[0], NU[1], agg_threshold=0.1) for _ in range(2)]
    A0, A1 = Xs[0].scipy_matrix(0), Xs[1].scipy_matrix(0)
    assert np.array_equal(A0.indptr, A1.indptr) and np.array_equal(A0.indices, A1.indices)
    assert np.array_equal(A0.data, A1.data)
    u = np.random.default_rng(8).standard_normal(Xs[0].N_full)
    y = [X.Evaluate(1.0, 0.0, np.zeros(X.N_full), u) for X in (Xs[0], Xs[0], Xs[1])]
    assert np.array_equal(y[0], y[1]) and np.array_equal(y[0], y[2])
    assert rel(y[0], A0 @ u) <= 1e-12


def _sphere_poly(D, c, r):
    C = np.zeros((3,) * D)
    for d in range(D):
        e = [0] * D
        e[d] = 2
        C[tuple(e)] += 1.0
        e[d] = 1
        C[tuple(e)] -= 2 * c[d]
        C[(0,) * D] += c[d] ** 2
    C[(0,) * D] -= r * r
    return C


def test_polynomial_level_set_matches_sphere_path(ctx):
    """The sphere given as the polynomial |x - c|^2 - r^2 (Saye rules, bh_xdg_create_poly)
    against the closed-form sphere path: same species masks and agglomeration map,
    agglomerated matrices within the rule accuracy, and the FGMRES solutions of the
    parabola problem within 1e-7 of each other."""
    _say("case xdg polynomial level set vs sphere")
    c, r = [0.5, 0.47, 0.52], 0.27
    x = np.linspace(0, 1, 11)
    g = B.Grid(ctx, x, x, x)
    Xs = B.XdgPoisson(g, 2, c, r, NU[0], NU[1], agg_threshold=0.1, quad_order=10)
    Xp = B.XdgPoisson.poly(g, 2, _sphere_poly(3, c, r), NU[0], NU[1], agg_threshold=0.1, quad_order=10)
    assert np.array_equal(Xs.cells()["spmask"], Xp.cells()["spmask"])
    assert np.array_equal(Xs.agglomeration(), Xp.agglomeration())
    As, Ap = Xs.scipy_matrix(1), Xp.scipy_matrix(1)
    assert As.shape == Ap.shape
    assert spla.norm(As - Ap) <= 1e-6 * spla.norm(As)
    u = _parabola(3, c, r, NU)
    sols = []
    for X in (Xs, Xp):
        X.set_rhs([1.0, 1.0], u(X.bnd_nodes(), 1))
        ur, b, st = _solve(X, 1e-11)
        assert st["converged"]
        sols.append(ur)
    _say(f"  |u_poly - u_sphere| / |u_sphere| = {rel(sols[1], sols[0]):.2e}")
    assert rel(sols[1], sols[0]) <= 1e-7


def test_polynomial_level_set_torus_solve(ctx):
    """A torus (quartic level set, genus 1) on 16^3, k = 2, contrast 1:1000: set-up,
    symmetric agglomerated operator, FGMRES + two-level preconditioner to 1e-8."""
    _say("case xdg torus solve")
    R, rr = 0.3, 0.12
    S = _sphere_poly(3, [0.5, 0.5, 0.5], 0.0)
    S[0, 0, 0] += R * R - rr * rr
    P = _sphere_poly(3, [0.5, 0.5, 0.0], 0.0)
    P[0, 0, 2] -= 1.0
    C = np.zeros((5, 5, 5))
    for ia in np.ndindex(S.shape):
        for ib in np.ndindex(S.shape):
            C[tuple(ia[d] + ib[d] for d in range(3))] += S[ia] * S[ib]
    C[:3, :3, :3] -= 4 * R * R * P
    x = np.linspace(0, 1, 17)
    g = B.Grid(ctx, x, x, x)
    X = B.XdgPoisson.poly(g, 2, C, NU[0], NU[1], agg_threshold=0.1)
    A = X.scipy_matrix(1)
    assert spla.norm(A - A.T) <= 1e-12 * spla.norm(A)
    X.set_rhs([1.0, 1.0], None)
    u, b, st = _solve(X, 1e-8)
    _say(f"  {X.ncut} cut cells, {X.npairs} agglomeration pairs, stats {st}")
    assert st["converged"]
    assert rel(A @ u, b) <= 1.5e-8


def test_schwarz_overlap_in_coarse_hierarchy(ctx):
    """Schwarz with one layer of overlap and overlap scaling (AS/Schwarz.cs:683-730,
    :791-815) in the coarse p/h hierarchy of the two-level XDG preconditioner: FGMRES
    converges to 1e-8 and gives the block-Jacobi preconditioner's solution."""
    _say("case xdg solve 16^3, schwarz overlap 1 in the coarse hierarchy")
    x = np.linspace(0, 1, 17)
    g = B.Grid(ctx, x, x, x)
    X = B.XdgPoisson(g, 2, [0.5, 0.47, 0.52], 0.27, NU[0], NU[1])
    X.set_rhs([1.0, 1.0], None)
    u0, b, st0 = _solve(X, 1e-8)
    P = B.XdgPrecond(X, coarse_degree=1, prec=B.BH_FP32)
    B.pmg_set_schwarz(P, True, overlap=1, overlap_scaling=True)
    S = B.XdgFlexGMRES(X, P, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=2000)
    u1 = np.zeros(X.N)
    S.Solve(u1, b)
    st1 = S.stats()
    _say("  block-Jacobi", st0, " schwarz overlap 1", st1)
    assert st0["converged"] and st1["converged"]
    A = X.scipy_matrix(1)
    assert rel(A @ u1, b) <= 1.5e-8
    assert rel(u1, u0) <= 1e-6


@pytest.mark.parametrize("D,n,q", [(3, 12, 6), (3, 10, 10), (2, 16, 8)])
def test_device_cut_rules_equal_host(ctx, D, n, q):
    """Sphere cut-cell, interface and cut-face rules built on the device (one thread
    per rule) equal the host set-up's bit for bit (the same construction,
    bh_sphrule.hpp, compiled without multiply-add contraction; IEEE add, multiply,
    divide and square root round alike): the same element lists, nodes, weights,
    normals, species masks and agglomeration.  n = 10 puts the sphere's centre
    coordinate 0.5 on a grid plane, where roots meet piece ends."""
    _say(f"case device cut rules D={D} n={n} q={q}")
    c = [0.5, 0.47, 0.52][:D]
    r = 0.27 if D == 3 else 0.31
    x = np.linspace(0, 1, n + 1)
    Xd = B.XdgPoisson(B.Grid(ctx, *([x] * D)), 2, c, r, NU[0], NU[1], agg_threshold=0.1, quad_order=q)
    Xh = B.XdgPoisson(B.Grid(None, *([x] * D)), 2, c, r, NU[0], NU[1], agg_threshold=0.1, quad_order=q)
    ed, eh = Xd.elements(), Xh.elements()
    assert np.array_equal(ed["ints"], eh["ints"])
    assert len(ed["nw"]) == len(eh["nw"]) > 0
    for key in ("nx", "nw", "nn", "dbls"):
        assert np.array_equal(ed[key], eh[key]), key
    cd, ch = Xd.cells(), Xh.cells()
    assert np.array_equal(cd["spmask"], ch["spmask"])
    assert np.array_equal(cd["vol"], ch["vol"])
    assert np.array_equal(Xd.agglomeration(), Xh.agglomeration())


@pytest.mark.parametrize("D,n", [(3, 10), (2, 16)])
def test_mass_matrix_device(ctx, D, n):
    """XDG mass matrix on the device (X/MassMatrixFactory.cs:35) vs the oracle's on the
    same cut rules, and the closed-form volumes and ball (disk) moments (q = 10)."""
    _say(f"case xdg mass matrix D={D} n={n} q=10")
    c, r = [0.5, 0.47, 0.52][:D], 0.27 if D == 3 else 0.31
    X, G, asm, tgt, e, L = _setup(ctx, D=D, n=n, c=c, r=r, q=10)
    M = X.mass_matrix()
    Mo = asm.mass_blocks(e["ints"], e["nx"], e["nw"])
    assert M.shape == Mo.shape
    assert np.abs(M - Mo).max() <= 1e-13
    mom, ball, cube = check_mass_blocks(M, G, asm, 2, np.array(c), r, 1e-11)
    assert np.allclose(mom[0], ball, rtol=1e-10)
    assert np.allclose(mom[1], cube - ball, rtol=1e-10)


def _torus_poly(R, rr):
    S = _sphere_poly(3, [0.5, 0.5, 0.5], 0.0)
    S[0, 0, 0] += R * R - rr * rr
    P = _sphere_poly(3, [0.5, 0.5, 0.0], 0.0)
    P[0, 0, 2] -= 1.0
    C = np.zeros((5, 5, 5))
    for ia in np.ndindex(S.shape):
        for ib in np.ndindex(S.shape):
            C[tuple(ia[d] + ib[d] for d in range(3))] += S[ia] * S[ib]
    C[:3, :3, :3] -= 4 * R * R * P
    return C


def _superellipse_poly(s):
    C = np.zeros((5, 5))      # (x - 1/2)^4 + (y - 1/2)^4 - s^4
    for i, v in enumerate([1 / 16, -1 / 2, 3 / 2, -2, 1]):
        C[i, 0] += v
        C[0, i] += v
    C[0, 0] -= s ** 4
    return C


@pytest.mark.parametrize("case", ["sphere3d", "torus", "plane", "superellipse2d"])
def test_device_poly_rules_equal_host(ctx, case):
    """Polynomial level sets: the cut-cell, interface and cut-face rules built on the
    device (Saye's algorithm, one thread per rule, bh_sayerule.hpp compiled without
    multiply-add contraction) equal the host set-up's bit for bit, including the boxes
    that need halvings (the superellipse's flat sides, the torus)."""
    if case == "sphere3d":
        D, n, q, C = 3, 10, 10, _sphere_poly(3, [0.5, 0.47, 0.52], 0.27)
    elif case == "torus":
        D, n, q, C = 3, 16, 8, _torus_poly(0.3, 0.12)
    elif case == "plane":
        D, n, q = 3, 10, 4
        C = np.zeros((2, 2, 2))
        C[0, 0, 1], C[0, 0, 0], C[1, 0, 0], C[0, 1, 0] = 1.0, -0.3, -0.2, -0.1
    else:
        D, n, q, C = 2, 20, 12, _superellipse_poly(0.3)
    _say(f"case device polynomial cut rules {case} D={D} n={n} q={q}")
    x = np.linspace(0, 1, n + 1)
    f0 = B.xdg_cut_rule_fallbacks()
    Xd = B.XdgPoisson.poly(B.Grid(ctx, *([x] * D)), 2, C, NU[0], NU[1], agg_threshold=0.1, quad_order=q)
    f1 = B.xdg_cut_rule_fallbacks()
    Xh = B.XdgPoisson.poly(B.Grid(None, *([x] * D)), 2, C, NU[0], NU[1], agg_threshold=0.1, quad_order=q)
    f2 = B.xdg_cut_rule_fallbacks()
    ed, eh = Xd.elements(), Xh.elements()
    assert Xd.ncut == Xh.ncut > 0
    assert np.array_equal(ed["ints"], eh["ints"])
    assert len(ed["nw"]) == len(eh["nw"]) > 0
    for key in ("nx", "nw", "nn", "dbls"):
        assert np.array_equal(ed[key], eh[key]), key
    cd, ch = Xd.cells(), Xh.cells()
    for key in ("spmask", "vol", "iface_area"):
        assert np.array_equal(cd[key], ch[key]), key
    assert np.array_equal(Xd.agglomeration(), Xh.agglomeration())
    assert f1 - f0 == f2 - f1

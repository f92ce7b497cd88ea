"""CPU checks that pin the oracle's p/h multigrid (oracle.cpp pmg_build /
agg_projection / pmg_vcycle) independently of its own code, following the
reference's MultigridTest (L4-application/MultigridTest/TestProgramLinearMG.cs):

* ProlongationTest (:131-166): a prolongated coarse vector is one polynomial on
  each aggregate (no jumps); here the stronger statement that it equals the coarse
  polynomial point by point;
* PolynomialRestAndPrlgTest (:231-262): restriction then prolongation of a
  polynomial field of degree <= p returns it (< 1e-8);
* RestictionMatrixTest (:276-320): the matrix form of the restriction equals the
  direct one; here R P = I and the level matrices equal dense products P^T A P
  formed in numpy from an independent projection;
* RestrictionOfSystemOpTest (:330-360): the restriction does not couple what it
  must not (modes above the coarse degree; cells outside the aggregate);
* one V-cycle equal to a dense numpy restatement of the cycle
  (AS/MultigridOperator.cs:573-640, AS/BlockJacobi.cs:196-255).

The numpy side uses only the basis and Gauss tables, which other tests pin to the
reference's own tables (F/Square.cs, F/Cube_polynomials.cs, F/LineQuadRules).
"""
import numpy as np
import pytest

import oracle as O
from tests.helpers import log_uniform_nu, nodes_3d


def _grid_nodes(D, n, seed):
    if D == 3:
        return nodes_3d(*n, jitter=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    out = []
    for m in n:
        a = np.linspace(0.0, 1.0, m + 1)
        a[1:-1] += 0.3 * (rng.random(m - 1) - 0.5) / m
        out.append(a)
    return out


def _cell_ijk(D, n, c):
    if D == 2:
        return (c // n[1], c % n[1], 0)
    return (c % n[0], (c // n[0]) % n[1], c // (n[0] * n[1]))


def _cell_num(D, n, i, j, k):
    return i * n[1] + j if D == 2 else i + j * n[0] + k * n[0] * n[1]


def _aggregate(D, n):
    """parent of each cell when 2^D cells are merged (reference numbering), and the
    coarse grid size."""
    nc = [(m + 1) // 2 for m in n[:D]] + [1] * (3 - D)
    J = int(np.prod(n[:D]))
    par = np.array([_cell_num(D, nc, *(v // 2 for v in _cell_ijk(D, n, c))) for c in range(J)])
    return par, nc


def _boxes(D, n, nodes, cells):
    lo = np.zeros((len(cells), D))
    h = np.zeros((len(cells), D))
    for t, c in enumerate(cells):
        idx = _cell_ijk(D, n, c)
        for d in range(D):
            lo[t, d] = nodes[d][idx[d]]
            h[t, d] = nodes[d][idx[d] + 1] - nodes[d][idx[d]]
    return lo, h


def _tensor_rule(D, n1):
    x, w = O.gauss(n1)
    X = np.stack([a.ravel() for a in np.meshgrid(*([x] * D), indexing="ij")], -1)
    W = np.prod(np.stack([a.ravel() for a in np.meshgrid(*([w] * D), indexing="ij")], -1), axis=1)
    return X, W


def _to_ref(x, lo, h):
    return 2.0 * (x - lo) / h - 1.0


def _project(D, deg, f, lo, h):
    """L2 projection of f onto the orthonormal basis of the box (phi = psi / sqrt(det J),
    F/BasisData.cs:43-50): c_n = sqrt(det J) sum_q w_q f(x_q) psi_n(xi_q)."""
    X, W = _tensor_rule(D, deg + 4)
    x = lo + 0.5 * (X + 1.0) * h
    psi, _ = O.basis_eval(D, deg, X)
    return np.sqrt(np.prod(0.5 * h)) * (psi * (W * f(x))[:, None]).sum(0)


def _field(D, deg, c, lo, h, x):
    psi, _ = O.basis_eval(D, deg, _to_ref(x, lo, h))
    return psi @ c / np.sqrt(np.prod(0.5 * h))


def _projection_blocks(D, deg, lof, hf, loc, hc):
    """P[m][n] = int_fine phi^f_m phi^c_n dx by numpy Gauss quadrature on the fine box."""
    X, W = _tensor_rule(D, deg + 2)
    x = lof + 0.5 * (X + 1.0) * hf
    pf, _ = O.basis_eval(D, deg, X)
    pc, _ = O.basis_eval(D, deg, _to_ref(x, loc, hc))
    s = np.sqrt(np.prod(0.5 * hf)) / np.sqrt(np.prod(0.5 * hc))   # det J_f / sqrt(det J_f det J_c)
    return s * (pf * W[:, None]).T @ pc


def _random_poly(D, deg, seed):
    rng = np.random.default_rng(seed)
    terms = [(rng.standard_normal(), e) for e in O.basis_exponents(D, deg)]

    def f(x):
        return sum(a * np.prod((x - 0.37) ** np.array(e), axis=-1) for a, e in terms)
    return f


def _hierarchy(D, n, degrees, seed, min_coarse):
    nodes = _grid_nodes(D, n, seed)
    g = O.Grid(*nodes)
    op = O.Sip(g, degrees[0], log_uniform_nu(g.J, seed))
    A, _ = op.assemble()
    P = O.Pmg(A, g, degrees, min_coarse=min_coarse, omega=0.7, pre=2, post=1)
    return nodes, g, A, P


CASES = [(2, [7, 5], 1), (2, [6, 6], 2), (2, [5, 7], 3), (3, [5, 4, 3], 1), (3, [4, 4, 3], 2), (3, [3, 3, 2], 3)]


@pytest.mark.parametrize("D,n,p", CASES)
def test_aggregation_levels_follow_the_grid(D, n, p):
    """Aggregation levels: 2^D cells per aggregate in the reference numbering, and each
    aggregate's box is the union of its children (F/CoarseningAlgorithms.cs:555-611)."""
    nodes, g, A, P = _hierarchy(D, n, [p], 11 + p, 1)
    nn = list(n) + [1] * (3 - D)
    l = 0
    cells_lo, cells_h = _boxes(D, nn, nodes, range(g.J))
    while P.level_info(l)["kind"] == 1:
        par, _ = P.level_transfer(l)
        epar, nc = _aggregate(D, nn)
        assert np.array_equal(par, epar)
        lo, h = P.level_boxes(l, D)
        assert np.allclose(lo, cells_lo, atol=1e-15) and np.allclose(h, cells_h, atol=1e-15)
        clo = np.full((par.max() + 1, D), np.inf)
        chi = np.full((par.max() + 1, D), -np.inf)
        for c in range(len(par)):
            clo[par[c]] = np.minimum(clo[par[c]], lo[c])
            chi[par[c]] = np.maximum(chi[par[c]], lo[c] + h[c])
        cells_lo, cells_h = clo, chi - clo
        nn = nc
        l += 1
        assert P.level_info(l)["n"][:D] == nn[:D]
    assert l >= 1


@pytest.mark.parametrize("D,n,p", CASES)
def test_prolongation_blocks_equal_numpy_projection(D, n, p):
    """P_i[m][n] = int_{cell i} phi_m phi^agg_n (AS/AggregationGridBasis.cs), every
    aggregation level, against numpy quadrature on each fine box."""
    nodes, g, A, P = _hierarchy(D, n, [p], 21 + p, 1)
    l = 0
    while P.level_info(l)["kind"] == 1:
        par, Pb = P.level_transfer(l)
        lof, hf = P.level_boxes(l, D)
        loc, hc = P.level_boxes(l + 1, D)
        for c in range(len(par)):
            ref = _projection_blocks(D, p, lof[c], hf[c], loc[par[c]], hc[par[c]])
            assert np.abs(Pb[c] - ref).max() < 1e-13
        l += 1


@pytest.mark.parametrize("D,n,p", CASES)
def test_prolongation_is_the_coarse_polynomial(D, n, p):
    """ProlongationTest (TestProgramLinearMG.cs:131-166): a random coarse vector,
    prolongated, is the aggregate's polynomial on every child cell (so it has no jumps
    inside an aggregate), at random points, every aggregation level."""
    nodes, g, A, P = _hierarchy(D, n, [p], 31 + p, 1)
    rng = np.random.default_rng(p)
    l = 0
    while P.level_info(l)["kind"] == 1:
        par, Pb = P.level_transfer(l)
        lof, hf = P.level_boxes(l, D)
        loc, hc = P.level_boxes(l + 1, D)
        Np = Pb.shape[1]
        xc = rng.random((par.max() + 1, Np))
        for c in range(len(par)):
            xf = Pb[c] @ xc[par[c]]
            pts = lof[c] + rng.random((9, D)) * hf[c]
            uf = _field(D, p, xf, lof[c], hf[c], pts)
            uc = _field(D, p, xc[par[c]], loc[par[c]], hc[par[c]], pts)
            assert np.abs(uf - uc).max() < 1e-11 * max(1.0, np.abs(uc).max())
        l += 1


@pytest.mark.parametrize("D,n,p", CASES)
def test_polynomial_restriction_and_prolongation(D, n, p):
    """PolynomialRestAndPrlgTest (TestProgramLinearMG.cs:231-262): for a polynomial of
    degree <= p, R = P^T applied to its fine projection is its projection on the
    aggregate, and P R returns the fine projection (reference bound 1e-8)."""
    nodes, g, A, P = _hierarchy(D, n, [p], 41 + p, 1)
    f = _random_poly(D, p, 7 + p)
    l = 0
    while P.level_info(l)["kind"] == 1:
        par, Pb = P.level_transfer(l)
        lof, hf = P.level_boxes(l, D)
        loc, hc = P.level_boxes(l + 1, D)
        cf = np.array([_project(D, p, f, lof[c], hf[c]) for c in range(len(par))])
        rc = np.zeros((par.max() + 1, Pb.shape[2]))
        for c in range(len(par)):
            rc[par[c]] += Pb[c].T @ cf[c]
        cc = np.array([_project(D, p, f, loc[I], hc[I]) for I in range(len(rc))])
        assert np.abs(rc - cc).max() < 1e-11 * np.abs(cc).max()
        back = np.array([Pb[c] @ rc[par[c]] for c in range(len(par))])
        assert np.linalg.norm(back - cf) < 1e-8
        assert np.abs(back - cf).max() < 1e-11 * np.abs(cf).max()
        l += 1


def _dense_prolongation(P, l):
    """dense P_l: level l+1 -> level l (truncation: injection of the leading modes)."""
    a, b = P.level_info(l), P.level_info(l + 1)
    Nf, Nc = a["nrows"] * a["bs"], b["nrows"] * b["bs"]
    M = np.zeros((Nf, Nc))
    if a["kind"] == 0:
        for c in range(a["nrows"]):
            for m in range(b["bs"]):
                M[c * a["bs"] + m, c * b["bs"] + m] = 1.0
    else:
        par, Pb = P.level_transfer(l)
        bs = a["bs"]
        for c in range(a["nrows"]):
            M[c * bs:(c + 1) * bs, par[c] * bs:(par[c] + 1) * bs] = Pb[c]
    return M


@pytest.mark.parametrize("D,n,degs", [(2, [7, 6], [2, 1]), (3, [4, 4, 3], [2, 1]), (3, [5, 3, 3], [3, 1])])
def test_restriction_identity_and_galerkin_products(D, n, degs):
    """RestictionMatrixTest / RAP (AS/MultigridOperator.cs:629): R P = I on every level
    and every level matrix equals P^T A P of the level above, with P built in numpy
    (projection blocks by quadrature, truncation by injection)."""
    nodes, g, A, P = _hierarchy(D, n, degs, 5, 2)
    nl = len(P.sizes)
    assert nl >= 3
    Ad = A.dense()
    for l in range(nl - 1):
        Pd = _dense_prolongation(P, l)
        a = P.level_info(l)
        if a["kind"] == 1:   # the oracle's blocks == independent quadrature (checked above); rebuild in numpy
            par, _ = P.level_transfer(l)
            lof, hf = P.level_boxes(l, D)
            loc, hc = P.level_boxes(l + 1, D)
            bs = a["bs"]
            Pd = np.zeros_like(Pd)
            for c in range(a["nrows"]):
                Pd[c * bs:(c + 1) * bs, par[c] * bs:(par[c] + 1) * bs] = _projection_blocks(
                    D, a["deg"], lof[c], hf[c], loc[par[c]], hc[par[c]])
        assert np.abs(Pd.T @ Pd - np.eye(Pd.shape[1])).max() < 1e-12
        Ad = Pd.T @ Ad @ Pd
        Ao = P.level_matrix(l + 1).dense()
        assert np.abs(Ao - Ad).max() < 1e-11 * np.abs(Ad).max()


@pytest.mark.parametrize("D,n", [(2, [6, 5]), (3, [4, 3, 3])])
def test_restriction_couples_only_what_it_must(D, n):
    """RestrictionOfSystemOpTest (TestProgramLinearMG.cs:330-360), single-field form:
    the truncation restriction drops exactly the modes above the coarse degree (a fine
    vector carried by those modes restricts to zero), an aggregate's restriction reads
    only its children, and the coarse block structure is the aggregate adjacency."""
    nodes, g, A, P = _hierarchy(D, n, [2, 1], 9, 1)
    a, b = P.level_info(0), P.level_info(1)
    assert a["kind"] == 0
    R0 = _dense_prolongation(P, 0).T
    hi = np.zeros(a["nrows"] * a["bs"])
    for c in range(a["nrows"]):
        hi[c * a["bs"] + b["bs"]:(c + 1) * a["bs"]] = np.random.default_rng(c).standard_normal(a["bs"] - b["bs"])
    assert np.all(R0 @ hi == 0.0)
    assert np.abs(R0 @ np.ones(len(hi))).min() == 1.0
    # aggregation level 1 -> 2
    par, _ = P.level_transfer(1)
    R1 = _dense_prolongation(P, 1).T
    bs = b["bs"]
    for c in range(b["nrows"]):
        rows = np.abs(R1[:, c * bs:(c + 1) * bs]).sum(1) > 0
        assert set(np.flatnonzero(rows) // bs) == {par[c]}
    nn = list(n) + [1] * (3 - D)
    rp, col, _ = P.level_matrix(2).export()
    adj = set()
    for c in range(b["nrows"]):
        i = _cell_ijk(D, nn, c)
        for d in range(D):
            j = list(i)
            j[d] += 1
            if j[d] < nn[d]:
                adj.add((par[c], par[_cell_num(D, nn, *j)]))
                adj.add((par[_cell_num(D, nn, *j)], par[c]))
        adj.add((par[c], par[c]))
    got = {(r, int(col[q])) for r in range(len(rp) - 1) for q in range(rp[r], rp[r + 1])}
    assert got == adj


def _dense_vcycle(Al, Pl, bsl, omega, pre, post, l, b):
    """dense numpy restatement of the V-cycle: block-Jacobi pre-smoothing from x = 0
    (AS/BlockJacobi.cs:196-255), residual restriction R = P^T, recursive coarse
    correction, post-smoothing; dense solve on the coarsest level."""
    A = Al[l]
    if l == len(Al) - 1:
        return np.linalg.solve(A, b)
    bs = bsl[l]
    nb = A.shape[0] // bs
    Dinv = [np.linalg.inv(A[c * bs:(c + 1) * bs, c * bs:(c + 1) * bs]) for c in range(nb)]

    def smooth(r):
        return omega * np.concatenate([Dinv[c] @ r[c * bs:(c + 1) * bs] for c in range(nb)])
    x = smooth(b)
    for _ in range(pre - 1):
        x = x + smooth(b - A @ x)
    xc = _dense_vcycle(Al, Pl, bsl, omega, pre, post, l + 1, Pl[l].T @ (b - A @ x))
    x = x + Pl[l] @ xc
    for _ in range(post):
        x = x + smooth(b - A @ x)
    return x


@pytest.mark.parametrize("D,n,degs,omega,pre,post", [
    (2, [7, 6], [2, 1], 0.7, 2, 1),
    (3, [4, 4, 3], [2, 1], 0.8, 1, 1),
    (3, [5, 3, 3], [3, 1], 0.6, 1, 2),
])
def test_vcycle_equals_dense_restatement(D, n, degs, omega, pre, post):
    """One V-cycle of the oracle (pmg_vcycle) equals the dense numpy cycle built from
    the fine matrix and numpy prolongations (Galerkin products, dense block inverses,
    dense coarsest solve)."""
    nodes, g, A, P = _hierarchy(D, n, degs, 3, 2)
    P.set(omega, pre, post)
    nl = len(P.sizes)
    Al = [A.dense()]
    Pl = []
    bsl = []
    for l in range(nl - 1):
        a = P.level_info(l)
        bsl.append(a["bs"])
        Pd = _dense_prolongation(P, l)
        if a["kind"] == 1:
            par, _ = P.level_transfer(l)
            lof, hf = P.level_boxes(l, D)
            loc, hc = P.level_boxes(l + 1, D)
            bs = a["bs"]
            Pd = np.zeros_like(Pd)
            for c in range(a["nrows"]):
                Pd[c * bs:(c + 1) * bs, par[c] * bs:(par[c] + 1) * bs] = _projection_blocks(
                    D, a["deg"], lof[c], hf[c], loc[par[c]], hc[par[c]])
        Pl.append(Pd)
        Al.append(Pd.T @ Al[-1] @ Pd)
    b = np.random.default_rng(17).standard_normal(A.nrows * A.bs)
    ref = _dense_vcycle(Al, Pl, bsl, omega, pre, post, 0, b)
    got = P.vcycle(b)
    assert np.abs(got - ref).max() < 1e-10 * np.abs(ref).max()


def test_vcycle_dense_restatement_detects_a_planted_error():
    """The dense restatement is sensitive to the kind of mistake it guards against:
    prolongation blocks stored the other way round ([coarse][fine] instead of
    [fine][coarse]) change the cycle."""
    D, n, degs = 3, [4, 4, 3], [2, 1]
    nodes, g, A, P = _hierarchy(D, n, degs, 3, 2)
    P.set(0.8, 1, 1)
    nl = len(P.sizes)
    Al, Pl, bsl = [A.dense()], [], []
    for l in range(nl - 1):
        a = P.level_info(l)
        bsl.append(a["bs"])
        Pd = _dense_prolongation(P, l)
        if a["kind"] == 1:
            par, Pb = P.level_transfer(l)
            bs = a["bs"]
            for c in range(a["nrows"]):
                Pd[c * bs:(c + 1) * bs, par[c] * bs:(par[c] + 1) * bs] = Pb[c].T
        Pl.append(Pd)
        Al.append(Pd.T @ Al[-1] @ Pd)
    b = np.random.default_rng(17).standard_normal(A.nrows * A.bs)
    ref = _dense_vcycle(Al, Pl, bsl, 0.8, 1, 1, 0, b)
    assert np.abs(P.vcycle(b) - ref).max() > 1e-6 * np.abs(ref).max()

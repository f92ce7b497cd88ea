"""GPU parity of the p/h multigrid preconditioner against the oracle hierarchy.

V-cycle: the device cycle (FP64) equals the oracle's restated cycle to 1e-11
(same hierarchy: Galerkin p-sub-blocks, degree-p aggregation, block-Jacobi,
dense coarsest solve).  K-cycle: equals a numpy restatement of
`kin` flexible-GMRES steps on the oracle's level operator.  FP32 cycle within
1e-4 of FP64.  End-to-end: FGMRES + multigrid reaches 1e-8 true residual.
"""
import numpy as np
import pytest

import bosss_amd as B
import oracle as O
from tests.helpers import log_uniform_nu, nodes_3d

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _problem(ctx, n=(6, 5, 4), k=3, seed=3):
    nodes = nodes_3d(*n, jitter=0.3, seed=seed)
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, seed)
    gg = B.Grid(ctx, *nodes)
    oo = O.Sip(og, k, nu)
    go = B.SipOperator(gg, k, nu)
    Ao, _ = oo.assemble()
    return og, gg, oo, go, Ao


@pytest.mark.parametrize("matrix_free", [False, True])
@pytest.mark.parametrize("degrees", [[3, 1], [3, 2, 1, 0]])
def test_vcycle_matches_oracle(ctx, matrix_free, degrees):
    _say(f"case vcycle degrees={degrees} matrix_free={matrix_free}")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(1).standard_normal(oo.N)
    ref, nl = Ao.pmg_vcycle(og, b, degrees, min_coarse=8, omega=0.7, pre=2, post=1)
    mg = B.PMultigrid(go, degrees=degrees, min_coarse=8, omega=0.7, pre=2, post=1, kcycle_inner=0,
                      matrix_free=matrix_free)
    out = mg.Solve(np.zeros(oo.N), b)   # x = 0 + V(b - 0)
    assert nl >= 3
    assert rel(out, ref) < 1e-11


@pytest.mark.parametrize("matrix_free", [False, True])
@pytest.mark.parametrize("prec,tol", [(B.BH_FP64, 1e-11), (B.BH_FP32, 1e-4)])
def test_vcycle_schwarz_matches_oracle(ctx, matrix_free, prec, tol):
    """Aggregate-block additive Schwarz smoothing on the k = 1 level and the h-levels
    (AS/Schwarz.cs with multigrid-aggregate blocks, no overlap): device V-cycle vs the
    oracle cycle with the same smoother."""
    _say(f"case vcycle schwarz matrix_free={matrix_free} prec={prec}")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(5).standard_normal(oo.N)
    P = O.Pmg(Ao, og, [3, 1], min_coarse=8, omega=0.7, pre=2, post=1)
    P.set_schwarz(True)
    ref = P.vcycle(b)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.7, pre=2, post=1, kcycle_inner=0,
                      matrix_free=matrix_free, prec=prec)
    B.pmg_set_schwarz(mg, True)
    out = mg.Solve(np.zeros(oo.N), b)
    assert rel(out, ref) < tol
    B.pmg_set_schwarz(mg, False)
    P.set_schwarz(False)
    assert rel(mg.Solve(np.zeros(oo.N), b), P.vcycle(b)) < tol


@pytest.mark.parametrize("overlap,scaling", [(1, False), (1, True), (2, True)])
@pytest.mark.parametrize("prec,tol", [(B.BH_FP64, 1e-11), (B.BH_FP32, 1e-4)])
def test_vcycle_schwarz_overlap_matches_oracle(ctx, overlap, scaling, prec, tol):
    """Schwarz with overlap (AS/Schwarz.cs:683-730: aggregate blocks grown by `overlap`
    face-neighbour sweeps; the block solutions summed): device V-cycle vs the oracle
    cycle.  On 6x5x4 cells at k = 1 one layer gives blocks of up to 32 cells (128 DOF,
    the limit); two layers exceed it on the fine level, which keeps block-Jacobi on
    both sides, while the coarser h-levels take the grown blocks.  With scaling the
    block sums are divided by the DOF's multiplicity (:791-815)."""
    _say(f"case vcycle schwarz overlap={overlap} scaling={scaling} prec={prec}")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(6).standard_normal(oo.N)
    P = O.Pmg(Ao, og, [3, 1], min_coarse=8, omega=0.5, pre=2, post=1)
    P.set_schwarz(True, overlap=overlap, scaling=scaling)
    ref = P.vcycle(b)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.5, pre=2, post=1, kcycle_inner=0, prec=prec)
    B.pmg_set_schwarz(mg, True, overlap=overlap, overlap_scaling=scaling)
    out = mg.Solve(np.zeros(oo.N), b)
    assert rel(out, ref) < tol
    # a second set-up on the same solver replaces the blocks
    B.pmg_set_schwarz(mg, True, overlap=0)
    P.set_schwarz(True, overlap=0)
    assert rel(mg.Solve(np.zeros(oo.N), b), P.vcycle(b)) < tol


def test_schwarz_overlap_fgmres_converges(ctx):
    """FGMRES + V-cycle with one layer of overlap and overlap scaling reaches 1e-8 and
    gives the solution of the smoother without overlap."""
    _say("case fgmres schwarz overlap 1, scaled")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(7).standard_normal(oo.N)
    its, sols = [], []
    for ov in (0, 1):
        mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.5, pre=2, post=2, kcycle_inner=0)
        B.pmg_set_schwarz(mg, True, overlap=ov, overlap_scaling=True)
        S = B.FlexGMRES(op=go, Precond=mg, MaxKrylovDim=40, ConvergenceCriterion=1e-8)
        u = S.Solve(np.zeros(oo.N), b)
        st = S.stats()
        _say(f"  overlap {ov}: {st['iterations']} iterations")
        assert st["converged"]
        its.append(st["iterations"])
        sols.append(u)
    assert rel(sols[1], sols[0]) < 1e-6
    with pytest.raises(B.BoSSSHipError):
        B.pmg_set_schwarz(mg, True, overlap=3)


@pytest.mark.parametrize("matrix_free", [False, True])
@pytest.mark.parametrize("prec,tol", [(B.BH_FP64, 1e-11), (B.BH_FP32, 1e-4)])
def test_level_pmg_matches_oracle(ctx, matrix_free, prec, tol):
    """LevelPmg (AS/LevelPmg.cs, UseDiagonalPmg): the low-order correction through the
    p/h hierarchy below, then each cell's high-order modes with the inverse of their
    diagonal sub-block; device cycle vs the oracle cycle."""
    _say(f"case level pmg matrix_free={matrix_free} prec={prec}")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(8).standard_normal(oo.N)
    P = O.Pmg(Ao, og, [3, 1], min_coarse=8, omega=0.7, pre=2, post=1)
    P.set_level_pmg(True)
    ref = P.vcycle(b)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.7, pre=2, post=1, kcycle_inner=0,
                      matrix_free=matrix_free, prec=prec)
    B.pmg_set_level_pmg(mg, True)
    out = mg.Solve(np.zeros(oo.N), b)
    assert rel(out, ref) < tol
    B.pmg_set_level_pmg(mg, False)
    P.set_level_pmg(False)
    assert rel(mg.Solve(np.zeros(oo.N), b), P.vcycle(b)) < tol


def test_level_pmg_fgmres_converges(ctx):
    """FGMRES preconditioned by LevelPmg over the K-cycle reaches 1e-8 in the true
    residual and gives the plain multigrid's solution."""
    _say("case fgmres level pmg")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(9).standard_normal(oo.N)
    sols = []
    for on in (False, True):
        mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=4, prec=B.BH_FP32)
        B.pmg_set_level_pmg(mg, on)
        S = B.FlexGMRES(op=go, Precond=mg, MaxKrylovDim=40, ConvergenceCriterion=1e-8)
        u = S.Solve(np.zeros(oo.N), b)
        st = S.stats()
        _say(f"  level pmg {on}: {st['iterations']} iterations")
        assert st["converged"]
        r = b - go.Evaluate(1.0, 0.0, np.zeros(oo.N), u, affine=False)
        assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(b)
        sols.append(u)
    assert rel(sols[1], sols[0]) < 1e-6


def _ilu_blocks(Ao, bs):
    """blocks {(i, j): bs x bs} and sorted block rows of the oracle's matrix."""
    M = Ao.dense()
    J = M.shape[0] // bs
    blocks, rows = {}, []
    for i in range(J):
        r = []
        for j in range(J):
            Bk = M[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs]
            if np.any(Bk != 0):
                blocks[(i, j)] = Bk.copy()
                r.append(j)
        rows.append(r)
    return blocks, rows


@pytest.mark.parametrize("p", [0, 1])
@pytest.mark.parametrize("prec,tol", [(B.BH_FP64, 1e-10), (B.BH_FP32, 1e-4)])
def test_cell_ilu_smoother_matches_oracle(ctx, p, prec, tol):
    """CellILU (AS/CellILU.cs) as the finest level's smoother: the device's level-scheduled
    factorisation and sweeps, x = omega U^-1 L^-1 r, against the numpy restatement on the
    oracle's matrix (6x5x4 cells, k = 3, 20x20 blocks)."""
    from oracle import ilu
    _say(f"case cell ilu({p}) prec={prec}")
    og, gg, oo, go, Ao = _problem(ctx)
    bs = oo.Np
    blocks, rows = _ilu_blocks(Ao, bs)
    W, pat = ilu.factor(blocks, rows, p)
    r = np.random.default_rng(10).standard_normal(oo.N)
    ref = 0.7 * ilu.solve(W, pat, r, bs)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.7, pre=1, post=1, kcycle_inner=0, prec=prec)
    B.pmg_set_cell_ilu(mg, p)
    out = B.pmg_smooth(mg, 0, r)
    _say(f"  rel err {rel(out, ref):.2e}")
    assert rel(out, ref) < tol
    B.pmg_set_cell_ilu(mg, -1)   # back to block-Jacobi
    assert rel(B.pmg_smooth(mg, 0, r), ref) > 1e-3


def test_cell_ilu_fgmres_converges(ctx):
    """FGMRES + V-cycle with the CellILU(1) smoother on the finest level reaches 1e-8 in
    the true residual and gives the block-Jacobi cycle's solution."""
    _say("case fgmres cell ilu")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(12).standard_normal(oo.N)
    sols = []
    for p in (-1, 1):
        mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.8, pre=1, post=1, kcycle_inner=4, prec=B.BH_FP32)
        B.pmg_set_cell_ilu(mg, p)
        S = B.FlexGMRES(op=go, Precond=mg, MaxKrylovDim=40, ConvergenceCriterion=1e-8)
        u = S.Solve(np.zeros(oo.N), b)
        st = S.stats()
        _say(f"  cell ilu {p}: {st['iterations']} iterations")
        assert st["converged"]
        rr = b - go.Evaluate(1.0, 0.0, np.zeros(oo.N), u, affine=False)
        assert np.linalg.norm(rr) <= 1.0001e-8 * np.linalg.norm(b)
        sols.append(u)
    assert rel(sols[1], sols[0]) < 1e-6


@pytest.mark.parametrize("matrix_free", [False, True])
def test_multigrid_mapping_matches_oracle(ctx, matrix_free):
    """The device hierarchy's multigrid mapping (levels, DOF per cell, degrees, logical
    grids, aggregation parents and prolongation blocks) equals the oracle's, and every
    Galerkin level matrix equals the oracle's to 1e-12 (AS/MultigridOperator.cs:573-640,
    AS/AggregationGridBasis.cs)."""
    _say(f"case multigrid mapping matrix_free={matrix_free}")
    og, gg, oo, go, Ao = _problem(ctx)
    P = O.Pmg(Ao, og, [3, 1], min_coarse=8, omega=0.7, pre=1, post=1)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, omega=0.7, pre=1, post=1, kcycle_inner=0, matrix_free=matrix_free)
    lv = B.pmg_levels(mg)
    assert len(lv) == len(P.sizes) >= 4
    for l, L in enumerate(lv):
        ref = P.level_info(l)
        assert (L["nrows"], L["bs"], L["deg"]) == (ref["nrows"], ref["bs"], ref["deg"])
        if l + 1 < len(lv):
            assert L["kind"] == ref["kind"]
            assert L["n"] == ref["n"]
        if L["kind"] == 1:
            par, Pb = B.pmg_level_transfer(mg, l)
            rpar, rPb = P.level_transfer(l)
            assert np.array_equal(par, rpar)
            assert np.abs(Pb - rPb).max() <= 1e-14 * np.abs(rPb).max()
        rp, col, blk = B.pmg_level_matrix(mg, l)
        orp, ocol, oblk = P.level_matrix(l).export()
        assert np.array_equal(rp, orp) and np.array_equal(col, ocol)
        assert np.abs(blk - oblk).max() <= 1e-12 * np.abs(oblk).max()


def _kcycle_numpy(P, level, b, kin):
    """kin flexible-GMRES steps on level `level` (CGS2), V-cycle preconditioned."""
    n = b.size
    beta = np.linalg.norm(b)
    V = [b / beta]
    Z = []
    H = np.zeros((kin + 1, kin))
    for j in range(kin):
        z = P.vcycle(V[j], level)
        Z.append(z)
        w = P.spmv(level, z)
        for _ in range(2):
            h = np.array([w @ v for v in V])
            w = w - sum(hh * v for hh, v in zip(h, V))
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V.append(w / H[j + 1, j])
    e = np.zeros(kin + 1)
    e[0] = beta
    y = np.linalg.lstsq(H, e, rcond=None)[0]
    return np.array(Z).T @ y


def test_kcycle_matches_numpy_restatement(ctx):
    _say("case kcycle")
    og, gg, oo, go, Ao = _problem(ctx)
    degrees = [3, 1]
    kin = 3
    P = O.Pmg(Ao, og, degrees, min_coarse=8, omega=0.7, pre=2, post=2)
    b = np.random.default_rng(2).standard_normal(oo.N)
    # fine level of the cycle in numpy: BJ pre-smoothing, truncation, K-cycle, post-smoothing
    rp, col, val = Ao.export()
    Np, J = oo.Np, og.J
    Dinv = np.array([np.linalg.inv(val[[p for p in range(rp[j], rp[j + 1]) if col[p] == j][0]]) for j in range(J)])
    A = lambda v: Ao.spmv(1.0, v, 0.0, np.zeros(oo.N))
    bj = lambda r: np.einsum("jab,jb->ja", Dinv, r.reshape(J, Np)).ravel()
    x = 0.7 * bj(b)
    x += 0.7 * bj(b - A(x))
    r = b - A(x)
    rc = r.reshape(J, Np)[:, :4].ravel()
    xc = _kcycle_numpy(P, 1, rc, kin)
    xx = x.reshape(J, Np).copy()
    xx[:, :4] += xc.reshape(J, 4)
    x = xx.ravel()
    for _ in range(2):
        x += 0.7 * bj(b - A(x))
    mg = B.PMultigrid(go, degrees=degrees, min_coarse=8, omega=0.7, pre=2, post=2, kcycle_inner=kin)
    out = mg.Solve(np.zeros(oo.N), b)
    assert rel(out, x) < 1e-9


@pytest.mark.parametrize("degrees", [[3, 1], [3, 2, 1, 0]])
def test_vcycle_diag_equilib_matches_oracle(ctx, degrees):
    """DiagBlockEquilib (every level but the coarsest in the basis where its diagonal
    blocks are diagonals of signs; minus the identity for the SIP form): with
    block-Jacobi smoothing the V-cycle is the same operator on the caller's vectors
    (a similarity per level: L A L^T, C^T P L^T, x~ += omega S r~), so it equals the
    oracle's untransformed cycle."""
    _say(f"case vcycle diag-equilib degrees={degrees}")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(7).standard_normal(oo.N)
    ref, nl = Ao.pmg_vcycle(og, b, degrees, min_coarse=8, omega=0.7, pre=2, post=1)
    mg = B.PMultigrid(go, degrees=degrees, min_coarse=8, omega=0.7, pre=2, post=1, kcycle_inner=0, matrix_free=False)
    B.pmg_set_diag_equilib(mg)
    out = mg.Solve(np.zeros(oo.N), b)
    assert nl >= 3
    assert rel(out, ref) < 1e-10
    m32 = B.PMultigrid(go, degrees=degrees, min_coarse=8, omega=0.7, pre=2, post=1, kcycle_inner=0, matrix_free=False,
                       prec=B.BH_FP32)
    B.pmg_set_diag_equilib(m32)
    assert rel(m32.Solve(np.zeros(oo.N), b), ref) < 1e-4


def test_diag_equilib_rejects_matrix_free(ctx):
    og, gg, oo, go, Ao = _problem(ctx, n=(3, 3, 3))
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=0, matrix_free=True)
    with pytest.raises(B.BoSSSHipError):
        B.pmg_set_diag_equilib(mg)


def test_fp32_cycle_close_to_fp64(ctx):
    _say("case fp32 cycle")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(4).standard_normal(oo.N)
    m64 = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=2, prec=B.BH_FP64)
    m32 = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=2, prec=B.BH_FP32)
    x64 = m64.Solve(np.zeros(oo.N), b)
    x32 = m32.Solve(np.zeros(oo.N), b)
    assert rel(x32, x64) < 1e-4


def test_fp16_mfma_cycle_close_to_fp64(ctx):
    """FP16-MFMA cycle (FP16 level-0 apply and FP16 level-0 smoother blocks, each
    block scaled by its largest entry) against the FP64 cycle: within the FP16
    rounding of the preconditioner."""
    _say("case fp16-mfma cycle")
    og, gg, oo, go, Ao = _problem(ctx)
    b = np.random.default_rng(6).standard_normal(oo.N)
    m64 = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=2, prec=B.BH_FP64, matrix_free=True)
    m16 = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=2, prec=B.BH_FP16_MFMA, matrix_free=True)
    x64 = m64.Solve(np.zeros(oo.N), b)
    x16 = m16.Solve(np.zeros(oo.N), b)
    assert rel(x16, x64) < 5e-3


@pytest.mark.parametrize("kin", [0, 2])
def test_graph_replay_matches_eager(ctx, kin):
    """The HIP-graph replay of the synchronisation-free V-cycles gives the same
    preconditioner output as eager launches, on every replay, with eager work
    and synchronisations of the program in between."""
    _say(f"case graph replay kin={kin}")
    og, gg, oo, go, Ao = _problem(ctx)
    mg_g = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=kin, prec=B.BH_FP32)
    mg_e = B.PMultigrid(go, degrees=[3, 1], min_coarse=8, kcycle_inner=kin, prec=B.BH_FP32)
    B.pmg_set_graphs(mg_e, False)
    rng = np.random.default_rng(9)
    for rep in range(5):
        b = rng.standard_normal(oo.N)
        xg = mg_g.Solve(np.zeros(oo.N), b)
        go.Evaluate(1.0, 0.0, np.zeros(go.N), b, affine=False)   # eager work between replays
        ctx.sync()
        xe = mg_e.Solve(np.zeros(oo.N), b)
        assert np.array_equal(xg, xe), f"replay {rep}"


@pytest.mark.parametrize("prec,equilib", [(B.BH_FP64, False), (B.BH_FP32, False), (B.BH_FP16_MFMA, False),
                                          (B.BH_FP64, True), (B.BH_FP32, True)])
def test_fgmres_multigrid_solve(ctx, prec, equilib):
    """FGMRES(30) + p/h multigrid (K-cycle) to 1e-8 true FP64 residual, 16^3 k=3, log-uniform nu
    (also with the block-equilibrated assembled hierarchy)."""
    _say(f"case fgmres+mg prec={prec} equilib={equilib}")
    n = 16
    x = np.linspace(0, 1, n + 1)
    og = O.Grid(x, x, x)
    nu = log_uniform_nu(og.J, 11)
    gg = B.Grid(ctx, x, x, x)
    go = B.SipOperator(gg, 3, nu)
    oo = O.Sip(og, 3, nu)
    b = np.random.default_rng(5).standard_normal(go.N)
    mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=64, kcycle_inner=2, prec=prec)
    if equilib:
        mg = B.PMultigrid(go, degrees=[3, 1], min_coarse=64, kcycle_inner=2, prec=prec, matrix_free=False)
        B.pmg_set_diag_equilib(mg)
    fg = B.FlexGMRES(op=go, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=300)
    xs = fg.Solve(np.zeros(go.N), b)
    st = fg.stats()
    _say("  stats", st)
    assert st["converged"]
    assert st["iterations"] < 150
    r = b - oo.evaluate(xs)
    assert np.linalg.norm(r) <= 1.0001e-8 * np.linalg.norm(b)

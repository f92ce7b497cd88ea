"""Variable-block BlockMsrMatrix in the XDG constant frame (bh_bsr_create_v,
bh_xdg_precond_create_v, bh_xdg_dof_frame).

BoSSS lays out an XDG field with a constant frame per cell, j * Np * (species
count), the present species' blocks first (X/XDGBasis.cs:46-53), and its
BlockMsrMatrix rows follow that frame with variable block lengths
(P/BlockMsrMatrix.cs:191-340).  This library stores the rows back to back
("compressed").  The tests check:

* the exported frame map equals the one derived in numpy from the species mask
  and the agglomeration pairs, bit-exact;
* SpMV of the caller's frame-layout matrix equals the library's compressed SpMV
  bitwise (FP64 and FP32), whatever the unused frame slots of x hold, and writes
  zeros into them;
* FGMRES with the two-level preconditioner built from the caller's frame-layout
  agglomerated matrix converges like the library's own XDG solve and gives the
  same solution.
"""
import numpy as np
import pytest

import bosss_amd as B

pytestmark = pytest.mark.gpu

NU = (1.0, 1000.0)


def _say(*a):
    print(*a, flush=True)


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _problem(ctx, n=12):
    x = np.linspace(0, 1, n + 1)
    g = B.Grid(ctx, x, x, x)
    X = B.XdgPoisson(g, 2, [0.5, 0.47, 0.52], 0.27, NU[0], NU[1], agg_threshold=0.1)
    return g, X


def _frame_matrix(ctx, X, which):
    """The library's matrix (full: cell rows; agglomerated: species units merged
    into cell rows) handed over in the constant frame, row-major blocks."""
    i0, ln, c0 = X.dof_frame(which)
    J = X.J
    if which == 0:
        # the full matrix has cell rows already: same blocks in the same order
        off, rp, col, voff, v = X.matrix(0)
        assert np.array_equal(off[:-1], c0)
        blocks = [v[voff[p]:voff[p + 1]].reshape(ln[col[p]], ln[r]).T.ravel()
                  for r in range(J) for p in range(rp[r], rp[r + 1])]
        M = B.BlockMsrMatrix.variable(ctx, rp, col, i0, ln, np.concatenate(blocks))
        return M, i0, ln, c0
    A = X.scipy_matrix(which).tocsr()
    cell_of = np.repeat(np.arange(J), ln)
    coo = A.tocoo()
    pairs = np.unique(np.stack([cell_of[coo.row], cell_of[coo.col]], 1), axis=0)
    rowptr = np.zeros(J + 1, np.int64)
    np.add.at(rowptr, pairs[:, 0] + 1, 1)
    rowptr = np.cumsum(rowptr)
    col = pairs[:, 1].astype(np.int32)
    blocks = []
    for r, c in pairs:
        blocks.append(A[c0[r]:c0[r] + ln[r], c0[c]:c0[c] + ln[c]].toarray().ravel())
    vals = np.concatenate(blocks) if blocks else np.zeros(0)
    M = B.BlockMsrMatrix.variable(ctx, rowptr, col, i0, ln, vals)
    return M, i0, ln, c0


def _to_frame(v, i0, ln, c0, nf, fill=0.0):
    out = np.full(nf, fill)
    for j in range(len(ln)):
        out[i0[j]:i0[j] + ln[j]] = v[c0[j]:c0[j] + ln[j]]
    return out


def _from_frame(vf, i0, ln, c0, n):
    out = np.zeros(n)
    for j in range(len(ln)):
        out[c0[j]:c0[j] + ln[j]] = vf[i0[j]:i0[j] + ln[j]]
    return out


def test_dof_frame_map(ctx):
    _say("case xdg frame map 12^3")
    g, X = _problem(ctx)
    Np = X.Np
    mask = X.cells()["spmask"]
    nsp = (mask & 1) + ((mask >> 1) & 1)
    i0, ln, c0 = X.dof_frame(0)
    assert np.array_equal(i0, np.arange(X.J, dtype=np.int64) * 2 * Np)
    assert np.array_equal(ln, nsp * Np)
    assert np.array_equal(c0, np.concatenate([[0], np.cumsum(ln)[:-1]]))
    assert int(ln.sum()) == X.N_full
    src = np.zeros(X.J, np.int64)
    for (s, t, sp) in X.agglomeration():
        src[s] += 1
    i1, l1, c1 = X.dof_frame(1)
    assert np.array_equal(i1, i0)
    assert np.array_equal(l1, (nsp - src) * Np)
    assert np.array_equal(c1, np.concatenate([[0], np.cumsum(l1)[:-1]]))
    assert int(l1.sum()) == X.N


def test_frame_spmv_bitwise(ctx):
    _say("case xdg frame spmv 12^3, fp64 and fp32")
    g, X = _problem(ctx)
    rng = np.random.default_rng(9)
    for which in (0, 1):
        M, i0, ln, c0 = _frame_matrix(ctx, X, which)
        n = int(ln.sum())
        nf = M.N
        # vectors span the frame up to the end of the last used slot
        assert nf == int((i0 + ln).max()) and nf <= X.J * 2 * X.Np
        x = rng.standard_normal(n)
        y0 = rng.standard_normal(n)
        ref = X.SpMV(1.7, x, 0.3, y0.copy(), which=which)
        xf = _to_frame(x, i0, ln, c0, nf, fill=np.nan)        # unused slots must not be read
        yf = _to_frame(y0, i0, ln, c0, nf, fill=np.nan)
        out = M.SpMV(1.7, xf, 0.3, yf)
        used = _to_frame(np.ones(n), i0, ln, c0, nf) > 0
        assert np.all(out[~used] == 0.0)
        if which == 1:
            # cell rows sum the species-unit blocks in another order: rounding level only
            assert rel(_from_frame(out, i0, ln, c0, n), ref) <= 1e-14
        else:
            assert np.array_equal(_from_frame(out, i0, ln, c0, n), ref)
        # FP32 on device vectors
        dx = B.DeviceArray.from_host(ctx, np.nan_to_num(xf, nan=7.0).astype(np.float32))
        dy = B.DeviceArray(ctx, nf, np.float32)
        M.SpMV(1.0, dx, 0.0, dy, prec=B.BH_FP32)
        dxc = B.DeviceArray.from_host(ctx, x.astype(np.float32))
        dyc = B.DeviceArray(ctx, n, np.float32)
        X.SpMV(1.0, dxc, 0.0, dyc, which=which, prec=B.BH_FP32)
        ctx.sync()
        o32 = dy.download()
        r32 = dyc.download()
        assert np.all(o32[~used] == 0.0)
        if which == 0:
            assert np.array_equal(_from_frame(o32.astype(np.float64), i0, ln, c0, n), r32.astype(np.float64))
        else:
            assert rel(_from_frame(o32.astype(np.float64), i0, ln, c0, n), r32.astype(np.float64)) <= 1e-5
        for a in (dx, dy, dxc, dyc):
            a.free()


def test_frame_solve_matches_library(ctx):
    _say("case xdg frame solve 12^3, FGMRES(30) + two-level FP32 preconditioner")
    g, X = _problem(ctx)
    X.set_rhs([1.0, 1.0], None)
    b = X.rhs(1)
    P = B.XdgPrecond(X, coarse_degree=1, prec=B.BH_FP32)
    S = B.XdgFlexGMRES(X, P, MaxKrylovDim=30, ConvergenceCriterion=1e-10)
    u = S.Solve(np.zeros(X.N), b)
    st = S.stats()
    M, i0, ln, c0 = _frame_matrix(ctx, X, 1)
    Pv = B.XdgPrecondV(g, M, X.Np, coarse_degree=1, prec=B.BH_FP32)
    Sv = B.FlexGMRES(M=M, Precond=Pv, MaxKrylovDim=30, ConvergenceCriterion=1e-10)
    bf = _to_frame(b, i0, ln, c0, M.N)
    uf = Sv.Solve(np.zeros(M.N), bf)
    sv = Sv.stats()
    _say(f"  library {st['iterations']} iterations, frame {sv['iterations']}")
    assert st["converged"] and sv["converged"]
    assert abs(st["iterations"] - sv["iterations"]) <= 1
    used = _to_frame(np.ones(X.N), i0, ln, c0, M.N) > 0
    assert np.all(uf[~used] == 0.0)
    A = X.scipy_matrix(1)
    uc = _from_frame(uf, i0, ln, c0, X.N)
    assert rel(A @ uc, b) <= 1.5e-10
    assert rel(uc, u) <= 1e-7


def test_coarse_matrix_of_the_preconditioner(ctx):
    """The preconditioner's coarse matrix, summed on the device from the species
    units (library matrix) or from the caller's cell blocks (frame matrix), equals
    R A R^T in numpy: R takes, per cell, the leading Np(1) modes of each present
    species block and adds them (species merged)."""
    _say("case xdg coarse matrix 12^3, unit blocks and caller's cell blocks")
    g, X = _problem(ctx)
    A = X.scipy_matrix(1).tocsr()
    i0, ln, c0 = X.dof_frame(1)
    Np, J = X.Np, X.J
    Npc = 4
    rows, cols = [], []
    for c in range(J):
        for a in range(ln[c] // Np):
            for m in range(Npc):
                rows.append(c * Npc + m)
                cols.append(c0[c] + a * Np + m)
    import scipy.sparse as sp
    R = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(J * Npc, X.N))
    ref = (R @ A @ R.T).toarray()
    scale = np.abs(ref).max()

    def dense(rp, col, blk):
        out = np.zeros((J * Npc, J * Npc))
        for r in range(J):
            for p in range(rp[r], rp[r + 1]):
                out[r * Npc:(r + 1) * Npc, col[p] * Npc:(col[p] + 1) * Npc] = blk[p]
        return out

    P = B.XdgPrecond(X, coarse_degree=1, prec=B.BH_FP32)
    rp, col, blk = B.xdg_precond_coarse_matrix(P)
    assert blk.shape[1:] == (Npc, Npc) and len(rp) == J + 1
    assert np.all(np.diff(col[rp[0]:rp[1]]) > 0)
    assert np.abs(dense(rp, col, blk) - ref).max() <= 1e-13 * scale
    M, _, _, _ = _frame_matrix(ctx, X, 1)
    Pv = B.XdgPrecondV(g, M, Np, coarse_degree=1, prec=B.BH_FP32)
    rpv, colv, blkv = B.xdg_precond_coarse_matrix(Pv)
    assert np.abs(dense(rpv, colv, blkv) - ref).max() <= 1e-13 * scale

"""GPU parity of the SIP operator path against the CPU oracle, through the C ABI.

Bars: index maps / sparsity bit-exact (integer work); FP64 operator apply and
assembled blocks within 1e-12 relative L2 (the north-star tolerance for the
FP64 drop-in); FP32 paths within 1e-5.
"""
import numpy as np
import pytest

import bosss_amd as B
import oracle as O
from tests.helpers import log_uniform_nu, nodes_2d, nodes_3d

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _grid_nodes(D, seed):
    if D == 2:
        return nodes_2d(7, 5, stretch=True, y0=-1.0, y1=1.0)
    return nodes_3d(5, 4, 3, jitter=0.4, seed=seed)


def _bnd_values(op_o, seed):
    return np.random.default_rng(seed).standard_normal(op_o.bnd_nodes().shape[0])


def _pair(ctx, D, k, seed, bc=None, affine=False):
    nodes = _grid_nodes(D, seed)
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, seed)
    gg = B.Grid(ctx, *nodes)
    oo = O.Sip(og, k, nu, bc=bc)
    go = B.SipOperator(gg, k, nu, bc=bc)
    if affine:
        gv = _bnd_values(oo, seed + 1)
        oo.set_bnd_values(gv)
        go.set_bnd_values(gv)
        f = np.random.default_rng(seed + 2).standard_normal(oo.N)
        oo.set_rhs(f)
        go.set_rhs(f)
    return og, gg, oo, go


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("D", [2, 3])
def test_grid_index_maps_bit_exact(ctx, D):
    _say(f"case grid maps D={D}")
    nodes = _grid_nodes(D, 3)
    og = O.Grid(*nodes)
    gg = B.Grid(ctx, *nodes)
    assert (og.J, og.E) == (gg.J, gg.E)
    for a, b in zip(og.edges(), gg.edges()):
        assert np.array_equal(a, b)
    assert np.allclose(og.cj(), gg.cj(), rtol=1e-15, atol=0)


CASES = [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)]


@pytest.mark.parametrize("D,k", CASES)
@pytest.mark.parametrize("affine", [False, True])
def test_evaluate_fp64_matches_oracle(ctx, D, k, affine):
    _say(f"case evaluate fp64 D={D} k={k} affine={affine}")
    bc = [0, 1] * D if affine else None
    og, gg, oo, go = _pair(ctx, D, k, 10 * D + k, bc=bc, affine=affine)
    u = np.random.default_rng(5).standard_normal(oo.N)
    ref = oo.evaluate(u)
    out = go.Evaluate(1.0, 0.0, np.zeros(go.N), u)
    assert rel(out, ref) < 1e-12
    # Evaluate(alpha, beta, output) semantics
    y0 = np.random.default_rng(6).standard_normal(oo.N)
    ref2 = oo.evaluate(u, alpha=-1.5, beta=0.25, y=y0.copy())
    out2 = go.Evaluate(-1.5, 0.25, y0.copy(), u)
    assert rel(out2, ref2) < 1e-12


MFMA64_CASES = [(2, 2), (2, 3), (3, 2), (3, 3)]


@pytest.mark.parametrize("D,k", MFMA64_CASES)
@pytest.mark.parametrize("affine", [False, True])
@pytest.mark.parametrize("path,waves,per_xcd", [("vector", 8, 64), ("modal", 8, 64), ("mfma", 4, 1), ("mfma", 8, 64),
                                               ("tiled", 2, 0), ("tiled", 4, 0), ("tiled", 8, 0)])
def test_evaluate_fp64_paths_match_oracle(ctx, D, k, affine, path, waves, per_xcd):
    """Both FP64 kernels (vector and FP64-MFMA, each pinned, several launch
    shapes) within 1e-12 of the oracle, with Neumann sides, the affine part and
    Evaluate(alpha, beta)."""
    _say(f"case evaluate fp64 path={path} waves={waves} per_xcd={per_xcd} D={D} k={k} affine={affine}")
    bc = [0, 1] * D if affine else None
    og, gg, oo, go = _pair(ctx, D, k, 60 + 10 * D + k, bc=bc, affine=affine)
    go.force_fp64_path(path, waves, per_xcd)
    u = np.random.default_rng(15).standard_normal(oo.N)
    assert rel(go.Evaluate(1.0, 0.0, np.zeros(go.N), u), oo.evaluate(u)) < 1e-12
    y0 = np.random.default_rng(16).standard_normal(oo.N)
    ref2 = oo.evaluate(u, alpha=-1.5, beta=0.25, y=y0.copy())
    assert rel(go.Evaluate(-1.5, 0.25, y0.copy(), u), ref2) < 1e-12
    assert go.fp64_tuning()[0] == path


def test_evaluate_fp64_autotuned_ragged(ctx):
    """Ragged 3D grid (13 x 11 x 9 cells, k = 3: last 16-cell tile partial, XCD
    ranges uneven): the autotuned FP64 apply equals the vector kernel to 1e-13
    and the oracle to 1e-12."""
    _say("case evaluate fp64 autotuned ragged 13x11x9 k=3")
    rng = np.random.default_rng(17)
    nodes = [np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for n in (13, 11, 9)]
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, 17)
    gg = B.Grid(ctx, *nodes)
    oo = O.Sip(og, 3, nu)
    go = B.SipOperator(gg, 3, nu)
    u = rng.standard_normal(oo.N)
    auto = go.Evaluate(1.0, 0.0, np.zeros(go.N), u, affine=False)
    path, waves, per_xcd, ms_v, ms_m = go.fp64_tuning()
    _say(f"  tuned: {path} waves={waves} per_xcd={per_xcd} vector {ms_v:.4f} ms, mfma {ms_m:.4f} ms")
    assert path in ("vector", "modal", "tiled", "mfma") and ms_v > 0 and ms_m > 0
    go.force_fp64_path("vector")
    vec = go.Evaluate(1.0, 0.0, np.zeros(go.N), u, affine=False)
    assert rel(auto, vec) < 1e-13
    assert rel(auto, oo.evaluate(u)) < 1e-12


@pytest.mark.parametrize("D,k", [(2, 1), (2, 2), (2, 3), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_evaluate_modal_faces(ctx, D, k):
    """The vector kernel with the face terms in modal form equals the one that
    evaluates the face rule node by node to 1e-13 and the oracle to 1e-12 in FP64
    (ragged stretched grid, Neumann sides, affine part, Evaluate(alpha, beta)), and
    its FP32 apply is within 1e-5 of the FP64 result."""
    _say(f"case evaluate modal faces D={D} k={k}")
    rng = np.random.default_rng(80 + 10 * D + k)
    cells = (23, 17) if D == 2 else (13, 11, 9)
    nodes = [np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for n in cells]
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, 81)
    gg = B.Grid(ctx, *nodes)
    bc = [0, 1] * D
    oo = O.Sip(og, k, nu, bc=bc)
    go = B.SipOperator(gg, k, nu, bc=bc)
    gv = _bnd_values(oo, 82)
    oo.set_bnd_values(gv)
    go.set_bnd_values(gv)
    u = rng.standard_normal(oo.N)
    y0 = rng.standard_normal(oo.N)
    go.force_fp64_path("vector")
    node = go.Evaluate(-1.5, 0.25, y0.copy(), u)
    go.force_fp64_path("modal")
    modal = go.Evaluate(-1.5, 0.25, y0.copy(), u)
    assert go.fp64_tuning()[0] == "modal"
    assert rel(modal, node) < 1e-13
    assert rel(modal, oo.evaluate(u, alpha=-1.5, beta=0.25, y=y0.copy())) < 1e-12
    ud = B.DeviceArray.from_host(ctx, u.astype(np.float32))
    yd = B.DeviceArray(ctx, go.N, np.float32)
    go.Evaluate(1.0, 0.0, yd, ud, affine=False, prec=B.BH_FP32)
    ref = go.Evaluate(1.0, 0.0, np.zeros(go.N), u.astype(np.float32).astype(np.float64), affine=False)
    assert rel(yd.download().astype(np.float64), ref) < 1e-5


@pytest.mark.parametrize("D,k", CASES)
def test_evaluate_tiled_long_rows(ctx, D, k):
    """The tiled modal kernel (neighbour traces from the adjacent lanes and, through
    LDS, from the adjacent waves) on ragged stretched grids whose rows of consecutive
    cells are longer than one wave (70 and 133 cells: the cells just before and after
    a wave come from the halo lanes) and whose row count is not a multiple of the
    waves per workgroup: FP64 equals the node-by-node kernel to 1e-13 and the oracle to
    1e-12 for 2, 4 and 8 waves per workgroup, with Neumann sides, the affine part and
    Evaluate(alpha, beta); the FP32 tiled apply is within 1e-5 of FP64."""
    _say(f"case evaluate tiled long rows D={D} k={k}")
    rng = np.random.default_rng(120 + 10 * D + k)
    cells = (11, 133) if D == 2 else (70, 9, 3)
    nodes = [np.cumsum(np.r_[0.0, rng.uniform(0.5, 1.5, n)]) for n in cells]
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, 121)
    gg = B.Grid(ctx, *nodes)
    bc = [1, 0] * D
    oo = O.Sip(og, k, nu, bc=bc)
    go = B.SipOperator(gg, k, nu, bc=bc)
    gv = _bnd_values(oo, 122)
    oo.set_bnd_values(gv)
    go.set_bnd_values(gv)
    u = rng.standard_normal(oo.N)
    y0 = rng.standard_normal(oo.N)
    ref = oo.evaluate(u, alpha=-1.5, beta=0.25, y=y0.copy())
    go.force_fp64_path("vector")
    node = go.Evaluate(-1.5, 0.25, y0.copy(), u)
    for w in (2, 4, 8):
        go.force_fp64_path("tiled", waves=w)
        tiled = go.Evaluate(-1.5, 0.25, y0.copy(), u)
        assert go.fp64_tuning()[:2] == ("tiled", w)
        assert rel(tiled, node) < 1e-13, w
        assert rel(tiled, ref) < 1e-12, w
        go.force_fp32_path("tiled", waves=w)
        ud = B.DeviceArray.from_host(ctx, u.astype(np.float32))
        yd = B.DeviceArray(ctx, go.N, np.float32)
        go.Evaluate(1.0, 0.0, yd, ud, affine=False, prec=B.BH_FP32)
        ref32 = go.Evaluate(1.0, 0.0, np.zeros(go.N), u.astype(np.float32).astype(np.float64), affine=False)
        assert rel(yd.download().astype(np.float64), ref32) < 1e-5, w


@pytest.mark.parametrize("D,k", [(2, 2), (3, 3)])
def test_evaluate_device_pointers(ctx, D, k):
    _say(f"case evaluate device pointers D={D} k={k}")
    og, gg, oo, go = _pair(ctx, D, k, 77)
    u = np.random.default_rng(8).standard_normal(oo.N)
    du = B.DeviceArray.from_host(ctx, u)
    dy = B.DeviceArray(ctx, go.N)
    go.Evaluate(1.0, 0.0, dy, du, affine=False)
    ctx.sync()
    assert rel(dy.download(), oo.evaluate(u)) < 1e-12


@pytest.mark.parametrize("D,k", [(2, 2), (3, 2), (3, 3)])
def test_evaluate_fp32(ctx, D, k):
    _say(f"case evaluate fp32 D={D} k={k}")
    og, gg, oo, go = _pair(ctx, D, k, 31)
    u = np.random.default_rng(9).standard_normal(oo.N)
    out = go.Evaluate(1.0, 0.0, np.zeros(go.N), u, affine=False, prec=B.BH_FP32)
    assert rel(out, oo.evaluate(u)) < 1e-5


MFMA_CASES = [(2, 2), (2, 3), (3, 1), (3, 2), (3, 3)]


@pytest.mark.parametrize("D,k", MFMA_CASES)
@pytest.mark.parametrize("affine", [False, True])
def test_evaluate_fp16_mfma(ctx, D, k, affine):
    """FP16 MFMA contractions with FP32 accumulation: within 2e-3 relative L2 of
    the FP64 oracle (FP16 tables and per-cell scaled FP16 operands carry a
    relative rounding of 4.9e-4 each), including Neumann sides, the affine part
    and the Evaluate(alpha, beta) semantics."""
    _say(f"case evaluate fp16-mfma D={D} k={k} affine={affine}")
    bc = [0, 1] * D if affine else None
    og, gg, oo, go = _pair(ctx, D, k, 40 + 10 * D + k, bc=bc, affine=affine)
    u = np.random.default_rng(11).standard_normal(oo.N)
    out = go.Evaluate(1.0, 0.0, np.zeros(go.N), u, affine=affine, prec=B.BH_FP16_MFMA)
    assert rel(out, oo.evaluate(u) if affine else oo.evaluate(u) - oo.evaluate(np.zeros(oo.N))) < 2e-3
    y0 = np.random.default_rng(12).standard_normal(oo.N)
    ref2 = oo.evaluate(u, alpha=-1.5, beta=0.25, y=y0.copy())
    if not affine:
        ref2 = ref2 + 1.5 * oo.evaluate(np.zeros(oo.N))
    out2 = go.Evaluate(-1.5, 0.25, y0.copy(), u, affine=affine, prec=B.BH_FP16_MFMA)
    assert rel(out2, ref2) < 2e-3


def test_evaluate_fp16_mfma_full_tiles(ctx):
    """16^3 cells, k = 3 (256 tiles over several workgroups): FP16-MFMA apply of a
    smooth and a rough vector vs the FP64 apply of the same operator."""
    _say("case evaluate fp16-mfma 16^3 k=3")
    x = np.linspace(0.0, 1.0, 17)
    gg = B.Grid(ctx, x, x, x)
    nu = log_uniform_nu(gg.J, 4)
    go = B.SipOperator(gg, 3, nu)
    rng = np.random.default_rng(13)
    for u in (rng.standard_normal(go.N), np.repeat(rng.standard_normal(gg.J), go.Np) * (np.arange(go.N) % go.Np == 0)):
        ref = go.Evaluate(1.0, 0.0, np.zeros(go.N), u, affine=False)
        du = B.DeviceArray.from_host(ctx, u.astype(np.float32))
        dy = B.DeviceArray(ctx, go.N, np.float32)
        go.Evaluate(1.0, 0.0, dy, du, affine=False, prec=B.BH_FP16_MFMA)
        ctx.sync()
        assert rel(dy.download().astype(np.float64), ref) < 2e-3


@pytest.mark.parametrize("D,k", CASES)
def test_assembly_matches_oracle(ctx, D, k):
    _say(f"case assembly D={D} k={k}")
    bc = [1, 0] * D
    og, gg, oo, go = _pair(ctx, D, k, 50 + D + k, bc=bc, affine=True)
    Ao, affo = oo.assemble()
    Ag, affg = go.ComputeMatrix()
    rpo, colo, valo = Ao.export()
    rpg, colg, valg = Ag.export()
    assert np.array_equal(rpo, rpg) and np.array_equal(colo, colg)
    assert rel(valg.ravel(), valo.ravel()) < 1e-12
    assert rel(affg, affo) < 1e-12


@pytest.mark.parametrize("D,k", [(2, 2), (3, 3)])
def test_spmv_matches_oracle(ctx, D, k):
    _say(f"case spmv D={D} k={k}")
    og, gg, oo, go = _pair(ctx, D, k, 90)
    Ao, _ = oo.assemble()
    rp, col, val = Ao.export()
    Ag = B.BlockMsrMatrix(ctx, nrows=Ao.nrows, bs=Ao.bs, rowptr=rp, col=col, vals=val)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(oo.N)
    y0 = rng.standard_normal(oo.N)
    ref = Ao.spmv(0.5, x, -2.0, y0.copy())
    out = Ag.SpMV(0.5, x, -2.0, y0.copy())
    assert rel(out, ref) < 1e-13
    # FP32 on device vectors
    dx = B.DeviceArray.from_host(ctx, x.astype(np.float32))
    dy = B.DeviceArray(ctx, oo.N, np.float32)
    Ag.SpMV(1.0, dx, 0.0, dy, prec=B.BH_FP32)
    ctx.sync()
    assert rel(dy.download().astype(np.float64), Ao.spmv(1.0, x, 0.0, np.zeros(oo.N))) < 1e-5


@pytest.mark.parametrize("prec", [B.BH_FP64, B.BH_FP32])
def test_block_jacobi_matches_oracle(ctx, prec):
    _say(f"case block jacobi prec={prec}")
    og, gg, oo, go = _pair(ctx, 3, 2, 123)
    Ao, _ = oo.assemble()
    Ag, _ = go.ComputeMatrix()
    b = np.random.default_rng(1).standard_normal(oo.N)
    xo = Ao.bjacobi(np.zeros(oo.N), b, omega=0.6, iters=3)
    bj = B.BlockJacobi(Ag, omega=0.6, NoOfIterations=3, prec=prec)
    xg = bj.Solve(np.zeros(oo.N), b)
    assert rel(xg, xo) < (1e-12 if prec == B.BH_FP64 else 1e-4)


def test_soft_gmres_block_jacobi_matches_oracle(ctx):
    """GMRES(30) + block-Jacobi (the first configuration's solver) on a small grid."""
    _say("case soft gmres")
    nodes = nodes_2d(12, 12)
    og = O.Grid(*nodes)
    gg = B.Grid(ctx, *nodes)
    oo = O.Sip(og, 2, 1.0)
    go = B.SipOperator(gg, 2, 1.0)
    Ao, _ = oo.assemble()
    Ag, _ = go.ComputeMatrix()
    b = np.random.default_rng(2).standard_normal(oo.N)
    xo, st = Ao.gmres(b, precond=1, restart=30, tol=1e-10)
    bj = B.BlockJacobi(Ag, omega=1.0, NoOfIterations=1)
    gm = B.SoftGMRES(M=Ag, Precond=bj, MaxKrylovDim=30, ConvergenceCriterion=1e-10)
    xg = gm.Solve(np.zeros(oo.N), b)
    sg = gm.stats()
    assert st["converged"] and sg["converged"]
    assert abs(sg["iterations"] - st["iters"]) <= 2
    assert rel(xg, xo) < 1e-7


@pytest.mark.parametrize("matrix_free,restart", [(False, 30), (True, 30), (True, 45)])
def test_flex_gmres_true_residual(ctx, matrix_free, restart):
    """FGMRES + block-Jacobi to 1e-10 true residual.  Restart 30 runs the fused
    Gram-Schmidt passes for every step; restart 45 also the path for more than
    31 basis vectors (separate passes)."""
    _say(f"case flex gmres matrix_free={matrix_free} restart={restart}")
    nodes = nodes_3d(6, 6, 6)
    og = O.Grid(*nodes)
    nu = log_uniform_nu(og.J, 4)
    gg = B.Grid(ctx, *nodes)
    oo = O.Sip(og, 2, nu)
    go = B.SipOperator(gg, 2, nu)
    Ag, _ = go.ComputeMatrix()
    b = np.random.default_rng(3).standard_normal(oo.N)
    bj = B.BlockJacobi(Ag, omega=1.0, NoOfIterations=1)
    fg = B.FlexGMRES(M=None if matrix_free else Ag, op=go if matrix_free else None, Precond=bj,
                     MaxKrylovDim=restart, ConvergenceCriterion=1e-10)
    x = fg.Solve(np.zeros(oo.N), b)
    st = fg.stats()
    _say("  stats", st)
    assert st["converged"]
    assert restart <= 31 or st["iterations"] > 32
    r = b - oo.evaluate(x)  # oracle operator: no affine part set
    assert np.linalg.norm(r) <= 1e-10 * np.linalg.norm(b) * 1.0001


def test_full_size_properties(ctx):
    """3D 64^3 k=3 (5.24M DOF): symmetry <Au,v> = <u,Av> and matrix-free ==
    assembled SpMV, both size-independent properties of the operator."""
    _say("case full size 64^3 k=3")
    n = 64
    x = np.linspace(0, 1, n + 1)
    gg = B.Grid(ctx, x, x, x)
    nu = log_uniform_nu(gg.J, 2024)
    go = B.SipOperator(gg, 3, nu)
    rng = np.random.default_rng(0)
    u = rng.standard_normal(go.N)
    v = rng.standard_normal(go.N)
    Au = go.Evaluate(1.0, 0.0, np.zeros(go.N), u)
    Av = go.Evaluate(1.0, 0.0, np.zeros(go.N), v)
    a, b = np.dot(Au, v), np.dot(u, Av)
    assert abs(a - b) <= 1e-12 * np.linalg.norm(Au) * np.linalg.norm(v)
    Ag, _ = go.ComputeMatrix()
    y = Ag.SpMV(1.0, u, 0.0, np.zeros(go.N))
    assert rel(y, Au) < 1e-12
    del Ag

"""Benchmark: DG operator apply (DOF/s) and time-to-solve on one MI355X.

Workload: 3D SIP Poisson on a 64^3 Cartesian hex grid on [0,1]^3, k = 3
(20 modes/cell, 5,242,880 DOF), per-cell diffusion log-uniform in [1e-2, 1e2]
(seeded).  A "step" is one FP64 matrix-free operator apply (Evaluate) of a
seeded N(0,1) DOF vector resident in HBM.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-solve]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12      # MI355X dense FP64 (vector and matrix), FLOP/s
PEAK_HBM = 8.0e12        # MI355X HBM3E, B/s
USE_GRAPHS = True        # HIP-graph replay of the multigrid V-cycles (--no-graphs: eager launches)


def sip_flops(n, D, k):
    """Algorithmic FLOPs of one matrix-free SIP apply: the multiply-adds the
    exact quadrature needs on the orthonormal Legendre tensor basis, without its
    structural zeros.  Volume: 2 per nonzero of the per-direction reference
    stiffness Kv[d] (the volume rule of an axis-aligned box factorises; Kv[d][m][n]
    is nonzero only for equal exponents outside d and exponents >= 1 of equal
    parity in d).  Faces of direction d: 8*Nqf*(Np + nd) (value and normal
    derivative on both sides, and the contraction back; nd modes have a nonzero
    d-derivative)."""
    from math import comb
    Np = comb(k + D, D)
    Nqf = (k + 1) ** (D - 1)
    J = n ** D
    ex = []
    for deg in range(k + 1):
        if D == 2:
            ex += [(a, deg - a) for a in range(deg + 1)]
        else:
            ex += [(a, b, deg - a - b) for a in range(deg + 1) for b in range(deg - a + 1)]
    assert len(ex) == Np
    vol = 0
    faces = 0
    Fd = (n - 1) * n ** (D - 1) + 2 * n ** (D - 1)     # faces per direction, boundary included
    for d in range(D):
        for em in ex:
            for en in ex:
                if em[d] >= 1 and en[d] >= 1 and (em[d] + en[d]) % 2 == 0 and all(em[e] == en[e] for e in range(D) if e != d):
                    vol += 2
        nd = sum(1 for e in ex if e[d] >= 1)
        faces += Fd * 8 * Nqf * (Np + nd)
    return J * vol + faces


def sip_flops_modal(n, D, k):
    """Algorithmic FLOPs of the apply in modal face form: the volume as in
    sip_flops; per face 8*(Np + nd) (value and d-derivative traces of both cells
    along the normal, and the contraction back) plus 24 per face mode e' (the
    flux terms of both sides), Nf = C(k + D - 1, D - 1) face modes."""
    from math import comb
    Np = comb(k + D, D)
    Nf = comb(k + D - 1, D - 1)
    J = n ** D
    ex = []
    for deg in range(k + 1):
        if D == 2:
            ex += [(a, deg - a) for a in range(deg + 1)]
        else:
            ex += [(a, b, deg - a - b) for a in range(deg + 1) for b in range(deg - a + 1)]
    vol = 0
    faces = 0
    Fd = (n - 1) * n ** (D - 1) + 2 * n ** (D - 1)
    for d in range(D):
        for em in ex:
            for en in ex:
                if em[d] >= 1 and en[d] >= 1 and (em[d] + en[d]) % 2 == 0 and all(em[e] == en[e] for e in range(D) if e != d):
                    vol += 2
        nd = sum(1 for e in ex if e[d] >= 1)
        faces += Fd * (8 * (Np + nd) + 24 * Nf)
    return J * vol + faces


def legendre_1d(k, xi):
    """orthonormal 1D Legendre sqrt(n+1/2) P_n at xi, shape (len(xi), k+1)."""
    P = np.zeros((len(xi), k + 1))
    P[:, 0] = 1.0
    if k > 0:
        P[:, 1] = xi
    for n in range(2, k + 1):
        P[:, n] = ((2 * n - 1) * xi * P[:, n - 1] - (n - 1) * P[:, n - 2]) / n
    return P * np.sqrt(np.arange(k + 1) + 0.5)


def sine_rhs_coeffs(n, k, seed, nmodes=16, kmax=6):
    """DG coefficients (cell-major, reference mode order) of
    f = sum_m a_m prod_d sin(pi w_md x_d), a_m ~ N(0,1), w_md in 1..kmax, on the
    uniform n^3 grid of [0,1]^3: the projection is separable per direction."""
    from math import comb
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(nmodes)
    w = rng.integers(1, kmax + 1, (nmodes, 3))
    xg, wg = np.polynomial.legendre.leggauss(8)
    L = legendre_1d(k, xg)                              # (8, k+1)
    h = 1.0 / n
    lo = np.arange(n) * h
    X = lo[:, None] + 0.5 * (xg[None, :] + 1.0) * h     # (n, 8)
    # 1D integrals S[m, d, i, e] = int_cell sin(pi w x) psi_e / sqrt(h/2) dx
    S = np.einsum("mdiq,q,qe->mdie", np.sin(np.pi * w[:, :, None, None] * X[None, None]), wg, L) * np.sqrt(h / 2)
    ex = [(a_, b_, d - a_ - b_) for d in range(k + 1) for a_ in range(d + 1) for b_ in range(d - a_ + 1)]
    ex = np.array(ex)                                   # (Np, 3) exponents (x, y, z)
    Np = comb(k + 3, 3)
    assert len(ex) == Np
    # cell c = i + j*n + kk*n*n
    C = np.zeros((n, n, n, Np))                         # [kk, j, i, mode]
    for m in range(nmodes):
        Sx = S[m, 0][:, ex[:, 0]]                       # (n_i, Np)
        Sy = S[m, 1][:, ex[:, 1]]
        Sz = S[m, 2][:, ex[:, 2]]
        C += a[m] * Sz[:, None, None, :] * Sy[None, :, None, :] * Sx[None, None, :, :]
    return C.reshape(-1, Np).ravel()


def lognormal_nu(n, seed, sigma=1.5, nmodes=16, ell=0.1):
    """Per-cell diffusion nu = exp(sigma g) on the uniform n^3 grid of [0,1]^3 (cell
    c = i + j n + kk n n, values at the cell centres): g = sqrt(2 / M) sum_m
    cos(2 pi k_m . x + phi_m), M = 16 seeded modes with k_m ~ N(0, 1/(2 pi ell)^2)
    per component (the spectral density of a Gaussian covariance of correlation
    length ell = 0.1) and phi_m ~ U(0, 2 pi), so that g has zero mean and unit
    variance over the random phases: log-normal nu with sigma = 1.5."""
    rng = np.random.default_rng(seed)
    kv = rng.normal(0.0, 1.0 / (2.0 * np.pi * ell), (nmodes, 3))
    ph = rng.uniform(0.0, 2.0 * np.pi, nmodes)
    xc = (np.arange(n) + 0.5) / n
    e = [np.exp(2j * np.pi * kv[:, d, None] * xc[None, :]) for d in range(3)]      # [m, cell index]
    z = np.exp(1j * ph)[:, None, None, None] * e[2][:, :, None, None] * e[1][:, None, :, None] * e[0][:, None, None, :]
    g = np.sqrt(2.0 / nmodes) * z.real.sum(0)                                       # [kk, j, i]
    return np.exp(sigma * g).ravel()


def solve_leg(B, ctx, grid, op, n, k, out):
    """time-to-solve: FP64 FGMRES to 1e-8 relative residual with the p/h-multigrid
    K-cycle preconditioner: the FP64 drop-in on the assembled matrix (plain), the
    FP64 matrix-free variant, and the mixed-precision variants (FP32 multigrid;
    FP16-MFMA level 0; with the learned correction)."""
    N = op.N
    f = sine_rhs_coeffs(n, k, seed=16)
    b = B.DeviceArray.from_host(ctx, f)
    t0 = time.perf_counter()
    M, _ = op.ComputeMatrix()
    ctx.sync()
    t_asm = time.perf_counter() - t0
    res = {"matrix_assembly_s": t_asm}
    # BlockMsrMatrix SpMV of the assembled operator (20x20 blocks, HBM-bound)
    nblk = M.nnzb
    rng = np.random.default_rng(7)
    for name, prec, dt in (("bsr_spmv_fp64", B.BH_FP64, np.float64), ("bsr_spmv_fp32", B.BH_FP32, np.float32)):
        xv = B.DeviceArray.from_host(ctx, rng.standard_normal(N).astype(dt))
        yv = B.DeviceArray(ctx, N, dt)
        for _ in range(3):
            M.SpMV(1.0, xv, 0.0, yv, prec=prec)
        ctx.sync()
        reps = 20
        ctx.record(2)
        for _ in range(reps):
            M.SpMV(1.0, xv, 0.0, yv, prec=prec)
        ctx.record(3)
        ctx.sync()
        ms = ctx.elapsed_ms(2, 3) / reps
        es = np.dtype(dt).itemsize
        nbytes = nblk * M.bs * M.bs * es + nblk * 4 + (M.nrows + 1) * 8 + 2 * N * es
        res[name] = {"ms": ms, "algorithmic_bytes": int(nbytes), "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "hbm_frac": nbytes / (ms * 1e-3) / PEAK_HBM}
        del xv, yv
    # one multigrid configuration for every variant (the fastest one found for both
    # precisions: single pre/post block-Jacobi sweep, 6-step K-cycle on the k = 1 level)
    mgp = dict(degrees=[k, 1], min_coarse=512, omega=0.8, pre=1, post=1, kcycle_inner=6)
    res["multigrid"] = dict(mgp)
    net = None
    variants = (
        # the drop-in: assembled BlockMsrMatrix SpMV for the Krylov operator and every
        # multigrid residual, FP64 throughout (the reference's method on the GPU)
        ("plain_fp64", B.BH_FP64, False, False),
        # same precision, matrix-free operator applies (FP64 MFMA on level 0, k = 1 kernel on level 1)
        ("fp64_matrix_free", B.BH_FP64, True, False),
        ("mixed_fp32", B.BH_FP32, True, False),
        ("mixed_fp16_mfma", B.BH_FP16_MFMA, True, False),
        ("mixed_fp16_mfma_learned", B.BH_FP16_MFMA, True, True),
    )
    for name, prec, mfree, use_net in variants:
        t0 = time.perf_counter()
        mg = B.PMultigrid(op, M=M, prec=prec, matrix_free=mfree, **mgp)
        B.pmg_set_graphs(mg, USE_GRAPHS)
        if use_net:
            t_mg = time.perf_counter() - t0
            net, res["learned_training"] = train_network(B, op, mg, n, k, steps=400, seed0=3000)
            t0 = time.perf_counter() - t_mg
            B.pmg_set_learned(mg, net)
        ctx.sync()
        t_init = time.perf_counter() - t0
        fg = (B.FlexGMRES(op=op, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=500)
              if mfree else
              B.FlexGMRES(M=M, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=500))
        x = B.DeviceArray.from_host(ctx, np.zeros(N))
        t_run = 1e30
        for rep in range(2):   # the first solve also captures the graphs and runs the autotuners
            x.upload(np.zeros(N))
            ctx.sync()
            t0 = time.perf_counter()
            fg.Solve(x, b)
            ctx.sync()
            t_run = min(t_run, time.perf_counter() - t0)
        st = fg.stats()
        # independent FP64 check of the true residual
        r = B.DeviceArray(ctx, N)
        op.Evaluate(1.0, 0.0, r, x, affine=False)
        rr = f - r.download()
        relres = float(np.linalg.norm(rr) / np.linalg.norm(f))
        res[name] = {"solver_init_s": t_init, "solver_run_s": t_run, "iterations": st["iterations"],
                     "converged": st["converged"], "true_rel_residual": relres}
        if use_net:
            acc, rej = B.pmg_learned_stats(mg)
            active, gain = B.pmg_learned_info(mg)
            res[name].update({"learned_steps_taken": acc, "learned_steps_declined": rej,
                              "learned_still_active": active,
                              "learned_mean_removed_residual_sq_fraction": gain,
                              "learned_probe_costs": B.pmg_learned_costs(mg),
                              "learned_fallback": "probe 8 calls; switch off below 1e-2 mean removed |r|^2 fraction "
                                                  "or when the step's log-reduction per ms is below the cycle's"})
        del fg, mg, x, r
    base = res["plain_fp64"]["solver_run_s"]
    base_all = base + res["plain_fp64"]["solver_init_s"]
    for name, _, _, _ in variants[1:]:
        res["speedup_" + name + "_vs_plain"] = base / res[name]["solver_run_s"]
        # with the multigrid setup (hierarchy, Galerkin products, smoother and coarse inverses) counted
        res["speedup_" + name + "_vs_plain_incl_setup"] = base_all / (res[name]["solver_run_s"] + res[name]["solver_init_s"])
    res["speedup_mixed_fp16_mfma_vs_fp64_matrix_free"] = (res["fp64_matrix_free"]["solver_run_s"] /
                                                          res["mixed_fp16_mfma"]["solver_run_s"])
    out["time_to_solve"] = res


def train_network(B, op, mg, n, k, steps, seed0):
    """Per-cell correction network for op's multigrid, trained online on the
    post-cycle residuals of multigrid iterations over fresh sine right-hand sides
    (seeds from seed0; 8 iterations per right-hand side), minimising the residual
    the correction leaves (bosss_amd.learned.train_resmin)."""
    from bosss_amd import learned
    t0 = time.perf_counter()
    net = learned.network_for(op, seed=1)
    hist, t_train, nrhs = learned.train_resmin(net, op, mg, lambda i: sine_rhs_coeffs(n, k, seed0 + i), steps=steps)
    info = {"residual_pairs": len(hist), "rhs": nrhs, "steps": steps,
            "loss": "min_a |r' - a A e|^2 / |r'|^2 (residual fraction the optimally scaled correction leaves)",
            "loss_first10": float(np.mean(hist[:10])), "loss_last50": float(np.mean(hist[-50:])),
            "train_s": t_train, "total_s": time.perf_counter() - t0}
    return net, info


def learned_leg(B, ctx, out, n=64, k=2, steps=2000, batch=32, nu_seed=5):
    """Learned correction workload: 3D SIP Poisson, 64^3 hexes, k = 2 (2.62M DOF),
    log-normal nu (sigma 1.5, 16 Fourier modes, correlation length 0.1).  The network
    (16 + 6 inputs over the face adjacency, 128-128 hidden, 10 outputs, BF16 weights /
    FP32 accumulation) is trained on 2,000 post-cycle residuals of plain multigrid
    iterations over fresh right-hand sides, then a batch of 32 fresh right-hand sides
    is solved to 1e-8 (true FP64 residual) with FGMRES(30) and the FP32 multigrid,
    plain and with the learned step."""
    x = np.linspace(0.0, 1.0, n + 1)
    g = B.Grid(ctx, x, x, x)
    nu = lognormal_nu(n, nu_seed)
    op = B.SipOperator(g, k, nu)
    N = op.N
    mgp = dict(degrees=[k, 1], min_coarse=512, omega=0.8, pre=1, post=1, kcycle_inner=6)
    mg = B.PMultigrid(op, prec=B.BH_FP32, matrix_free=True, **mgp)
    B.pmg_set_graphs(mg, USE_GRAPHS)
    res = {"grid": f"{n}^3", "k": k, "dof": N, "nu": "log-normal, sigma 1.5, 16 Fourier modes, correlation length 0.1",
           "nu_range": [float(nu.min()), float(nu.max())], "multigrid": dict(mgp, prec="fp32"), "batch": batch}
    net, res["training"] = train_network(B, op, mg, n, k, steps, seed0=100000)
    rhs = [sine_rhs_coeffs(n, k, 200000 + i) for i in range(batch)]
    xd = B.DeviceArray(ctx, N)
    bd = B.DeviceArray(ctx, N)
    yd = B.DeviceArray(ctx, N)
    for name, use in (("plain", False), ("learned", True)):
        B.pmg_set_learned(mg, net if use else None)
        fg = B.FlexGMRES(op=op, Precond=mg, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=500)
        bd.upload(rhs[0])                         # warm-up solve (graph capture)
        xd.upload(np.zeros(N))
        fg.Solve(xd, bd)
        ctx.sync()
        tot_t, tot_it, worst = 0.0, 0, 0.0
        for b in rhs:
            bd.upload(b)
            xd.upload(np.zeros(N))
            ctx.sync()
            t0 = time.perf_counter()
            fg.Solve(xd, bd)
            ctx.sync()
            tot_t += time.perf_counter() - t0
            st = fg.stats()
            tot_it += st["iterations"]
            yd.copy_from(bd)
            op.Evaluate(-1.0, 1.0, yd, xd, affine=False)
            worst = max(worst, float(np.sqrt(B.dot(ctx, yd, yd)) / np.linalg.norm(b)))
        r = {"batch_solve_s": tot_t, "iterations": tot_it, "worst_true_rel_residual": worst}
        if use:
            acc, rej = B.pmg_learned_stats(mg)
            active, gain = B.pmg_learned_info(mg)
            r.update({"learned_accepted": acc, "learned_rejected": rej, "learned_still_active": active,
                      "learned_mean_removed_residual_sq_fraction": gain,
                      "learned_probe_costs": B.pmg_learned_costs(mg)})
        res[name] = r
        del fg
    B.pmg_set_learned(mg, None)
    res["speedup_learned_vs_plain"] = res["plain"]["batch_solve_s"] / res["learned"]["batch_solve_s"]
    out["learned_correction"] = res
    del mg, op, net


def measured_traffic(n, k, kernel_prefix):
    """HBM bytes per launch (read + write) of a kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (counters scaled by a copy
    kernel of known size); None if not profiled for this workload."""
    path = os.path.join(ROOT, "profiles", f"sip_apply_k{k}_{n}cubed_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    for name, v in d["kernels"].items():
        if kernel_prefix in name:
            return v["read_bytes"] + v["write_bytes"]
    return None


def sip2d_leg(B, ctx, out, n=64, k=2, seed=1):
    """2D SIP Poisson on uniform 64 x 64 quads, k = 2 (24,576 DOF), nu = 1, Dirichlet 0,
    RHS = projection of 8 seeded sine modes: the FP64 Evaluate rate (the vectors fit in
    L2, so no roofline is claimed), device assembly, BlockMsrMatrix SpMV and the solve
    with SoftGMRES(30) + block-Jacobi to 1e-10 under the reference's termination."""
    from math import comb
    x = np.linspace(0.0, 1.0, n + 1)
    g = B.Grid(ctx, x, x)
    op = B.SipOperator(g, k, 1.0)
    N = op.N
    Np = comb(k + 2, 2)
    # L2 projection of the sines: separable per direction (cell i = x index, j = y index)
    rng = np.random.default_rng(seed)
    xg, wg = np.polynomial.legendre.leggauss(8)
    L = legendre_1d(k, xg)
    h = 1.0 / n
    X = (np.arange(n) * h)[:, None] + 0.5 * (xg[None, :] + 1.0) * h
    ex = np.array([(a_, d - a_) for d in range(k + 1) for a_ in range(d + 1)])
    f = np.zeros((n, n, Np))
    for _ in range(8):
        a = rng.standard_normal()
        w = rng.integers(1, 7, 2)
        Sx = (np.sin(np.pi * w[0] * X) * wg) @ L * np.sqrt(h / 2)      # [i, e]
        Sy = (np.sin(np.pi * w[1] * X) * wg) @ L * np.sqrt(h / 2)      # [j, e]
        f += a * Sx[:, None, ex[:, 0]] * Sy[None, :, ex[:, 1]]
    op.set_rhs(f.ravel())
    u = B.DeviceArray.from_host(ctx, np.random.default_rng(seed + 1).standard_normal(N))
    y = B.DeviceArray(ctx, N)
    for _ in range(5):
        op.Evaluate(1.0, 0.0, y, u)
    ctx.sync()
    reps = 200
    ctx.record(2)
    for _ in range(reps):
        op.Evaluate(1.0, 0.0, y, u)
    ctx.record(3)
    ctx.sync()
    ms = ctx.elapsed_ms(2, 3) / reps
    t0 = time.perf_counter()
    M, aff = op.ComputeMatrix()
    ctx.sync()
    t_asm = time.perf_counter() - t0
    for _ in range(5):
        M.SpMV(1.0, u, 0.0, y)
    ctx.sync()
    ctx.record(2)
    for _ in range(reps):
        M.SpMV(1.0, u, 0.0, y)
    ctx.record(3)
    ctx.sync()
    ms_spmv = ctx.elapsed_ms(2, 3) / reps
    b = -aff
    bj = B.BlockJacobi(M, omega=1.0, NoOfIterations=1)
    gm = B.SoftGMRES(M=M, Precond=bj, MaxKrylovDim=30, ConvergenceCriterion=1e-10)
    t_run = 1e30
    for _ in range(2):
        xs = np.zeros(N)
        t0 = time.perf_counter()
        xs = gm.Solve(xs, b)
        t_run = min(t_run, time.perf_counter() - t0)
    st = gm.stats()
    r = b - M.SpMV(1.0, xs, 0.0, np.zeros(N))
    out["sip_2d_64x64_k2"] = {
        "dof": N, "evaluate_ms": ms, "evaluate_dof_per_s": N / (ms * 1e-3), "assembly_s": t_asm,
        "bsr_blocks": M.nnzb, "spmv_ms": ms_spmv,
        "soft_gmres30_block_jacobi": {"solver_run_s_host_vectors": t_run, "iterations": st["iterations"],
                                      "converged": st["converged"],
                                      "termination": "left-preconditioned estimate < R0 tol + tol (the reference's)",
                                      "true_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(b))}}


def xdg_leg(B, ctx, out, n=48, seed=3):
    """Two-phase XDG Poisson, sphere in a 48^3 grid (k = 2, diffusion 1 inside /
    1000 outside, agglomeration threshold 0.1): set-up (cut-cell rules,
    device assembly, agglomeration), variable-block SpMV and the solve to 1e-8
    with FGMRES + FP32 two-level preconditioner."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.4, 0.6, 3)
    r = float(rng.uniform(0.2, 0.3))
    x = np.linspace(0.0, 1.0, n + 1)
    g = B.Grid(ctx, x, x, x)
    t0 = time.perf_counter()
    X = B.XdgPoisson(g, 2, c, r, 1.0, 1000.0, agg_threshold=0.1)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    X.set_rhs([1.0, 1.0], None)
    res = {"grid": f"{n}^3", "center": c.tolist(), "radius": r, "cut_cells": X.ncut, "agglomeration_pairs": X.npairs,
           "dof_full": X.N_full, "dof_agglomerated": X.N, "cut_rule_nodes": X.nnodes, "setup_s": t_setup}
    # variable-block SpMV on the agglomerated matrix (10 / 20 DOF blocks)
    for name, prec, dt in (("spmv_fp64", B.BH_FP64, np.float64), ("spmv_fp32", B.BH_FP32, np.float32)):
        xv = B.DeviceArray.from_host(ctx, rng.standard_normal(X.N).astype(dt))
        yv = B.DeviceArray(ctx, X.N, dt)
        for _ in range(3):
            X.SpMV(1.0, xv, 0.0, yv, which=1, prec=prec)
        ctx.sync()
        reps = 50
        ctx.record(2)
        for _ in range(reps):
            X.SpMV(1.0, xv, 0.0, yv, which=1, prec=prec)
        ctx.record(3)
        ctx.sync()
        ms = ctx.elapsed_ms(2, 3) / reps
        es = np.dtype(dt).itemsize
        nbytes = X.nvals * es + 2 * X.N * es
        res[name] = {"ms": ms, "algorithmic_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                     "hbm_frac": nbytes / (ms * 1e-3) / PEAK_HBM}
        del xv, yv
    # matrix-free FP64 evaluate of the full (cell, species) operator over the element list
    # (volume, face and interface items in colour groups, tensor and cut rules)
    uf = B.DeviceArray.from_host(ctx, rng.standard_normal(X.N_full))
    yf = B.DeviceArray(ctx, X.N_full)
    for _ in range(3):
        X.Evaluate(1.0, 0.0, yf, uf)
    ctx.sync()
    reps = 20
    ctx.record(2)
    for _ in range(reps):
        X.Evaluate(1.0, 0.0, yf, uf)
    ctx.record(3)
    ctx.sync()
    ms = ctx.elapsed_ms(2, 3) / reps
    nel = X.nelem
    ints = X.elements()["ints"]
    cut_items = ints[:, 8] > 0
    two_sided = (ints[:, 0] == 1) | (ints[:, 0] == 3)
    emat_values = int((np.where(two_sided, 4, 1)[cut_items]).sum()) * X.Np * X.Np
    # u + y, the element records (64 B each) and the cut items' element matrices
    nbytes = 2 * X.N_full * 8 + nel * 64 + emat_values * 8
    res["evaluate_fp64"] = {"ms": ms, "dof_per_s": X.N_full / (ms * 1e-3), "elements": nel,
                            "cut_elements": int(cut_items.sum()), "colour_groups": 14,
                            "algorithmic_bytes": nbytes, "GBps": nbytes / (ms * 1e-3) / 1e9,
                            "hbm_frac": nbytes / (ms * 1e-3) / PEAK_HBM,
                            "note": "uncut items in closed modal form (one lane per item), cut items from the "
                                    "element matrices kept by the assembly (one lane per item row)"}
    del uf, yf
    b = X.rhs(1)
    # one configuration for both precisions, the fastest of a sweep over pre/post sweeps
    # (1, 2), K-cycle steps (1, 2, 4, 6, 8), coarsest size (64 to 512 cells) and Schwarz
    # on the coarse aggregates (on, off); in FP32 the cut-cell blocks are applied with
    # their FP64 inverse (ill-conditioned at 1:1000)
    for name, prec in (("solve_fp32_precond", B.BH_FP32), ("solve_fp64_precond", B.BH_FP64)):
        t0 = time.perf_counter()
        P = B.XdgPrecond(X, coarse_degree=1, min_coarse=512, pre=1, post=1, kcycle_inner=4, prec=prec)
        B.pmg_set_graphs(P, USE_GRAPHS)
        if prec == B.BH_FP32:
            B.xdg_precond_set_fp64_blocks(P, True)
        ctx.sync()
        t_init = time.perf_counter() - t0
        S = B.XdgFlexGMRES(X, P, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=1000)
        t_run = 1e30
        for rep in range(2):   # the first solve also captures the graphs
            u = np.zeros(X.N)
            t0 = time.perf_counter()
            S.Solve(u, b)
            t_run = min(t_run, time.perf_counter() - t0)
        st = S.stats()
        yv = X.SpMV(1.0, u, 0.0, np.zeros(X.N), which=1)
        res[name] = {"precond_init_s": t_init, "solver_run_s": t_run, "iterations": st["iterations"],
                     "converged": st["converged"],
                     "preconditioner": "two-level: cell block-Jacobi (1 pre / 1 post sweep), p = 1 and aggregation "
                                       "levels down to 512 cells with a 4-step K-cycle, dense coarsest solve",
                     "true_rel_residual": float(np.linalg.norm(b - yv) / np.linalg.norm(b))}
        del S, P
    out["xdg"] = res


def euler_initial_state(n, k, seed, gamma=1.4, mach=0.2, nmodes=8, kmax=3):
    """Conserved-variable DG coefficients [cell][rho, m_x, m_y, m_z, E][mode] on the
    periodic n^3 grid of [0,1]^3: density 1 + 0.1 * (8 seeded Fourier modes,
    normalised), a seeded divergence-free Fourier velocity field scaled to
    max |u| = mach (sound speed 1), pressure 1/gamma; L2 projection with 4 Gauss
    points per direction, one z-layer of cells at a time."""
    from math import comb
    rng = np.random.default_rng(seed)
    kv = rng.integers(-kmax, kmax + 1, (nmodes, 3))
    kv[np.all(kv == 0, axis=1)] = [1, 0, 0]
    A = rng.standard_normal(nmodes)
    ph = rng.uniform(0, 2 * np.pi, nmodes)
    av = rng.standard_normal((nmodes, 3))
    dirs = np.cross(av, kv)                         # (a x k) . k = 0: divergence-free
    psi = rng.uniform(0, 2 * np.pi, nmodes)
    xg, wg = np.polynomial.legendre.leggauss(4)
    L = legendre_1d(k, xg)
    ex = np.array([(a_, b_, d - a_ - b_) for d in range(k + 1) for a_ in range(d + 1) for b_ in range(d - a_ + 1)])
    Np = comb(k + 3, 3)
    # reference basis at the 64 tensor nodes, node order (qx fastest)
    qi = np.stack(np.meshgrid(np.arange(4), np.arange(4), np.arange(4), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    Psi = L[qi[:, 0]][:, ex[:, 0]] * L[qi[:, 1]][:, ex[:, 1]] * L[qi[:, 2]][:, ex[:, 2]]   # (64, Np)
    W = wg[qi[:, 0]] * wg[qi[:, 1]] * wg[qi[:, 2]]
    h = 1.0 / n
    sc = np.sqrt((h / 2) ** 3)
    PW = Psi * W[:, None] * sc                     # (64, Np)
    c1 = np.arange(n) * h
    # scale factors from a coarse sampling of the fields
    xs = np.random.default_rng(seed + 1).random((20000, 3))
    srho = np.abs(np.sin(2 * np.pi * xs @ kv.T + ph) @ A).max()
    sv = np.sqrt(((np.sin(2 * np.pi * xs @ kv.T + psi) @ dirs) ** 2).sum(1)).max()
    U = np.zeros((n, n, n, 5, Np))
    # the modes are separable: sin(2 pi k.x + phi) = Im(e^{i phi} e^{2 pi i kx x} e^{2 pi i ky y} e^{2 pi i kz z}),
    # so per direction a table e[m, cell, node] and per layer a broadcast product
    xn = c1[:, None] + 0.5 * (xg[None, :] + 1) * h                      # [cell, node] coordinate per direction
    ed = [np.exp(2j * np.pi * kv[:, d, None, None] * xn[None]) for d in range(3)]   # [m, cell, node]
    for kz in range(n):
        # T[m, j, i, qz, qy, qx] -> [m, j, i, q] with q = qz 16 + qy 4 + qx
        T = (ed[2][:, kz, None, None, :, None, None] * ed[1][:, :, None, None, :, None] * ed[0][:, None, :, None, None, :])
        T = T.reshape(len(A), n, n, 64)
        rho = 1.0 + 0.1 * np.einsum("m,mjiq->jiq", A, (np.exp(1j * ph)[:, None, None, None] * T).imag) / srho
        u = np.einsum("mc,mjiq->jiqc", dirs, (np.exp(1j * psi)[:, None, None, None] * T).imag) * (mach / sv)
        p = 1.0 / gamma
        E = p / (gamma - 1) + 0.5 * rho * (u ** 2).sum(-1)
        F = np.stack([rho, rho * u[..., 0], rho * u[..., 1], rho * u[..., 2], E], -2)   # [j, i, 5, q]
        U[kz] = F @ PW
    return U.ravel()


def euler_rhs_flops(n, k, D=3, nq1=4):
    """Algorithmic FLOPs of one Euler/HLLC right-hand side on the periodic n^D grid:
    per cell, states at the volume nodes (2 NV Np NQV), flux to modes
    (2 NV D NQV Np) and the physical flux (~30 per direction and node); per face
    (D per cell), traces and flux to modes on both sides (8 NV Np NQF) and the
    HLLC flux (~120 per node)."""
    from math import comb
    Np, NV = comb(k + D, D), D + 2
    NQV, NQF = nq1 ** D, nq1 ** (D - 1)
    cell = 2 * NV * Np * NQV + 2 * NV * D * NQV * Np + 30 * D * NQV
    face = 8 * NV * Np * NQF + 120 * NQF
    return n ** D * (cell + D * face)


def euler_leg(B, ctx, out, n=96, k=2, steps=20, seed=4):
    """Explicit DG Euler / HLLC, periodic 96^3, k = 2 (44.2M DOF): 20 TVD3 steps in
    FP64 and FP32 from the same state; throughput and FP32-vs-FP64 difference."""
    x = np.linspace(0.0, 1.0, n + 1)
    g = B.Grid(ctx, x, x, x)
    e = B.EulerOperator(g, k=k, nq1=4, gamma=1.4)
    t0 = time.perf_counter()
    U0 = euler_initial_state(n, k, seed)
    t_init = time.perf_counter() - t0
    dt = 0.3 * (1.0 / n) / ((2 * k + 1) * 1.2)
    res = {"grid": f"{n}^3 periodic", "k": k, "dof": int(e.N), "steps": steps, "dt": dt, "initial_projection_s": t_init}
    finals = {}
    for name, prec, dtp in (("fp64", B.BH_FP64, np.float64), ("fp32", B.BH_FP32, np.float32)):
        dU = B.DeviceArray.from_host(ctx, U0.astype(dtp))
        e.tvd3(dU, dt, 1, prec=prec)      # warm-up step (then restart from U0)
        dU.upload(U0.astype(dtp))
        ctx.sync()
        ctx.record(4)
        e.tvd3(dU, dt, steps, prec=prec)
        ctx.record(5)
        ctx.sync()
        ms = ctx.elapsed_ms(4, 5)
        finals[name] = dU.download().astype(np.float64)
        res[name] = {"ms_total": ms, "ms_per_step": ms / steps, "dof_steps_per_s": e.N * steps / (ms * 1e-3),
                     "finite": bool(np.isfinite(finals[name]).all())}
        # the right-hand-side kernel alone (one launch per evaluation), against the matrix-core peak
        dR = B.DeviceArray(ctx, e.N, dtp)
        reps = 10
        ctx.record(4)
        for _ in range(reps):
            e.rhs(dU, dR, prec=prec)
        ctx.record(5)
        ctx.sync()
        rms = ctx.elapsed_ms(4, 5) / reps
        fl = euler_rhs_flops(n, k)
        peak = 78.6e12 if prec == B.BH_FP64 else 157.3e12
        res[name]["rhs_kernel_ms"] = rms
        res[name]["rhs_roofline"] = {"bound": "mfma", "algorithmic_flop": fl, "achieved_TFLOPs": fl / (rms * 1e-3) / 1e12,
                                     "peak_TFLOPs": peak / 1e12, "frac": fl / (rms * 1e-3) / peak,
                                     "hbm_GBps_at_algorithmic_bytes": 2 * e.N * np.dtype(dtp).itemsize / (rms * 1e-3) / 1e9}
        del dU, dR
    res["fp32_vs_fp64_rel_l2"] = float(np.linalg.norm(finals["fp32"] - finals["fp64"]) / np.linalg.norm(finals["fp64"]))
    # conservation of the cell-mean totals over the run (FP64)
    t0_ = U0.reshape(-1, 5, e.Np)[:, :, 0].sum(0)
    t1_ = finals["fp64"].reshape(-1, 5, e.Np)[:, :, 0].sum(0)
    res["fp64_conservation_rel"] = float(np.abs(t1_ - t0_).max() / np.abs(t0_).max())
    out["euler"] = res


def cpu_baseline(k, target_s=12.0, threads=None):
    """Oracle (CPU restatement of the reference's quadrature evaluate) on a
    bounded sample: 3D 24^3 cells, same k and coefficient law; all host threads
    the oracle's OpenMP runtime offers, or `threads`."""
    import oracle as O
    if threads is not None:
        O.lib().orc_set_threads(int(threads))
    n = 24
    x = np.linspace(0.0, 1.0, n + 1)
    g = O.Grid(x, x, x)
    nu = 10.0 ** np.random.default_rng(2024).uniform(-2, 2, g.J)
    op = O.Sip(g, k, nu)
    u = np.random.default_rng(1).standard_normal(op.N)
    op.evaluate(u)
    reps, t0 = 0, time.perf_counter()
    while True:
        op.evaluate(u)
        reps += 1
        dt = time.perf_counter() - t0
        if dt > target_s or reps >= 2000:
            break
    return dict(value=op.N * reps / dt, unit="DOF/s", cores=int(O.lib().orc_num_threads()), kind="port",
                sample=f"oracle FP64 SIP evaluate, 3D {n}^3 cells k={k} ({op.N} DOF), {reps} applies in {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--k", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-xdg", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-euler", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",  # kept for old command lines: the probe always runs
                    help="no effect: the timed 1 GiB device copy (copy16_k) always runs")
    ap.add_argument("--no-graphs", action="store_true",
                    help="eager kernel launches in the multigrid (for kernel-trace profilers that cannot follow graphs)")
    args = ap.parse_args()
    global USE_GRAPHS
    USE_GRAPHS = not args.no_graphs
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (single-GPU kernels; the multi-GPU "
              "sharding is a design, not a measured path). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import bosss_amd as B

    n, k, D = args.n, args.k, 3
    ctx = B.Context(0)
    x = np.linspace(0.0, 1.0, n + 1)
    grid = B.Grid(ctx, x, x, x)
    nu = 10.0 ** np.random.default_rng(2024).uniform(-2, 2, grid.J)
    op = B.SipOperator(grid, k, nu)
    N = op.N
    u = B.DeviceArray.from_host(ctx, np.random.default_rng(1).standard_normal(N))
    y = B.DeviceArray(ctx, N)

    def step():
        op.Evaluate(1.0, 0.0, y, u, affine=False)

    for _ in range(args.warmup):
        step()
    ctx.sync()
    t0 = time.perf_counter()
    ctx.record(0)
    for _ in range(args.steps):
        step()
    ctx.record(1)
    ctx.sync()
    wall = time.perf_counter() - t0
    kern_ms = ctx.elapsed_ms(0, 1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = N * args.steps / wall

    hbm_bytes = 2 * N * 8 + grid.J * 8
    # the FP64 apply picks its kernel (vector FMA, node by node or modal, or FP64 MFMA) by timing on first use
    f64_path, f64_w, f64_x, f64_ms_v, f64_ms_m = op.fp64_tuning()
    modal = f64_path in ("modal", "tiled")
    flops = sip_flops_modal(n, D, k) if modal else sip_flops(n, D, k)
    achieved = flops / (kern_ms * 1e-3)
    if f64_path == "mfma":
        kname = f"sip_apply_mfma64_kernel<{D}, {k}, {f64_w}>"
    elif f64_path == "tiled":
        kname = f"sip_apply_tile_kernel<{D}, {k}, double, {f64_w}>"
    else:
        kname = f"sip_apply_kernel<{D}, {k}, double, {'true' if modal else 'false'}>"
    if modal:
        # the modal form needs 1/8 of the quadrature's FLOPs: HBM is the bound
        gbps = hbm_bytes / (kern_ms * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": gbps, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                    "frac": gbps * 1e9 / PEAK_HBM, "traffic": measured_traffic(n, k, kname),
                    "kernel": kname, "kernel_ms": kern_ms, "algorithmic_bytes": hbm_bytes,
                    "algorithmic_flop": flops, "fp64_frac_at_algorithmic_flop": achieved / PEAK_FP64}
    else:
        # compute-bound: FP64 peak, the same for vector FMA and matrix cores on MI355X
        roofline = {"bound": "mfma", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                    "frac": achieved / PEAK_FP64, "traffic": measured_traffic(n, k, kname),
                    "kernel": kname, "kernel_ms": kern_ms,
                    "algorithmic_flop": flops, "algorithmic_bytes": hbm_bytes,
                    "hbm_frac_at_algorithmic_bytes": hbm_bytes / (kern_ms * 1e-3) / PEAK_HBM}

    out = {
        "metric": "DG op-apply DOF/s; time-to-1e-8 residual, mixed+learned vs plain FP64 MG",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded N(0,1) DOF vector, log-uniform per-cell diffusion)",
        "config": {"workload": f"3D SIP Poisson {n}^3 hexes k={k} ({N} DOF), FP64 matrix-free operator apply",
                   "grid": f"{n}^3", "k": k, "dof": N},
        "roofline": roofline,
        "fp64_autotune": {"chosen": f64_path, "vector_ms": f64_ms_v, "mfma64_ms": f64_ms_m,
                          "mfma64_waves_per_workgroup": f64_w, "mfma64_workgroups_per_xcd": f64_x},
    }
    # known-size copy kernel (copy16_k), timed: the HBM rate a plain streaming kernel
    # reaches on this GPU (read + write bytes per second), beside the 8 TB/s peak; its
    # FETCH_SIZE / WRITE_SIZE also calibrate the rocprofv3 traffic counters
    nbytes = 1024 * 1024 * 1024
    src = B.DeviceArray.from_host(ctx, np.ones(nbytes // 4, np.float32))
    dst = B.DeviceArray(ctx, nbytes // 4, np.float32)
    dst.copy_from(src)
    ctx.sync()
    reps = 10
    ctx.record(2)
    for _ in range(reps):
        dst.copy_from(src)
    ctx.record(3)
    ctx.sync()
    cms = ctx.elapsed_ms(2, 3) / reps
    out["copy_probe"] = {"kernel": "copy16_k", "bytes_read": nbytes, "bytes_written": nbytes, "launches": reps + 1,
                         "ms": cms, "GBps": 2 * nbytes / (cms * 1e-3) / 1e9,
                         "frac_of_peak": 2 * nbytes / (cms * 1e-3) / PEAK_HBM}
    out["roofline"]["frac_of_measured_copy"] = out["roofline"].get("achieved", 0) * 1e9 / (2 * nbytes / (cms * 1e-3)) \
        if out["roofline"]["unit"] == "GB/s" else None
    del src, dst
    # reduced-precision applies of the same operator (preconditioner use): FP32
    # vectors; vector FP32 FMA vs FP16 MFMA contractions with FP32 accumulation
    u64 = u.download()
    ref = y.download()
    u32 = B.DeviceArray.from_host(ctx, u64.astype(np.float32))
    y32 = B.DeviceArray(ctx, N, np.float32)
    bytes32 = 2 * N * 4 + grid.J * 4
    for name, prec, mfma_peak in (("apply_fp32", B.BH_FP32, 157.3e12), ("apply_fp16_mfma", B.BH_FP16_MFMA, 2.5e15)):
        for _ in range(3):
            op.Evaluate(1.0, 0.0, y32, u32, affine=False, prec=prec)
        ctx.sync()
        ctx.record(0)
        for _ in range(args.steps):
            op.Evaluate(1.0, 0.0, y32, u32, affine=False, prec=prec)
        ctx.record(1)
        ctx.sync()
        ms = ctx.elapsed_ms(0, 1) / args.steps
        err = float(np.linalg.norm(y32.download().astype(np.float64) - ref) / np.linalg.norm(ref))
        if prec == B.BH_FP16_MFMA:
            w16 = op.mfma_tuning()[0]      # waves, + 100 for the neighbour-prefetch variant
            kname32 = f"sip_apply_mfma_kernel<{D}, {k}, {w16 % 100}, {'true' if w16 >= 100 else 'false'}>"
            modal32 = False
        else:
            p32, w32, _ = op.fp32_tuning()
            modal32 = p32 in ("modal", "tiled")
            kname32 = (f"sip_apply_tile_kernel<{D}, {k}, float, {w32}>" if p32 == "tiled"
                       else f"sip_apply_kernel<{D}, {k}, float, {'true' if modal32 else 'false'}>")
        # the FP16-MFMA kernel evaluates the face rule node by node; the FP32 vector
        # kernel the form its own tuner picked
        fl = sip_flops_modal(n, D, k) if modal32 else sip_flops(n, D, k)
        out[name] = {"kernel_ms": ms, "dof_per_s": N / (ms * 1e-3), "algorithmic_bytes": bytes32,
                     "kernel": kname32, "traffic": measured_traffic(n, k, kname32),
                     "hbm_GBps": bytes32 / (ms * 1e-3) / 1e9, "hbm_frac": bytes32 / (ms * 1e-3) / PEAK_HBM,
                     "algorithmic_flop": fl, "TFLOPs": fl / (ms * 1e-3) / 1e12,
                     "compute_frac": fl / (ms * 1e-3) / mfma_peak, "rel_l2_vs_fp64": err}
        if prec == B.BH_FP16_MFMA:
            w_, x_, t_ = op.mfma_tuning()
            out[name]["autotuned_launch"] = {"waves_per_workgroup": w_ % 100, "prefetch_neighbours": w_ >= 100,
                                             "workgroups_per_xcd": x_, "tuning_ms": t_}
    del u32, y32
    if not args.no_solve:
        solve_leg(B, ctx, grid, op, n, k, out)
        sip2d_leg(B, ctx, out)
    if not args.no_learned:
        learned_leg(B, ctx, out)
    if not args.no_xdg:
        xdg_leg(B, ctx, out)
    if not args.no_euler:
        euler_leg(B, ctx, out)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(k)
        # the same sample on one host thread (the reference's evaluate is single-threaded per MPI rank)
        out["cpu_baseline"]["single_core"] = cpu_baseline(k, target_s=6.0, threads=1)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()

/* bosss_hip.h -- C ABI of libbosss_hip.so, the MI355X (gfx950) backend for the
 * BoSSS DG-operator / block-solver hot path.
 *
 * Conventions (mirroring the native bindings BoSSS already uses through
 * DynLibLoader, e.g. ilPSP.LinSolvers.PARDISO/Wrapper_MKL.cs:33-115):
 *   - every entry point returns an int status: 0 = OK, < 0 = error; the text of
 *     the last error of the calling thread is read with bh_last_error();
 *   - handles are opaque; the caller owns every array it passes in; the library
 *     copies what it keeps into device memory it owns and frees in *_destroy;
 *   - vector arguments are host pointers unless BH_DEVICE_PTR is set in the
 *     `flags` argument, in which case they are device pointers (HBM-resident,
 *     e.g. from bh_malloc) and no host<->device copy happens;
 *   - all work of one context is ordered on one HIP stream; calls return when
 *     the result is complete (host pointers) or enqueued (device pointers;
 *     bh_sync waits);
 *   - DOF vectors are cell-major: index = cell * Np + mode, cells in the
 *     numbering of the reference's Cartesian grid generators, modes in the
 *     order of the reference's orthonormal polynomial tables
 *     (F/UnsetteledCoordinateMapping.cs:696-711, F/Square.cs:192-260,
 *     F/Cube_polynomials.cs:13-60).
 * Reference paths are relative to the BoSSS source tree `src/`:
 *   F/ = L2-foundation/BoSSS.Foundation, P/ = L1-platform/BoSSS.Platform,
 *   AS/ = L3-solution/BoSSS.Solution.AdvancedSolvers, IL/ = ilPSP/layer_1.2-ilPSP/ilPSP.
 */
#ifndef BOSSS_HIP_H
#define BOSSS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bh_ctx bh_ctx;
typedef struct bh_grid bh_grid;
typedef struct bh_op bh_op;
typedef struct bh_bsr bh_bsr;
typedef struct bh_solver bh_solver;
typedef struct bh_mlp bh_mlp;

#define BH_DEVICE_PTR 1 /* vector arguments are device pointers */

/* precisions */
#define BH_FP64 0
#define BH_FP32 1
#define BH_FP16_MFMA 2 /* FP16 inputs, FP32 accumulation on MFMA */

/* ---------------------------------------------------------------- context
 * Replaces the library load + init of DynLibLoader (MW/DynLibLoader.cs:30). */
int bh_init(int device, bh_ctx** ctx);
int bh_finalize(bh_ctx* ctx);
int bh_last_error(char* buf, int len);
int bh_sync(bh_ctx* ctx);
int bh_malloc(bh_ctx* ctx, size_t bytes, void** dptr);
int bh_free(bh_ctx* ctx, void* dptr);
int bh_memcpy_h2d(bh_ctx* ctx, void* dst, const void* src, size_t bytes);
int bh_memcpy_d2h(bh_ctx* ctx, void* dst, const void* src, size_t bytes);
/* device-to-device copy on the context stream by a plain copy kernel (16-byte
 * aligned pointers); also the known-traffic kernel that calibrates HBM counters */
int bh_memcpy_d2d(bh_ctx* ctx, void* dst, const void* src, size_t bytes);
/* BLAS.ddot (IL/BLAS.cs:1216-1358) on device vectors: *result = x . y (FP64,
 * fixed-order reduction, returns after the result is on the host) */
int bh_dot(bh_ctx* ctx, int64_t n, const double* x, const double* y, double* result);
/* elapsed device time (ms) of the last `bh_timed_*` region on the context stream */
int bh_event_record(bh_ctx* ctx, int slot);
int bh_event_elapsed(bh_ctx* ctx, int slot0, int slot1, float* ms);

/* ---------------------------------------------------------------- grid
 * Cartesian grid from node arrays, as Grid2D.Cartesian2DGrid
 * (BoSSS.Foundation.Grid/Grid2D.cs:396-600, GlobalID = i*nY + j at :586) and
 * Grid3D.Cartesian3DGrid (Grid3D.cs:88-300, GlobalID = i + j*nX + k*nX*nY at
 * :261).  z may be NULL for D = 2.  ctx may be NULL: such a grid serves the
 * host-side set-up only (XDG cut-cell rules and agglomeration, bh_xdg_create). */
int bh_grid_cartesian(bh_ctx* ctx, int D, const double* x, int nx1, const double* y, int ny1,
                      const double* z, int nz1, bh_grid** grid);
int bh_grid_destroy(bh_grid* grid);
int bh_grid_info(bh_grid* grid, int* D, int* J, int* E);
/* edges (IN cell, OUT cell or -1, normal direction, boundary side or -1) and
 * cj (F/GridData_EdgeData.cs:2337-2604 CellIndices/EdgeTags;
 * F/GridData_CellData.cs:666-700 cj) -- exported for bit-exact index checks */
int bh_grid_edges(bh_grid* grid, int* in, int* out, int* dir, int* side);
int bh_grid_cj(bh_grid* grid, double* cj);

/* ---------------------------------------------------------------- SIP operator
 * DifferentialOperator with one SIPLaplace component (ipLaplace.cs:30-250;
 * ipPoisson/SipPoissonMain.cs:188-205) plus an RHS source (-rhs*v,
 * SipPoissonMain.cs:405-436).  nu: per-cell diffusion (J values);
 * bc[2*D]: 0 = Dirichlet, 1 = Neumann per boundary side (-x,+x,-y,+y,-z,+z);
 * penalty_base as the ipFlux ctor argument (SipControl.cs:76 default 1.3). */
int bh_sip_create(bh_grid* grid, int k, double penalty_base, const double* nu, const int* bc, bh_op** op);
int bh_op_destroy(bh_op* op);
int bh_op_info(bh_op* op, int* Np, int64_t* N);
/* boundary values g at the boundary-edge quadrature nodes (layout: boundary
 * edges in edge order x face nodes); bh_sip_bnd_nodes returns their physical
 * coordinates (count * D) so the host can evaluate g there (ipFlux.g_Diri /
 * g_Neum, SipPoissonMain.cs:370-385). */
int bh_sip_num_bnd_nodes(bh_op* op, int* count);
int bh_sip_bnd_nodes(bh_op* op, double* X);
int bh_sip_set_bnd_values(bh_op* op, const double* g);
/* RHS source field as DG coefficients (N values; NULL clears). */
int bh_sip_set_rhs(bh_op* op, const double* f, int flags);
/* IEvaluatorNonLin.Evaluate(alpha, beta, output) (F/IEvaluator.cs;
 * F/DifferentialOperator.cs:1601-1685):  y <- beta*y + alpha*Op(u), with the
 * affine (boundary-value and source) part included when `affine` != 0.
 * prec: BH_FP64 (drop-in), BH_FP32 or BH_FP16_MFMA (preconditioner use). */
int bh_op_evaluate(bh_op* op, double alpha, const double* u, double beta, double* y, int affine,
                   int prec, int flags);
/* No counterpart in the reference: the launch shape the autotuner picked for the
 * FP16-MFMA apply on its first call (cfg[0] waves per workgroup, plus 100 when
 * the variant that prefetches all neighbour blocks won; cfg[1] workgroups per
 * XCD; 0 before the first call) and its timed duration (ms) in ms[0].
 * FP64 apply (picked on its first call): cfg[2] = 0 not yet, 1 vector kernel
 * with the face rule evaluated node by node, 2 FP64-MFMA kernel, 3 vector kernel
 * with the face terms in modal form, 4 the same with the neighbours' traces
 * exchanged across lanes and waves (tiled); cfg[3] the chosen kernel's waves per
 * workgroup (MFMA or tiled), cfg[4] the MFMA kernel's workgroups per XCD; ms[1],
 * ms[2] the timed best vector and best MFMA durations.  The FP32 apply is tuned
 * separately, over the modal forms, on its own first call.  A decision holds for
 * every operator of the same device, dimension, degree, grid shape and precision
 * in the process (later operators of that shape take it without timing), so equal
 * inputs give bitwise equal outputs across operators.
 * cfg holds 5 ints, ms 3 floats. */
int bh_op_tuning(bh_op* op, int* cfg, float* ms);
/* No counterpart in the reference: the FP32 apply's choice (picked on its first
 * call): path 0 not yet, 1 face rule node by node, 3 modal, 4 tiled modal with
 * `waves` waves per workgroup; ms its timed duration.  Any pointer may be NULL. */
int bh_op_tuning_fp32(bh_op* op, int* path, int* waves, float* ms);
/* No counterpart in the reference: pin the FP64 apply to the vector kernel
 * (path 1: face rule node by node; path 3: modal face terms; path 4: tiled modal
 * face terms, waves in {2, 4, 8} per workgroup) or the FP64-MFMA kernel (path 2,
 * launch shape waves in {4, 8} per workgroup, wg_per_xcd >= 1) instead of timing
 * them on the first call; path 0 re-arms the autotuner (and forgets the decision
 * kept for this operator's shape).  All kernels give the
 * same result up to summation order. */
int bh_op_set_fp64_path(bh_op* op, int path, int waves, int wg_per_xcd);
/* No counterpart in the reference: pin the FP32 apply (multigrid levels) to path 1
 * (face rule node by node), 3 (modal face terms) or 4 (tiled modal face terms, waves
 * in {2, 4, 8} per workgroup); path 0 re-arms its own autotuner. */
int bh_op_set_fp32_path(bh_op* op, int path, int waves);
/* IEvaluatorLinear.ComputeMatrix (F/DifferentialOperator.cs:1787-1860):
 * assemble the operator matrix on the device and return the affine vector
 * (N values, host or device per flags; may be NULL). */
int bh_op_assemble(bh_op* op, bh_bsr** mat, double* affine, int flags);

/* ---------------------------------------------------------------- BlockMsrMatrix
 * Block-sparse matrix over cell blocks (P/BlockMsrMatrix.cs:191).  rowptr has
 * nrows+1 entries, col[nnzb] block-column indices (ascending per row), vals
 * nnzb blocks of bs*bs doubles, each block row-major. */
int bh_bsr_create(bh_ctx* ctx, int nrows, int bs, const int64_t* rowptr, const int* col,
                  const double* vals, bh_bsr** mat);
/* BlockMsrMatrix with variable block sizes (P/BlockMsrMatrix.cs:191-340, block row r
 * of length blk_len[r] at row/column offset blk_i0[r] of the vector layout, as the
 * XDG constant frame X/XDGBasis.cs:46-53 places cell j at j * Np * (species count)
 * with the present species' blocks first and unused slots after them).  rowptr and
 * col as in bh_bsr_create; vals lists every stored block (blk_len[r] x blk_len[c],
 * row-major) in rowptr order; blocks of length 0 are skipped.  Internally the rows
 * are stored back to back ("compressed": row r at the prefix sum of the lengths);
 * SpMV and solvers take and return vectors in the caller's layout, whose unused
 * slots read as zero rows and are written as zero.  Lengths 0..64, rows must not
 * overlap.  bh_bsr_info reports bs = 0 for such a matrix. */
int bh_bsr_create_v(bh_ctx* ctx, int nrows, const int64_t* rowptr, const int* col, const int64_t* blk_i0,
                    const int* blk_len, const double* vals, bh_bsr** mat);
int bh_bsr_destroy(bh_bsr* mat);
int bh_bsr_info(bh_bsr* mat, int* nrows, int* bs, int64_t* nnzb);
int bh_bsr_export(bh_bsr* mat, int64_t* rowptr, int* col, double* vals);
/* BlockMsrMatrix.SpMV(alpha, x, beta, y) (P/BlockMsrMatrix.cs:2970-3420):
 * y <- beta*y + alpha*A*x; prec BH_FP64 or BH_FP32 (FP32 copy of the values
 * and FP32 vectors, made on first use). */
int bh_bsr_spmv(bh_bsr* mat, double alpha, const void* x, double beta, void* y, int prec, int flags);

/* ---------------------------------------------------------------- solvers
 * ISolverSmootherTemplate (AS/ISmootherTemplate.cs): Init(op) at create time,
 * Solve(X, B) via bh_solver_solve, statistics via bh_solver_stats. */
/* BlockJacobi (AS/BlockJacobi.cs:35-260): omega, NoOfIterations (sweeps =
 * NoOfIterations + 1, as the reference's `iter <= NoOfIterations`). */
int bh_bjacobi_create(bh_bsr* mat, double omega, int iterations, int prec, bh_solver** solver);
/* SoftGMRES (AS/SoftGMRES.cs:75-470), left preconditioned, restart m; tol as
 * ConvergenceCriterion of IterativeSolverConfig.DefaultTermination
 * (AS/IterativeSolverConfig.cs:48-58).  precond may be NULL.  The operator is
 * the matrix `mat`, or the matrix-free SIP operator `op` when mat is NULL; a
 * variable-block `mat` (bh_bsr_create_v) makes the solver work on its layout. */
int bh_gmres_create(bh_bsr* mat, bh_op* op, bh_solver* precond, int restart, double tol, int maxit,
                    bh_solver** solver);
/* FlexGMRES (AS/FlexGMRES.cs): right preconditioned, flexible (the
 * preconditioner may change per iteration, e.g. FP32 / learned), convergence
 * judged on the true FP64 residual ||b - A x|| / ||b|| <= tol. */
int bh_fgmres_create(bh_bsr* mat, bh_op* op, bh_solver* precond, int restart, double tol, int maxit,
                     bh_solver** solver);
/* p/h-multigrid preconditioner (one V-cycle per Solve), the MI355X form of
 * the reference's p-multigrid (AS/PmgConfig.cs:80-200, AS/LevelPmg.cs) on the
 * MultigridOperator hierarchy (AS/MultigridOperator.cs:473-840) with
 * aggregation h-levels (AS/AggregationGridBasis.cs, F/CoarseningAlgorithms.cs:555-611):
 *   degrees[0..ndeg-1]: DG degree per p-level, degrees[0] = op's degree, strictly
 *     decreasing; the level operators are the Galerkin sub-blocks (leading modes);
 *   if the last degree is 0, the cells are then aggregated 2^D at a time
 *     (Galerkin P^T A P) until at most min_coarse cells remain; the coarsest
 *     level is solved with a dense inverse;
 *   smoothing: block-Jacobi (AS/BlockJacobi.cs), `pre`/`post` sweeps, omega;
 *   kcycle_inner > 0: the last p-level is solved by that many flexible GMRES
 *     steps preconditioned by its own V-cycle (K-cycle, the Krylov-accelerated
 *     coarse solve of AS/OrthonormalizationMultigrid.cs / LevelPmg's low-order
 *     solve, AS/LevelPmg.cs:411); 0: plain V-cycle;
 *   prec BH_FP64 or BH_FP32 (whole cycle in that precision); BH_FP16_MFMA: FP32
 *     cycle whose matrix-free finest level runs on the FP16 matrix cores;
 *   matrix_free != 0: the finest level's residuals use the matrix-free operator.
 * fine_mat: the assembled operator of `op` (NULL: assembled internally). */
int bh_pmg_create(bh_op* op, bh_bsr* fine_mat, const int* degrees, int ndeg, int min_coarse, double omega,
                  int pre, int post, int kcycle_inner, int prec, int matrix_free, bh_solver** solver);
/* No counterpart in the reference (launch strategy): enable (default) or disable
 * the HIP-graph replay of the host-synchronisation-free V-cycles of a
 * bh_pmg_create or bh_xdg_precond_create preconditioner (the whole cycle without
 * a K-cycle, else the K-cycle's inner cycle).  Results are identical either way. */
int bh_pmg_set_graphs(bh_solver* pmg, int enable);
/* OrthonormalizationMultigrid (AS/OrthonormalizationMultigrid.cs:797-1080) on the
 * finest level of the multigrid `pmg`: per iteration the pre-smoother's correction
 * (pre = 1; `pmg`'s level-0 smoother with its `pre` sweeps from a zero guess), the
 * coarse-grid correction `coarse_reps` times (restriction, `pmg` from level 1,
 * prolongation) and `post_sweeps` post-smoother corrections (`post` sweeps each);
 * each correction c is orthonormalised in the A-image against all earlier ones of
 * the solve (AddSol, :219-340: |A c| normalised, classical Gram-Schmidt in two
 * passes, re-orthonormalisation from the rescaled remainder on severe cancellation,
 * a random vector when nothing is left) and added with the residual-minimising
 * coefficient (MinimizeResidual, :149-215; the residual recomputed as b - A x after
 * each correction); a residual increase beyond 20 % fails with status -5 (the
 * reference's ArithmeticException).
 * A cancellation in a post-smoother correction skips the next pre-smoother
 * (:1047-1055).  Termination as IterativeSolverConfig.DefaultTermination
 * (tol, maxit).  The operator is `mat` (FP64 BSR), or the matrix-free `op` when
 * mat is NULL; `op` also supplies the learned correction's features.  max_hist
 * (1..128) bounds the history (FP64 vector pairs); when it is full the
 * minimisation restarts from the current iterate (the reference keeps every
 * correction of a solve). */
int bh_orthomg_create(bh_solver* pmg, bh_bsr* mat, bh_op* op, int pre, int post_sweeps, int coarse_reps,
                      int max_hist, double tol, int maxit, bh_solver** solver);
/* No counterpart in the reference: attach (NULL: detach) a network (bh_mlp_create
 * with Np + 4D inputs and Np outputs) whose correction of the current residual is
 * added after the post-smoother corrections of every iteration, orthonormalised
 * and residual-minimised like them. */
int bh_orthomg_set_learned(bh_solver* solver, bh_mlp* mlp);
/* info[4]: corrections added, severe cancellations, random directions, history
 * restarts (over the solver's life). */
int bh_orthomg_info(bh_solver* solver, int* info);
/* ---------------------------------------------------------------- learned correction
 * No counterpart in the reference: a per-cell MLP (in -> 128 -> 128 -> out,
 * ReLU; BF16 weights, FP32 accumulation on MFMA for inference; FP32 backward
 * and Adam on the device) used as an extra correction after the multigrid
 * cycle, accepted only when it reduces the residual (the acceptance rule of
 * OrthonormalizationMultigrid.MinimizeResidual, AS/OrthonormalizationMultigrid.cs:149-215).
 * Parameter layout (FP32, flat): W1[128][in_p], b1[128], W2[128][128], b2[128],
 * W3[out_p][128], b3[out_p], in_p/out_p = in/out rounded up to 16 (padding zero).
 * Feature / target arrays are row-major [cell][in_p] / [cell][out_p] FP32. */
int bh_mlp_create(bh_ctx* ctx, int n_in, int n_hidden, int n_out, uint64_t seed, bh_mlp** mlp);
int bh_mlp_destroy(bh_mlp* mlp);
int bh_mlp_info(bh_mlp* mlp, int* n_params, int* in_p, int* out_p);
int bh_mlp_get_params(bh_mlp* mlp, float* params);
int bh_mlp_set_params(bh_mlp* mlp, const float* params);
/* Y = network(X) for ncells rows (inference path: BF16 weights, MFMA) */
int bh_mlp_forward(bh_mlp* mlp, int ncells, const float* X, float* Y, int flags);
/* mean-squared-error loss over the first n_out outputs and its gradient
 * (FP32, weights rounded to BF16 as in inference); grad may be NULL */
int bh_mlp_gradient(bh_mlp* mlp, int ncells, const float* X, const float* T, float* grad, double* loss, int flags);
/* No counterpart in the reference.  Allocation-free variant of bh_mlp_gradient,
 * for callers that cannot allocate device memory (code under HIP graph capture,
 * tight training loops).  X [ncells][in_p], T [ncells][out_p], grad [n_params]
 * and loss (one double) are device pointers; grad and loss may be NULL.  The
 * gradient also lands in the network's own buffer for bh_mlp_adam_step.  The
 * per-block partials live in a fixed workspace of 25.5 MB per GPU (256 blocks x
 * 24,864 parameters, the largest network the widths allow) that the library
 * holds until the process ends; a network that does not fit is rejected.  No
 * allocation, no memset and no host synchronisation: results are stream-ordered
 * on the context's stream.  Every call on one GPU shares that workspace, so calls
 * that overlap on one GPU (several contexts or host threads) race. */
int bh_mlp_gradient_noalloc(bh_mlp* mlp, int ncells, const float* X, const float* T, float* grad, double* loss);
/* one Adam step with the gradient of the last bh_mlp_gradient(_noalloc) call */
int bh_mlp_adam_step(bh_mlp* mlp, float lr);
/* parameter gradient of a loss defined outside the network, from its upstream
 * gradient dY[ncells][out_p] = dLoss/dY (first n_out columns used; FP32 forward
 * with BF16-rounded weights, as bh_mlp_gradient); grad may be NULL, the gradient
 * also lands in the network's buffer for bh_mlp_adam_step */
int bh_mlp_backward(bh_mlp* mlp, int ncells, const float* X, const float* dY, float* grad, int flags);
/* feature map of a residual r (N values, FP64) on op's grid -> X[J][in_p], scales s[J]:
 * per cell its Np residual modes, the mean residual mode of its 2D face neighbours
 * (0 on the boundary), both over the patch norm s, and log10 of the neighbours'
 * diffusion ratios (2 on the boundary): Np + 4D inputs, in_p 16 or 32;
 * target map of an error e -> T[J][out_p] with the same scales */
int bh_mlp_features(bh_op* op, const double* r, float* X, float* s, int in_p, int flags);
int bh_mlp_targets(bh_op* op, const float* s, const double* e, float* T, int out_p, int flags);
/* correction of network outputs Y[J][out_p]: e = s_j h_j^2 / nu_j * Y (N values, FP64) */
int bh_mlp_correction(bh_op* op, const float* s, const float* Y, double* e, int out_p, int flags);
/* upstream gradient through that map: dY[j][m] = c * s_j h_j^2 / nu_j * g[j*Np + m]
 * (g = dLoss/de / c, N values FP64; columns m >= Np zero) */
int bh_mlp_upstream(bh_op* op, const float* s, const double* g, double c, float* dY, int out_p, int flags);
/* attach (mlp != NULL) or detach the learned correction of a bh_pmg_create solver:
 * after each cycle (correction z1 from zero, residual r' = b - A z1), e = network(r'),
 * and x = g1 z1 + g2 e with the coefficients minimising |b - A x|: A e orthonormalised
 * against A z1 and both added with their optimal coefficients, as
 * OrthonormalizationMultigrid.AddSol / MinimizeResidual do with their history
 * (AS/OrthonormalizationMultigrid.cs:149-300).  Counted as accepted when e removes
 * residual beyond the best multiple of z1, else rejected (x = z1, the plain cycle). */
int bh_pmg_set_learned(bh_solver* pmg, bh_mlp* mlp);
int bh_pmg_learned_stats(bh_solver* pmg, int* accepted, int* rejected);
/* No counterpart in the reference.  Fallback of the learned correction: after
 * `probe` preconditioner calls (0: never), a mean removed fraction of |r'|^2 per
 * step below `min_gain` switches the network off and the plain multigrid takes
 * over (defaults 8 and 1e-2).  bh_pmg_learned_info reports whether it is still
 * active and the mean gain: the fraction of the residual^2 of the best multiple of
 * the cycle's correction that the network's direction removes (over the probe
 * window once it is closed, else over the calls so far). */
int bh_pmg_set_learned_fallback(bh_solver* pmg, int probe, double min_gain);
/* No counterpart in the reference.  Cost rule of the fallback (on by default): when
 * the probe window closes with the mean gain above min_gain, 12 iterations of
 * flexible GMRES (FP64) on the last call's right-hand side are run twice on the
 * finest level, preconditioned by the cycle alone and by the cycle with the learned
 * step, and the network stays only if its run reduces |b - A x| at least as fast per
 * millisecond: what the network removes after a cycle may be residual the Krylov
 * iteration would remove anyway.  bh_pmg_learned_costs writes [ms per iteration
 * without, ms per iteration with the network, log-reduction per iteration without,
 * with] (zeros until the rule has run). */
int bh_pmg_set_learned_cost_rule(bh_solver* pmg, int on);
int bh_pmg_learned_costs(bh_solver* pmg, double* out4);
/* Schwarz (AS/Schwarz.cs:56-1454) with the multigrid-aggregate blocking and no
 * overlap: on every level that is aggregated (the k = 1 level and the h-levels
 * below it), the smoother becomes x (+)= omega sum_I R_I^T A_I^-1 R_I r over
 * the 2^D-cell aggregates I instead of block-Jacobi over cells (enable = 1), or
 * back to block-Jacobi (0).  Levels whose aggregate blocks exceed 64 DOF keep
 * block-Jacobi.  Applies to bh_pmg_create and to the coarse hierarchy of
 * bh_xdg_precond_create. */
int bh_pmg_set_schwarz(bh_solver* pmg, int enable);
/* LevelPmg with UseDiagonalPmg (AS/LevelPmg.cs:22-60 config, :232-330 set-up, :397-470
 * CoarseSolve / FineSolve, :561-600 Solve) as the finest level of a bh_pmg_create
 * hierarchy whose second level is a degree truncation: x = P x_c with x_c the
 * next level's solve of the low-order part of the residual (the modes of degree <=
 * that level's), then the high-order modes of every cell corrected with the inverse
 * of their part of the cell's diagonal block against the new residual.  Replaces
 * the finest level's pre- and post-smoothing (enable = 1), or restores them (0).
 * The low-order system is solved by the next levels (V- or K-cycle), not by a
 * sparse direct solver. */
int bh_pmg_set_level_pmg(bh_solver* pmg, int enable);
/* CellILU (AS/CellILU.cs:1229-1310 ComputeILU, :1490-1560 GetPattern, :182-250 Solve)
 * as the smoother of the finest level of a bh_pmg_create hierarchy: block ILU on the
 * pattern of A^(ilu_level + 1) (0, 1 or 2; the reference's default is 1), factorised
 * on the device level by level, applied as x (+)= omega U^-1 L^-1 r (FP32 factors in
 * an FP32 cycle); ilu_level < 0 restores block-Jacobi.  The sweeps are sequential
 * over the dependency levels of the pattern (one launch per level and sweep). */
int bh_pmg_set_cell_ilu(bh_solver* pmg, int ilu_level);
/* No counterpart in the reference (for checks): x = S(r) with the smoother S of level
 * `level` from a zero guess, in the hierarchy's precision (host FP64 vectors). */
int bh_pmg_smooth(bh_solver* pmg, int level, const double* r, double* x);
/* The multigrid mapping of a bh_pmg_create hierarchy (MultigridMapping /
 * AggregationGridBasis, AS/MultigridOperator.cs:573-640, AS/AggregationGridBasis.cs),
 * for checks.  level < 0: info[0] = number of levels.  Else info[7] = {cells, DOF per
 * cell, transfer to the next level (0 degree truncation, 1 aggregation, -1 coarsest),
 * degree, logical grid n0, n1, n2} and *nnzb the blocks of the level matrix.
 * bh_pmg_level_transfer: per fine cell its aggregate and its prolongation block
 * P[fine mode][coarse mode] (row-major); bh_pmg_level_matrix: the Galerkin level
 * matrix (row-major blocks, as bh_bsr_export). */
int bh_pmg_level_info(bh_solver* pmg, int level, int* info7, int64_t* nnzb);
int bh_pmg_level_transfer(bh_solver* pmg, int level, int* parent, double* P);
int bh_pmg_level_matrix(bh_solver* pmg, int level, int64_t* rowptr, int* col, double* vals);
/* Schwarz with overlap (Schwarz.Config.Overlap, AS/Schwarz.cs:491-512; block growth
 * :683-730; the additive sum of the block solutions :1026-1085): each aggregate block
 * is grown by `overlap` (0, 1 or 2) sweeps over the face neighbours of its cells and
 * sorted; a DOF in several blocks receives the sum of their local solutions, divided
 * by their number when overlap_scaling is set (EnableOverlapScaling, :518-524,
 * :791-815; the reference scales the whole iterate, which is the same for a smoother
 * started from zero).  Levels whose largest block exceeds 128 DOF keep block-Jacobi.
 * overlap = 0 is bh_pmg_set_schwarz. */
int bh_pmg_set_schwarz_overlap(bh_solver* pmg, int enable, int overlap, int overlap_scaling);
/* DiagBlockEquilib (AS/MultigridOperator_ChangeOfBasis.cs:819-828; the factor of
 * SymmInv, :624): every multigrid level but the coarsest is taken into the basis
 * in which its diagonal blocks are diagonals of signs (D_j = C_j S_j C_j^T, level
 * matrix L A L^T with L = C^-1, transfers C^T P L^T), so block-Jacobi there is
 * x~ (+)= omega S r~ and the K-cycle's inner products are those of that basis.
 * The preconditioner's input and output stay in the caller's basis.  For
 * hierarchies created with matrix_free = 0; irreversible; exclusive with the
 * Schwarz smoother and the learned correction. */
int bh_pmg_set_diag_equilib(bh_solver* pmg);
int bh_pmg_learned_info(bh_solver* pmg, int* active, double* mean_gain);

/* ---------------------------------------------------------------- compressible Euler
 * Explicit DG for the Euler equations on a periodic Cartesian grid: the CNS
 * operator with the HLLC fluxes OptimizedHLLC{Density,Momentum,Energy}Flux
 * (L3-solution/BoSSS.Solution.CompressibleFlowCommon/Convection/OptimizedHLLC*.cs,
 * MachScaling = gamma*mach^2) and RungeKutta with RungeKuttaScheme.TVD3
 * (L3-solution/BoSSS.Solution/TimeStepping/RungeKuttaScheme.cs:145-167).
 * State layout [cell][variable][mode], variables rho, m_1..m_D, E.
 * nq1 Gauss points per direction (CNS uses order 3k: nq1 = 4 for k = 2).
 * Supported (D, k, nq1): (3,1,2), (3,1,3), (3,2,4), (2,1,3), (2,2,4), (2,3,5). */
typedef struct bh_euler bh_euler;
int bh_euler_create(bh_grid* grid, int k, int nq1, double gamma, double mach, bh_euler** e);
int bh_euler_destroy(bh_euler* e);
int bh_euler_info(bh_euler* e, int64_t* N, int* Np);
/* R = dU/dt (prec BH_FP64: double vectors; BH_FP32: float device vectors) */
int bh_euler_rhs(bh_euler* e, const void* U, void* R, int prec, int flags);
/* `steps` TVD3 steps of size dt, in place */
int bh_euler_tvd3(bh_euler* e, void* U, double dt, int steps, int prec, int flags);
/* Newton (AS/Newton.cs:195-370, 880-960, 1546-1600) on the backward-Euler step
 * F(U) = U - Un - dt R(U) = 0 of this operator: matrix-free Jacobian by the
 * reference's directional difference (dirder), each step's J s = -F solved by
 * GMRES(restart) without preconditioner to 1e-5 |F| (the inexact-Newton
 * threshold of its matrix-free option), Armijo line search with parabolic
 * backtracking (at most 30 cuts); stops at |F| <= conv |F(U_start)| + conv or
 * after maxit steps.  U holds the initial guess (e.g. Un) and returns the solution;
 * info[3]: Newton steps, GMRES iterations, line-search cuts (may be NULL);
 * res[2]: |F| at the start and at the end (may be NULL).  FP64. */
int bh_euler_implicit_step(bh_euler* e, const double* Un, double* U, double dt, double conv, int maxit, int restart,
                           int* info, double* res, int flags);

/* ---------------------------------------------------------------- XDG two-phase Poisson
 * Cut-cell SIP Poisson with a spherical level set phi = |x - c| - r, species
 * A (phi < 0, diffusion nu_in) and B (phi > 0, nu_out): XLaplace_Bulk and
 * XLaplace_Interface in SIP mode (L3-solution/BoSSS.Solution.XNSECommon/
 * XLaplace.cs:50-346), assembled as XEvaluatorLinear.ComputeMatrix does
 * (X/XSpatialOperatorMk2_XEvaluators.cs:163) into (cell, species) blocks;
 * cut-cell rules by dimension reduction (X/Quadrature/SayeGaussRule_*.cs);
 * agglomeration of cut cells with volume fraction <= agg_threshold
 * (X/AgglomerationAlgorithm.cs:393-460, 2064-2140) applied as
 * CellAgglomerator.ManipulateMatrixAndRHS (X/CellAgglomerator.cs:1755).
 * X/ = L2-foundation/BoSSS.Foundation.XDG.  Dirichlet boundary everywhere.
 * DOF layout: cell-major, per cell the present species blocks (A then B),
 * Np modes each; the agglomerated ("reduced") system drops source blocks. */
typedef struct bh_xdg bh_xdg;
/* quad_order: Gauss points per 1D piece of the cut-cell rules (<= 0: k + 1).
 * Supported degrees: D = 2, k = 1..3; D = 3, k = 1..2 (other values are refused
 * before any set-up work).  Agglomeration: direct, chain and group passes
 * (X/AgglomerationAlgorithm.cs:1699-2330); a source with no source neighbour and
 * no target fails the call. */
int bh_xdg_create(bh_grid* grid, int k, const double* center, double radius, double nu_in, double nu_out,
                  double penalty_base, double agg_threshold, int quad_order, bh_xdg** x);
/* The same XDG problem for a level set given as a polynomial on the whole domain,
 * phi(x) = sum over i0, i1, i2 <= deg of coeffs[i0 + (deg + 1) (i1 + (deg + 1) i2)]
 * x^i0 y^i1 z^i2 (2D: (deg + 1)^2 coefficients), species A (nu_in) where phi < 0.
 * The cut-cell, cut-face and interface rules come from Saye's dimension reduction
 * (X/Quadrature/SayeGaussRule_Algorithm.cs, SayeGaussRule_Cube.cs): Bernstein bounds
 * decide which boxes are cut and in which directions phi is monotone, the roots are
 * found by a Newton iteration safeguarded by bisection, boxes without a monotone
 * direction are halved (up to 10 times, then a low-order rule is used and counted,
 * bh_xdg_cut_rule_fallbacks); deg <= 8.  On a grid with a device, deg <= 4 runs
 * the classification and the rules on the device (one thread per rule, the same rules
 * bit for bit as the host's).  The level set is one polynomial on the
 * whole grid, so both sides of every face see the same phi; BoSSS's per-cell DG
 * level-set field is not taken in this form. */
int bh_xdg_create_poly(bh_grid* g, int k, int deg, const double* coeffs, double nu_in, double nu_out,
                       double penalty_base, double agg_threshold, int quad_order, bh_xdg** x);
/* boxes (process-wide) where the polynomial cut rules used the low-order fallback */
int bh_xdg_cut_rule_fallbacks(long* count);
int bh_xdg_destroy(bh_xdg* x);
/* info[13]: N_full, N_red, cut cells, agglomeration pairs, failed sources, elements,
 * cut-rule nodes, full matrix values, reduced matrix values, boundary nodes, Np, J,
 * block rows of the agglomerated matrix (species units) */
int bh_xdg_info(bh_xdg* x, int64_t* info);
/* per cell: species mask (bit 0 A, bit 1 B); volumes [2][J]; interface area [J];
 * agglomerated length scales [2][J] (X/MultiphaseAgglomerator.cs:870-1030) */
int bh_xdg_cells(bh_xdg* x, int* spmask, double* vol, double* iface_area, double* len);
/* agglomeration pairs (source cell, target cell, species) in discovery order */
int bh_xdg_agglomeration(bh_xdg* x, int* pairs);
/* element list (ints[9]: type 0 volume / 1 interior face / 2 boundary face / 3 interface,
 * cellP, cellQ, speciesP, speciesQ, dir, side, rule start, rule length (0: tensor Gauss);
 * dbls[3]: nuP, nuQ, penalty) and the cut-rule nodes (xyz), weights and interface normals */
int bh_xdg_elements(bh_xdg* x, int* ints, double* dbls, double* nx, double* nw, double* nn);
/* which = 0: the full matrix over cells (blocks of the cell's present species);
 * which = 1: the agglomerated matrix over species units (uniform Np x Np blocks,
 * the DOF layout of the agglomerated vectors).  Exported as DOF offsets [rows+1],
 * block rows [rows+1], block columns, value offsets [nblocks+1] and column-major
 * block values (FP64). */
int bh_xdg_matrix(bh_xdg* x, int which, int64_t* dof_off, int64_t* rowptr, int* col, int64_t* voff, double* vals);
/* Dirichlet data: physical boundary quadrature nodes (3 per node) */
int bh_xdg_bnd_nodes(bh_xdg* x, double* xyz);
/* right-hand side from species-constant sources f[2] and Dirichlet values g at the boundary nodes (or NULL: 0) */
int bh_xdg_set_rhs(bh_xdg* x, const double* f, const double* g);
/* which = 0: full right-hand side (host); 1: agglomerated right-hand side */
int bh_xdg_rhs(bh_xdg* x, int which, double* b, int flags);
/* y <- alpha A x + beta y with the full (0) or agglomerated (1) matrix (BlockMsrMatrix.SpMV) */
int bh_xdg_spmv(bh_xdg* x, int which, double alpha, const void* xv, double beta, void* y, int prec, int flags);
/* XEvaluatorNonlin.Evaluate for the linear XLaplace forms (X/XSpatialOperatorMk2_XEvaluators.cs:467-646):
 * y <- beta y + alpha A u on the full (cell, species) DOF layout over the element list: uncut items
 * (whole cells or faces of one species) in closed modal form, equal to their tensor Gauss quadrature
 * up to rounding; cut items through the element matrices their own rules gave at set-up; one launch
 * per colour group, no atomics, bitwise reproducible */
int bh_xdg_evaluate(bh_xdg* x, double alpha, const double* u, double beta, double* y, int flags);
/* full-space DG coefficients from agglomerated ones (host vectors) */
int bh_xdg_extend(bh_xdg* x, const double* ured, double* ufull);
/* two-level preconditioner of the agglomerated system: block-Jacobi (FP32/FP64) on
 * the cell blocks and a species-merged degree-`coarse_degree` coarse space solved
 * by the aggregation multigrid (K-cycle with kcycle_inner inner steps).  Measured
 * fastest for a sphere in 48^3 at k = 2: coarse_degree 1, min_coarse 512, omega 0.8,
 * pre = post = 1, kcycle_inner 4, FP32 with FP64 cut-cell blocks */
int bh_xdg_precond_create(bh_xdg* x, int coarse_degree, int min_coarse, double omega, int pre, int post,
                          int kcycle_inner, int prec, bh_solver** solver);
/* The same two-level preconditioner for a caller's variable-block matrix of cell
 * blocks (bh_bsr_create_v, one block row per cell of g, each of length 0, Np or 2 Np:
 * the species blocks of the cell, species A first) in the caller's vector layout;
 * use it with bh_gmres_create on the same matrix (flexible = 1). */
int bh_xdg_precond_create_v(bh_grid* g, bh_bsr* mat, int Np, int coarse_degree, int min_coarse, double omega,
                            int pre, int post, int kcycle_inner, int prec, bh_solver** solver);
/* The DOF layouts of the full (which = 0) or agglomerated (1) XDG system: per cell,
 * its offset in the XDG constant frame (X/XDGBasis.cs:46-53: j * 2 Np), its length
 * (present, non-agglomerated species times Np) and its offset in this library's
 * compressed layout (bh_xdg_matrix's dof_off); any pointer may be NULL. */
int bh_xdg_dof_frame(bh_xdg* x, int which, int64_t* frame_i0, int* len, int64_t* compressed_i0);
/* The XDG mass matrix (X/MassMatrixFactory.cs:35), computed on the device: the block of
 * cell j and species s is M_mn = int over K_j and the species' domain of phi_m phi_n
 * (the identity for an uncut cell), over the full (cell, species) layout: row r of
 * that layout holds its block row in out[r * Np .. r * Np + Np) (N_full * Np values). */
int bh_xdg_mass_matrix(bh_xdg* x, double* out);
/* No counterpart in the reference: in an FP32 preconditioner, apply the cell
 * diagonal blocks with their FP64 inverse and FP64 accumulation (vectors stay
 * FP32); cut cells that couple two species at high contrast have blocks too
 * ill-conditioned for FP32-rounded inverses.  on = 1 / 0. */
int bh_xdg_precond_set_fp64_blocks(bh_solver* solver, int on);
/* No counterpart in the reference: the coarse matrix of an XDG preconditioner (cell
 * rows, blocks of Np(coarse degree), species merged: the sum over the species blocks
 * of a cell pair of their leading Npc x Npc parts), for checks.  Sizes first (arrays
 * NULL), then rowptr (nrows + 1), col (nnzb) and row-major vals (nnzb bs^2). */
int bh_xdg_precond_coarse_export(bh_solver* solver, int* nrows, int* bs, int64_t* nnzb, int64_t* rowptr, int* col,
                                 double* vals);
/* FGMRES on the agglomerated system (AS/FlexGMRES.cs), FP64 */
int bh_xdg_fgmres_create(bh_xdg* x, bh_solver* precond, int restart, double tol, int maxit, bh_solver** solver);

int bh_solver_solve(bh_solver* solver, double* x, const double* b, int flags);
int bh_solver_stats(bh_solver* solver, int* iterations, int* converged, double* res0, double* res);
int bh_solver_destroy(bh_solver* solver);

#ifdef __cplusplus
}
#endif
#endif /* BOSSS_HIP_H */

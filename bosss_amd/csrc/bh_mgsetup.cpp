// p/h multigrid preconditioner (host orchestration of the bh_mg.hip kernels).
//
// Hierarchy (AS/PmgConfig.cs:80-200 p-levels, AS/MultigridOperator.cs:573-640
// Galerkin setup, AS/AggregationGridBasis.cs + F/CoarseningAlgorithms.cs:555-611
// aggregation sequence):
//   level 0         : the fine DG operator, degree k (matrix-free or BSR)
//   p-levels        : degrees given by the caller (e.g. 3 -> 1); the operator
//                     is the Galerkin sub-block (leading modes of each cell)
//   h-levels        : cells of the last p-level are aggregated 2^D at a time
//                     (Cartesian index halving); the coarse space holds the
//                     polynomials of the same degree on the aggregate box;
//                     A_c = P^T A P, until at most `min_coarse` cells remain
//   coarsest        : dense inverse (applied as a GEMV on the device)
// Cycle: V(pre, post) with block-Jacobi smoothing x <- x + omega Dinv (b - A x)
// on every level (AS/BlockJacobi.cs:196-255), in FP64 or FP32.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>

#include "bh_mg.hpp"
#include "bh_mlp.hpp"
#include "bh_solvers.hpp"

namespace bh {

PMultigrid::~PMultigrid() {
    drop_graphs();
    for (auto& L : lv)
        if (L.ownA) delete L.A;
    delete eqA0;
    delete ownedAf;
}

void PMultigrid::drop_graphs() {
    for (int t = 0; t < 2; ++t)
        for (int l = 0; l < 16; ++l) {
            if (gexec[t][l]) (void)hipGraphExecDestroy(gexec[t][l]);
            gexec[t][l] = nullptr;
            gcalls[t][l] = 0;
        }
}

// The first call runs eagerly (it also builds lazily created tables), the second
// captures the cycle into a graph; later calls replay it.  The cycle's buffers
// are fixed per level, so the captured kernel arguments stay valid.
template <typename T>
void PMultigrid::cycle_body(int l) {
    vcycle<T>(l);
    if (l == 0 && learned && !learned_off) learned_correction<T>();
}

template <typename T>
void PMultigrid::vcycle_g(int l) {
    const int ti = sizeof(T) == 4 ? 1 : 0;
    if (!use_graphs || capturing || l >= 16 || gcalls[ti][l]++ == 0) {
        cycle_body<T>(l);
        return;
    }
    if (!gexec[ti][l]) {
        hipGraph_t g = nullptr;
        BH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        capturing = true;
        try {
            cycle_body<T>(l);
        } catch (...) {
            capturing = false;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        capturing = false;
        BH_HIP(hipStreamEndCapture(ctx->stream, &g));
        const hipError_t e = hipGraphInstantiate(&gexec[ti][l], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        BH_HIP(e);
    }
    BH_HIP(hipGraphLaunch(gexec[ti][l], ctx->stream));
}

static bh_bsr* make_bsr(bh_ctx* ctx, int nrows, int bs, const std::vector<int64_t>& rp, const std::vector<int>& col) {
    bh_bsr* M = new bh_bsr();
    M->ctx = ctx;
    M->nrows = nrows;
    M->bs = bs;
    M->h_rowptr = rp;
    M->h_col = col;
    M->nnzb = rp[nrows];
    M->rowptr.alloc((nrows + 1) * sizeof(int64_t));
    M->col.alloc(std::max<int64_t>(M->nnzb, 1) * sizeof(int));
    ctx->upload(M->rowptr.p, rp.data(), (nrows + 1) * sizeof(int64_t));
    ctx->upload(M->col.p, col.data(), M->nnzb * sizeof(int));
    M->vals.alloc(std::max<int64_t>(M->nnzb, 1) * bs * bs * sizeof(double));
    return M;
}

static int cell_of(int D, const int* n, int i, int j, int k) {
    return D == 2 ? i * n[1] + j : i + j * n[0] + k * n[0] * n[1];
}
static void ijk_of(int D, const int* n, int c, int* idx) {
    if (D == 2) { idx[0] = c / n[1]; idx[1] = c % n[1]; idx[2] = 0; }
    else { idx[0] = c % n[0]; idx[1] = (c / n[0]) % n[1]; idx[2] = c / (n[0] * n[1]); }
}

// L2-projection block of the coarse-box basis onto a fine cell:
// P[m][n] = int_f phi^f_m phi^c_n = sqrt(detJ_f/detJ_c) sum_q w_q psi_m(xi_q) psi_n(xi_c(x_q))
static void projection_block(int D, int deg, const double* lof, const double* hf, const double* loc, const double* hc,
                             double* P) {
    const int Np = basis_count(D, deg), n1 = deg + 1;
    std::vector<double> x(n1), w(n1), pf(Np), pc(Np);
    gauss_rule(n1, x.data(), w.data());
    double dJf = 1, dJc = 1;
    for (int d = 0; d < D; ++d) { dJf *= 0.5 * hf[d]; dJc *= 0.5 * hc[d]; }
    const double sc = std::sqrt(dJf / dJc);
    std::fill(P, P + Np * Np, 0.0);
    int nq = 1;
    for (int d = 0; d < D; ++d) nq *= n1;
    for (int q = 0; q < nq; ++q) {
        double xi[3], xc[3], wq = 1;
        int r = q;
        for (int d = 0; d < D; ++d) {
            const int id = r % n1;
            r /= n1;
            xi[d] = x[id];
            wq *= w[id];
            const double X = lof[d] + 0.5 * (xi[d] + 1.0) * hf[d];
            xc[d] = 2.0 * (X - loc[d]) / hc[d] - 1.0;
        }
        basis_eval(D, deg, xi, pf.data(), nullptr);
        basis_eval(D, deg, xc, pc.data(), nullptr);
        for (int m = 0; m < Np; ++m)
            for (int n = 0; n < Np; ++n) P[m * Np + n] += sc * wq * pf[m] * pc[n];
    }
}

void PMultigrid::build(bh_op* op, bh_bsr* Af, const std::vector<int>& degrees, int min_coarse, bool mf) {
    BH_CHECK(!degrees.empty() && degrees[0] == op->k, "multigrid: first degree must be the operator's degree");
    BH_CHECK(Af->bs == op->Np && Af->nrows == op->grid->J, "multigrid: matrix does not match the operator");
    fine_op = op;
    build_grid(op->ctx, op->grid, Af, degrees, min_coarse, mf ? op : nullptr);
}

void PMultigrid::build_grid(bh_ctx* c, bh_grid* grid, bh_bsr* Af, const std::vector<int>& degrees, int min_coarse, bh_op* mfop) {
    ctx = c;
    const int D = grid->D;
    dim = D;
    N = (int64_t)grid->J * basis_count(D, degrees[0]);
    matrix_free = mfop != nullptr;
    for (size_t i = 1; i < degrees.size(); ++i)
        BH_CHECK(degrees[i] < degrees[i - 1] && degrees[i] >= 0, "multigrid: degrees must decrease");
    BH_CHECK(Af->bs == basis_count(D, degrees[0]) && Af->nrows == grid->J, "multigrid: matrix does not match the grid");
    const int J = grid->J;
    // p-levels
    for (size_t i = 0; i < degrees.size(); ++i) {
        MgLevel L;
        L.J = J;
        L.deg = degrees[i];
        L.bs = basis_count(D, degrees[i]);
        L.N = (int64_t)J * L.bs;
        for (int d = 0; d < D; ++d) L.n[d] = grid->n[d];
        if (i == 0) {
            L.A = Af;
            L.ownA = false;
            L.op = mfop;
        } else {
            L.A = make_bsr(ctx, J, L.bs, Af->h_rowptr, Af->h_col);
            L.ownA = true;
            extract_subblocks(ctx, Af->nnzb, Af->bs, L.bs, Af->vals.as<double>(), L.A->vals.as<double>());
            if (mfop && degrees[i] >= 1) {
                // the truncated-basis Galerkin operator of a SIP matrix is the SIP operator of the
                // lower degree with the finest level's penalty factor (orthonormal hierarchical
                // basis; both quadratures are exact for it): apply it matrix-free as well
                L.ownop = make_sip_op(mfop->grid, degrees[i], mfop->penalty_base, mfop->penalty_deg,
                                      mfop->h_nu.data(), mfop->bc);
                L.op = L.ownop.get();
            }
        }
        L.kind = 0;
        lv.push_back(std::move(L));
    }
    // h-levels: aggregation of the last p-level
    {
        MgLevel& L0 = lv.back();
        L0.lo.resize((size_t)J * D);
        L0.h.resize((size_t)J * D);
        for (int c = 0; c < J; ++c) {
            int idx[3];
            grid->cell_ijk(c, idx);
            for (int d = 0; d < D; ++d) {
                L0.lo[(size_t)c * D + d] = grid->nodes[d][idx[d]];
                L0.h[(size_t)c * D + d] = grid->width[d][idx[d]];
            }
        }
    }
    while (lv.back().J > min_coarse) {
        MgLevel& F = lv.back();
        int nc[3] = {1, 1, 1};
        bool can = false;
        for (int d = 0; d < D; ++d) { nc[d] = (F.n[d] + 1) / 2; can |= F.n[d] > 1; }
        if (!can) break;
        const int bs = F.bs, deg = F.deg;
        const int Jc = nc[0] * nc[1] * (D == 3 ? nc[2] : 1);
        std::vector<int> parent(F.J);
        std::vector<double> loc((size_t)Jc * D, 1e300), hic((size_t)Jc * D, -1e300);
        for (int c = 0; c < F.J; ++c) {
            int idx[3];
            ijk_of(D, F.n, c, idx);
            const int I = cell_of(D, nc, idx[0] / 2, idx[1] / 2, idx[2] / 2);
            parent[c] = I;
            for (int d = 0; d < D; ++d) {
                loc[(size_t)I * D + d] = std::min(loc[(size_t)I * D + d], F.lo[(size_t)c * D + d]);
                hic[(size_t)I * D + d] = std::max(hic[(size_t)I * D + d], F.lo[(size_t)c * D + d] + F.h[(size_t)c * D + d]);
            }
        }
        std::vector<double> hc((size_t)Jc * D);
        for (size_t t = 0; t < hc.size(); ++t) hc[t] = hic[t] - loc[t];
        // prolongation blocks
        // (the block depends only on the fine box relative to the coarse box:
        // cache by that relative geometry, quantised far below round-off of the setup)
        std::vector<double> P;                  // the distinct blocks
        std::vector<int> pidx(F.J);             // block of each fine cell
        std::map<std::vector<long long>, int> seen;
        for (int c = 0; c < F.J; ++c) {
            const double* lf = &F.lo[(size_t)c * D];
            const double* hf = &F.h[(size_t)c * D];
            const double* lc = &loc[(size_t)parent[c] * D];
            const double* hcc = &hc[(size_t)parent[c] * D];
            std::vector<long long> key(2 * D);
            for (int d = 0; d < D; ++d) {
                key[2 * d] = std::llround((lf[d] - lc[d]) / hcc[d] * 1e13);
                key[2 * d + 1] = std::llround(hf[d] / hcc[d] * 1e13);
            }
            auto it = seen.find(key);
            if (it != seen.end()) {
                pidx[c] = it->second;
            } else {
                const int u = (int)(P.size() / ((size_t)bs * bs));
                P.resize(P.size() + (size_t)bs * bs);
                projection_block(D, deg, lf, hf, lc, hcc, &P[(size_t)u * bs * bs]);
                seen.emplace(std::move(key), u);
                pidx[c] = u;
            }
        }
        std::vector<int> cptr(Jc + 1, 0), cidx(F.J);
        for (int c = 0; c < F.J; ++c) cptr[parent[c] + 1]++;
        for (int I = 0; I < Jc; ++I) cptr[I + 1] += cptr[I];
        {
            std::vector<int> fillp(cptr.begin(), cptr.end() - 1);
            for (int c = 0; c < F.J; ++c) cidx[fillp[parent[c]]++] = c;
        }
        // Galerkin A_c = P^T A P (blocks: Ac[I][J] = sum P_i^T A_ij P_j over the children i of I,
        // j of J).  Host: the coarse sparsity and, per coarse block, the list of its fine
        // blocks (integer work); device: the block products, gathered in that fixed order.
        std::vector<int64_t> rp(Jc + 1, 0);
        std::vector<int> col;
        const std::vector<int64_t>& frp = F.A->h_rowptr;
        const std::vector<int>& fcl = F.A->h_col;
        std::vector<int> qof((size_t)F.A->nnzb);
        {
            std::vector<int> mark(Jc, -1), pos(Jc), rowc;
            for (int I = 0; I < Jc; ++I) {
                rowc.clear();
                for (int t = cptr[I]; t < cptr[I + 1]; ++t) {
                    const int i = cidx[t];
                    for (int64_t p = frp[i]; p < frp[i + 1]; ++p) {
                        const int Jb = parent[fcl[p]];
                        if (mark[Jb] != I) { mark[Jb] = I; rowc.push_back(Jb); }
                    }
                }
                std::sort(rowc.begin(), rowc.end());
                for (size_t t = 0; t < rowc.size(); ++t) pos[rowc[t]] = (int)t;
                for (int t = cptr[I]; t < cptr[I + 1]; ++t) {
                    const int i = cidx[t];
                    for (int64_t p = frp[i]; p < frp[i + 1]; ++p) qof[p] = (int)(rp[I] + pos[parent[fcl[p]]]);
                }
                col.insert(col.end(), rowc.begin(), rowc.end());
                rp[I + 1] = (int64_t)col.size();
            }
        }
        const int64_t nq = rp[Jc];
        std::vector<int64_t> qptr(nq + 1, 0);
        for (int64_t p = 0; p < F.A->nnzb; ++p) qptr[qof[p] + 1]++;
        for (int64_t q = 0; q < nq; ++q) qptr[q + 1] += qptr[q];
        std::vector<int> plist((size_t)F.A->nnzb), prow((size_t)F.A->nnzb);
        {
            std::vector<int64_t> fillq(qptr.begin(), qptr.end() - 1);
            for (int I = 0; I < Jc; ++I)                       // children in order, then their blocks
                for (int t = cptr[I]; t < cptr[I + 1]; ++t) {
                    const int i = cidx[t];
                    for (int64_t p = frp[i]; p < frp[i + 1]; ++p) {
                        const int64_t k = fillq[qof[p]]++;
                        plist[k] = (int)p;
                        prow[k] = i;
                    }
                }
        }
        F.kind = 1;
        F.cptr.alloc(cptr.size() * sizeof(int));
        F.cidx.alloc(cidx.size() * sizeof(int));
        F.parent.alloc(parent.size() * sizeof(int));
        F.pidx.alloc(pidx.size() * sizeof(int));
        F.P64.alloc(P.size() * sizeof(double));
        F.P32.alloc(P.size() * sizeof(float));
        ctx->upload(F.cptr.p, cptr.data(), cptr.size() * sizeof(int));
        ctx->upload(F.cidx.p, cidx.data(), cidx.size() * sizeof(int));
        ctx->upload(F.parent.p, parent.data(), parent.size() * sizeof(int));
        ctx->upload(F.pidx.p, pidx.data(), pidx.size() * sizeof(int));
        ctx->upload(F.P64.p, P.data(), P.size() * sizeof(double));
        convert_f64_f32(ctx, F.P64.as<double>(), F.P32.as<float>(), (int64_t)P.size());
        MgLevel C;
        C.J = Jc;
        C.deg = deg;
        C.bs = bs;
        C.N = (int64_t)Jc * bs;
        for (int d = 0; d < 3; ++d) C.n[d] = nc[d];
        C.lo = loc;
        C.h = hc;
        C.A = make_bsr(ctx, Jc, bs, rp, col);
        C.ownA = true;
        {
            DevBuf dq((nq + 1) * sizeof(int64_t)), dl(plist.size() * sizeof(int)), dr(prow.size() * sizeof(int));
            ctx->upload(dq.p, qptr.data(), qptr.size() * sizeof(int64_t));
            ctx->upload(dl.p, plist.data(), plist.size() * sizeof(int));
            ctx->upload(dr.p, prow.data(), prow.size() * sizeof(int));
            galerkin_gather(ctx, nq, bs, dq.as<int64_t>(), dl.as<int>(), dr.as<int>(), F.A->col.as<int>(),
                            F.A->vals.as<double>(), F.pidx.as<int>(), F.P64.as<double>(), C.A->vals.as<double>());
            ctx->sync();
        }
        lv.push_back(std::move(C));
    }
    // smoother inverses and work vectors
    const bool f32 = prec == BH_FP32;
    const size_t es = f32 ? sizeof(float) : sizeof(double);
    for (size_t l = 0; l + 1 < lv.size(); ++l) {
        MgLevel& L = lv[l];
        bsr_invert_diag(L.A, L.dinv64, f32 ? &L.dinv32 : nullptr);
        if (f32) { L.dinv64.release(); if (!L.op) L.A->ensure_fp32(); }
        // FP16-MFMA cycle: the fine level's smoother blocks in FP16 as well (half the bytes of the
        // largest smoother sweep; the blocks are scaled per block).  Small blocks (the k = 1 level)
        // stay FP32: there the FP16 sweep measured slower.
        const bool hbs = L.bs == 10 || L.bs == 15 || L.bs == 20 || L.bs == 21 || L.bs == 28 || L.bs == 35;
        if (f32 && mfma && l == 0 && hbs) {
            L.dinv16.alloc((size_t)L.J * L.bs * L.bs * sizeof(_Float16));
            L.dscale.alloc((size_t)L.J * sizeof(float));
            blocks_to_f16(ctx, L.J, L.bs, L.dinv32.as<float>(), L.dinv16.as<_Float16>(), L.dscale.as<float>());
        }
    }
    for (auto& L : lv) {
        L.b.alloc(L.N * es);
        L.x.alloc(L.N * es);
        L.r.alloc(L.N * es);
    }
    // FP16-MFMA cycle: the level-0 operator apply on the FP16 matrix cores or the FP32
    // vector kernel (modal face form), whichever is faster on this operator; the FP16
    // smoother blocks stay either way
    l0_mfma_apply = true;
    if (f32 && mfma && lv[0].op) {
        MgLevel& L = lv[0];
        fill32(ctx, L.x.as<float>(), 1.0f, L.N);
        hipEvent_t e0, e1;
        BH_HIP(hipEventCreate(&e0));
        BH_HIP(hipEventCreate(&e1));
        auto time_it = [&](int p) {
            sip_apply(L.op, 1.0, L.x.p, 0.0, L.r.p, false, p);   // first call: the kernel's own tuner
            BH_HIP(hipEventRecord(e0, ctx->stream));
            for (int i = 0; i < 3; ++i) sip_apply(L.op, 1.0, L.x.p, 0.0, L.r.p, false, p);
            BH_HIP(hipEventRecord(e1, ctx->stream));
            BH_HIP(hipEventSynchronize(e1));
            float ms = 0;
            BH_HIP(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        const float t16 = time_it(BH_FP16_MFMA), t32 = time_it(BH_FP32);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        l0_mfma_apply = t16 < t32;
    }
    if (kcycle_level >= (int)lv.size() - 1) kcycle_level = -1;   // the coarsest level is solved directly
    if (kcycle_level >= 0 && kin > 0) {
        MgLevel& K = lv[kcycle_level];
        K.kv.alloc((size_t)(kin + 1) * K.N * es);
        K.kz.alloc((size_t)kin * K.N * es);
        K.kw.alloc(K.N * es);
        BH_CHECK(kin <= 16, "multigrid: at most 16 inner K-cycle steps");
        K.ks.alloc(KcycleScalars(kin).total * sizeof(double));
        fill(ctx, K.ks.as<double>(), 0.0, KcycleScalars(kin).total);
    }
    // coarsest: dense inverse
    MgLevel& Cc = lv.back();
    ncoarse = (int)Cc.N;
    BH_CHECK(ncoarse <= 4096, "multigrid: coarsest level too large for the dense solve");
    std::vector<double> dense((size_t)ncoarse * ncoarse, 0.0);
    {
        const int bs = Cc.bs;
        std::vector<double> v((size_t)Cc.A->nnzb * bs * bs);
        ctx->download(v.data(), Cc.A->vals.p, v.size() * sizeof(double));
        for (int r = 0; r < Cc.J; ++r)
            for (int64_t p = Cc.A->h_rowptr[r]; p < Cc.A->h_rowptr[r + 1]; ++p) {
                const int c = Cc.A->h_col[p];
                for (int i = 0; i < bs; ++i)
                    for (int j = 0; j < bs; ++j)
                        dense[(size_t)(r * bs + i) * ncoarse + c * bs + j] = v[(size_t)p * bs * bs + (size_t)j * bs + i];
            }
    }
    {
        DevBuf Ad(dense.size() * sizeof(double));
        ctx->upload(Ad.p, dense.data(), dense.size() * sizeof(double));
        Ainv64.alloc(dense.size() * sizeof(double));
        dense_inverse_dev(ctx, Ad.as<double>(), Ainv64.as<double>(), ncoarse);
    }
    if (prec == BH_FP32) {
        Ainv32.alloc(dense.size() * sizeof(float));
        convert_f64_f32(ctx, Ainv64.as<double>(), Ainv32.as<float>(), (int64_t)dense.size());
    }
    ctx->sync();
}

template <typename T>
void PMultigrid::residual(int l, const T* b, const T* x, T* r) {
    MgLevel& L = lv[l];
    const int p = sizeof(T) == 4 ? BH_FP32 : BH_FP64;
    if (L.op) {
        sip_apply(L.op, -1.0, x, 0.0, r, false, op_prec(l, p == BH_FP32), b);   // r = b - A x in one pass
    } else {
        bsr_residual(L.A, b, x, r, p);
    }
}

template <typename T>
void PMultigrid::apply(int l, const T* x, T* y) {
    MgLevel& L = lv[l];
    const bool f32 = sizeof(T) == 4;
    if (L.op) sip_apply(L.op, 1.0, x, 0.0, y, false, op_prec(l, f32));
    else bsr_spmv(L.A, 1.0, x, 0.0, y, f32 ? BH_FP32 : BH_FP64);
}

template <typename T>
void PMultigrid::vcycle(int l, const T* b, T* x) {
    MgLevel& L = lv[l];
    T* r = vec<T>(L.r);
    const bool f32 = sizeof(T) == 4;
    if (l + 1 == (int)lv.size()) {
        if (sizeof(T) == 4 && Ainv32.p) dense_gemv<T, float>(ctx, ncoarse, Ainv32.as<float>(), b, x);
        else dense_gemv<T, double>(ctx, ncoarse, Ainv64.as<double>(), b, x);
        return;
    }
    if (l == 0 && level_pmg) {
        // LevelPmg (AS/LevelPmg.cs:561-600): the low-order correction from x = 0 (the
        // residual is b), then the high-order modes against the new residual
        restrict_to<T>(0, b);
        solve_level<T>(1);
        if constexpr (sizeof(T) == 4) fill32(ctx, (float*)x, 0.0f, L.N);
        else fill(ctx, (double*)x, 0.0, L.N);
        prolong_add<T>(0, x);
        residual<T>(0, b, x, r);
        hi_smooth<T>(ctx, L.J, L.bs, lp_lo, (const T*)(f32 ? lp32.p : lp64.p), r, x);
        return;
    }
    // pre-smoothing from a zero guess: the first sweep is x = omega M^-1 b
    smooth<T>(l, b, x, false);
    for (int s = 1; s < pre; ++s) {
        residual<T>(l, b, x, r);
        smooth<T>(l, r, x, true);
    }
    residual<T>(l, b, x, r);
    restrict_to<T>(l, r);
    solve_level<T>(l + 1);
    prolong_add<T>(l, x);
    for (int s = 0; s < post; ++s) {
        residual<T>(l, b, x, r);
        smooth<T>(l, r, x, true);
    }
}

template <typename T>
void PMultigrid::restrict_to(int l, const T* r) {
    MgLevel& L = lv[l];
    MgLevel& C = lv[l + 1];
    const bool f32 = sizeof(T) == 4;
    const T* P = (const T*)(f32 ? L.P32.p : L.P64.p);
    const T* Pt = (const T*)(f32 ? L.eqP32.p : L.eqP64.p);
    if (equilib) rect_restrict<T>(ctx, C.J, L.bs, C.bs, L.eqcptr.as<int>(), L.eqcidx.as<int>(), Pt, r, vec<T>(C.b));
    else if (L.kind == 0) p_restrict<T>(ctx, L.J, L.bs, C.bs, r, vec<T>(C.b));
    else agg_restrict<T>(ctx, C.J, L.bs, L.cptr.as<int>(), L.cidx.as<int>(), L.pidx.as<int>(), P, r, vec<T>(C.b));
}

template <typename T>
void PMultigrid::prolong_add(int l, T* x) {
    MgLevel& L = lv[l];
    MgLevel& C = lv[l + 1];
    const bool f32 = sizeof(T) == 4;
    const T* P = (const T*)(f32 ? L.P32.p : L.P64.p);
    const T* Pt = (const T*)(f32 ? L.eqP32.p : L.eqP64.p);
    if (equilib) rect_prolong_add<T>(ctx, L.J, L.bs, C.bs, L.eqparent.as<int>(), Pt, vec<T>(C.x), x);
    else if (L.kind == 0) p_prolong_add<T>(ctx, L.J, L.bs, C.bs, vec<T>(C.x), x);
    else agg_prolong_add<T>(ctx, L.J, L.bs, L.parent.as<int>(), L.pidx.as<int>(), P, vec<T>(C.x), x);
}

// LevelPmg with UseDiagonalPmg (AS/LevelPmg.cs:232-330): the coarse system is the next
// level (the modes of degree <= its degree: the first lv[1].bs modes of a cell); the
// high-order modes of each cell are solved with the inverse of their part of the
// cell's diagonal block (the reference factorises it by LU; here it is inverted once,
// like the block-Jacobi blocks).  Replaces the finest level's smoothing.
void PMultigrid::set_level_pmg(bool on) {
    drop_graphs();
    level_pmg = false;
    lp64.release();
    lp32.release();
    if (!on) return;
    BH_CHECK(!equilib, "multigrid: LevelPmg and DiagBlockEquilib are exclusive");
    BH_CHECK(lv.size() >= 2 && lv[0].kind == 0, "multigrid: LevelPmg needs a p-level (degree truncation) below the finest");
    MgLevel& L = lv[0];
    const int lo = lv[1].bs, nh = L.bs - lo;
    BH_CHECK(nh > 0 && nh <= 128, "multigrid: LevelPmg needs 1 to 128 high-order modes per cell");
    L.A->ensure_diag();
    DevBuf dense((size_t)L.J * nh * nh * sizeof(double));
    hi_blocks(ctx, L.J, L.bs, lo, L.A, dense.as<double>());
    lp64.alloc((size_t)L.J * nh * nh * sizeof(double));
    if (prec == BH_FP32) lp32.alloc((size_t)L.J * nh * nh * sizeof(float));
    invert_dense_blocks(ctx, L.J, nh, dense.as<double>(), lp64.as<double>(), prec == BH_FP32 ? lp32.as<float>() : nullptr);
    if (prec == BH_FP32) lp64.release();
    lp_lo = lo;
    level_pmg = true;
}

// CellILU (AS/CellILU.cs:1229-1310 ComputeILU, :1490-1560 GetPattern, :182-250 Solve) as
// the finest level's smoother: the ILU(p) pattern is the block pattern of A^(p+1); the
// factorisation restricted to it runs row by row in IKJ order (for every block the same
// operations in the same order as the reference's k-i-j loops), rows grouped by
// dependency level (a row depends on the rows k < i of its pattern row), one launch per
// level; the diagonal blocks are inverted (the reference keeps their LU factors).  The
// smoother is x (+)= omega U^-1 L^-1 r: a forward sweep over the levels, a backward one
// in reverse order; in the FP32 cycle the factors are rounded to FP32 (the reference
// back-substitutes in single precision too).  p < 0 switches it off.
void PMultigrid::set_cell_ilu(int p) {
    drop_graphs();
    ilu_p = -1;
    ilu_nlev = 0;
    ilu_lvptr.clear();
    for (DevBuf* b : {&ilu_rows, &ilu_prp, &ilu_pcol, &ilu_pdiag, &ilu_W64, &ilu_W32, &ilu_U64, &ilu_U32, &ilu_y, &ilu_z})
        b->release();
    if (p < 0) return;
    BH_CHECK(p <= 2, "multigrid: cell ILU level must be 0, 1 or 2");
    BH_CHECK(!equilib, "multigrid: cell ILU and DiagBlockEquilib are exclusive");
    BH_CHECK(lv.size() >= 2, "multigrid: cell ILU needs a hierarchy");
    MgLevel& L = lv[0];
    bh_bsr* A = L.A;
    const int J = L.J, bs = L.bs, E = bs * bs;
    BH_CHECK(bs <= 56, "multigrid: cell ILU supports blocks up to 56 (k <= 5 in 3D)");
    // pattern rows
    std::vector<std::vector<int>> base(J), pat;
    for (int i = 0; i < J; ++i) base[i].assign(A->h_col.begin() + A->h_rowptr[i], A->h_col.begin() + A->h_rowptr[i + 1]);
    for (auto& r : base) std::sort(r.begin(), r.end());
    pat = base;
    std::vector<char> mark(J, 0);
    for (int it = 0; it < p; ++it) {
        std::vector<std::vector<int>> nxt(J);
        for (int i = 0; i < J; ++i) {
            for (int k : pat[i])
                for (int j : base[k])
                    if (!mark[j]) { mark[j] = 1; nxt[i].push_back(j); }
            for (int j : nxt[i]) mark[j] = 0;
            std::sort(nxt[i].begin(), nxt[i].end());
        }
        pat.swap(nxt);
    }
    std::vector<int64_t> prp(J + 1, 0), pdiag(J, -1);
    std::vector<int> pcol;
    for (int i = 0; i < J; ++i) {
        for (int j : pat[i]) {
            if (j == i) pdiag[i] = (int64_t)pcol.size();
            pcol.push_back(j);
        }
        prp[i + 1] = (int64_t)pcol.size();
        BH_CHECK(pdiag[i] >= 0, "multigrid: cell ILU needs every diagonal block");
    }
    // dependency levels and the rows grouped by level (ascending row index within a level)
    std::vector<int> lev(J, 0);
    int nlev = 0;
    for (int i = 0; i < J; ++i) {
        int m = -1;
        for (int64_t q = prp[i]; q < pdiag[i]; ++q) m = std::max(m, lev[pcol[q]]);
        lev[i] = m + 1;
        nlev = std::max(nlev, lev[i] + 1);
    }
    std::vector<int> lvptr(nlev + 1, 0), rows(J);
    for (int i = 0; i < J; ++i) ++lvptr[lev[i] + 1];
    for (int l = 0; l < nlev; ++l) lvptr[l + 1] += lvptr[l];
    {
        std::vector<int> pos(lvptr.begin(), lvptr.end() - 1);
        for (int i = 0; i < J; ++i) rows[pos[lev[i]]++] = i;
    }
    // A's blocks at their pattern positions
    std::vector<int64_t> wpos(std::max<int64_t>(A->nnzb, 1));
    for (int i = 0; i < J; ++i)
        for (int64_t q = A->h_rowptr[i]; q < A->h_rowptr[i + 1]; ++q) {
            const auto it = std::lower_bound(pcol.begin() + prp[i], pcol.begin() + prp[i + 1], A->h_col[q]);
            wpos[q] = it - pcol.begin();
        }
    ilu_rows.alloc(J * sizeof(int));
    ilu_prp.alloc((J + 1) * sizeof(int64_t));
    ilu_pcol.alloc(pcol.size() * sizeof(int));
    ilu_pdiag.alloc(J * sizeof(int64_t));
    ctx->upload(ilu_rows.p, rows.data(), J * sizeof(int));
    ctx->upload(ilu_prp.p, prp.data(), (J + 1) * sizeof(int64_t));
    ctx->upload(ilu_pcol.p, pcol.data(), pcol.size() * sizeof(int));
    ctx->upload(ilu_pdiag.p, pdiag.data(), J * sizeof(int64_t));
    DevBuf dwpos(wpos.size() * sizeof(int64_t)), st(sizeof(int));
    ctx->upload(dwpos.p, wpos.data(), wpos.size() * sizeof(int64_t));
    const int64_t nnzp = (int64_t)pcol.size();
    ilu_W64.alloc((size_t)nnzp * E * sizeof(double));
    ilu_U64.alloc((size_t)J * E * sizeof(double));
    fill(ctx, ilu_W64.as<double>(), 0.0, nnzp * E);
    ilu_scatter(ctx, A->nnzb, bs, dwpos.as<int64_t>(), A->vals.as<double>(), ilu_W64.as<double>());
    const int zero = 0;
    ctx->upload(st.p, &zero, sizeof(int));
    for (int l = 0; l < nlev; ++l)
        ilu_factor_level(ctx, ilu_rows.as<int>() + lvptr[l], lvptr[l + 1] - lvptr[l], bs, ilu_prp.as<int64_t>(),
                         ilu_pcol.as<int>(), ilu_pdiag.as<int64_t>(), ilu_W64.as<double>(), ilu_U64.as<double>(), st.as<int>());
    int h = 0;
    ctx->download(&h, st.p, sizeof(int));
    BH_CHECK(h == 0, "multigrid: cell ILU: singular diagonal block in row " + std::to_string(h - 1));
    const size_t es = prec == BH_FP32 ? sizeof(float) : sizeof(double);
    if (prec == BH_FP32) {
        ilu_W32.alloc((size_t)nnzp * E * sizeof(float));
        ilu_U32.alloc((size_t)J * E * sizeof(float));
        convert_f64_f32(ctx, ilu_W64.as<double>(), ilu_W32.as<float>(), nnzp * E);
        convert_f64_f32(ctx, ilu_U64.as<double>(), ilu_U32.as<float>(), (int64_t)J * E);
        ctx->sync();
        ilu_W64.release();
        ilu_U64.release();
    }
    ilu_y.alloc(L.N * es);
    ilu_z.alloc(L.N * es);
    ilu_lvptr = lvptr;
    ilu_nlev = nlev;
    ilu_p = p;
}

// smoother of level l: block-Jacobi on the cell blocks (AS/BlockJacobi.cs:196-255), or
// additive Schwarz on the aggregates of the level when enabled (AS/Schwarz.cs, MG-aggregate
// blocking, no overlap)
template <typename T>
void PMultigrid::smooth(int l, const T* r, T* x, bool add) {
    MgLevel& L = lv[l];
    const bool f32 = sizeof(T) == 4;
    if (l == 0 && ilu_nlev > 0) {
        ilu_sweeps<T>(ctx, ilu_nlev, ilu_lvptr, ilu_rows.as<int>(), L.bs, ilu_prp.as<int64_t>(), ilu_pcol.as<int>(),
                      ilu_pdiag.as<int64_t>(), (const T*)(f32 ? ilu_W32.p : ilu_W64.p), (const T*)(f32 ? ilu_U32.p : ilu_U64.p), r,
                      (T*)ilu_y.p, (T*)ilu_z.p, x, omega, add);
    } else if (equilib && L.eqL64.p) {
        // block-equilibrated level: the diagonal blocks are diag(S) (signs), block-Jacobi is x (+)= omega S r
        sign_smooth<T>(ctx, L.N, omega, (const T*)(f32 ? L.eqS32.p : L.eqS64.p), r, x, add);
    } else if (L.sw_cmax > 0 && L.sw_overlap > 0)
        schwarz_apply_overlap<T>(ctx, L.sw_nagg, L.bs, L.sw_cmax, L.sw_bptr.as<int>(), L.sw_bidx.as<int>(), L.sw_iptr.as<int>(),
                                 L.sw_iidx.as<int>(), L.J, (const T*)(f32 ? L.sw32.p : L.sw64.p), r,
                                 (T*)(f32 ? L.sw_y32.p : L.sw_y64.p), x, omega, add, L.sw_scale);
    else if (L.sw_cmax > 0)
        schwarz_apply<T>(ctx, L.sw_nagg, L.bs, L.sw_cmax, L.cptr.as<int>(), L.cidx.as<int>(),
                         (const T*)(f32 ? L.sw32.p : L.sw64.p), r, x, omega, add);
    else if (f32 && L.dinv16.p)
        block_gemv_h(ctx, L.J, L.bs, L.dinv16.as<_Float16>(), L.dscale.as<float>(), (const float*)r, (float*)x, omega, add);
    else
        block_gemv<T>(ctx, L.J, L.bs, (const T*)(f32 ? L.dinv32.p : L.dinv64.p), r, x, omega, add);
}

// DiagBlockEquilib (AS/MultigridOperator_ChangeOfBasis.cs:819-828, SymmInv :624): every level
// but the coarsest runs in the basis in which its diagonal blocks are diagonal matrices of
// signs (the identity for an SPD operator, minus it for the SIP form's sign).  Per level,
// L_j = C_j^-1 with D_j = C_j S_j C_j^T; the level matrix becomes L A L^T (blocks
// L_i A_ij L_j^T), vectors x~ = C^T x, right-hand sides b~ = L b, the transfers
// Pt_i = C_i^T P_i L_I^T (and Pt^T), and block-Jacobi is x~ (+)= omega S r~.  The coarsest
// level stays in the original basis (L = I).  For assembled hierarchies; one-way.
void PMultigrid::set_equilib() {
    if (equilib) return;
    BH_CHECK(!matrix_free, "multigrid: DiagBlockEquilib needs the assembled hierarchy (matrix_free = 0)");
    BH_CHECK(!learned, "multigrid: DiagBlockEquilib and the learned correction are exclusive");
    for (auto& L : lv) BH_CHECK(L.sw_cmax == 0, "multigrid: DiagBlockEquilib and the Schwarz smoother are exclusive");
    drop_graphs();
    const int nl = (int)lv.size();
    std::vector<DevBuf> C(nl);
    for (int l = 0; l + 1 < nl; ++l) {
        chol_inv_blocks(lv[l].A, C[l], lv[l].eqL64, lv[l].eqS64);
        if (prec == BH_FP32) {
            lv[l].eqS32.alloc((size_t)lv[l].N * sizeof(float));
            convert_f64_f32(ctx, lv[l].eqS64.as<double>(), lv[l].eqS32.as<float>(), lv[l].N);
        }
    }
    // transfer blocks from the original P and the factors
    for (int l = 0; l + 1 < nl; ++l) {
        MgLevel& F = lv[l];
        MgLevel& Cc = lv[l + 1];
        const size_t nb = (size_t)F.J * F.bs * Cc.bs;
        F.eqP64.alloc(nb * sizeof(double));
        F.eqP32.alloc(nb * sizeof(float));
        const double* Lc = (l + 2 < nl) ? Cc.eqL64.as<double>() : nullptr;   // the coarsest level is not transformed
        std::vector<int> cp, ci, par;
        if (F.kind == 1) {
            transfer_blocks(ctx, F.J, F.bs, Cc.bs, C[l].as<double>(), F.pidx.as<int>(), F.P64.as<double>(), F.parent.as<int>(), Lc,
                            F.eqP64.as<double>(), F.eqP32.as<float>());
        } else {
            transfer_blocks(ctx, F.J, F.bs, Cc.bs, C[l].as<double>(), nullptr, nullptr, nullptr, Lc, F.eqP64.as<double>(),
                            F.eqP32.as<float>());
        }
        // children lists / parents for the rectangular transfers (degree truncation: identity)
        if (F.kind == 1) {
            F.eqcptr.alloc(((size_t)Cc.J + 1) * sizeof(int));
            F.eqcidx.alloc((size_t)F.J * sizeof(int));
            F.eqparent.alloc((size_t)F.J * sizeof(int));
            copy_bytes(ctx, F.eqcptr.p, F.cptr.p, ((size_t)Cc.J + 1) * sizeof(int));
            copy_bytes(ctx, F.eqcidx.p, F.cidx.p, (size_t)F.J * sizeof(int));
            copy_bytes(ctx, F.eqparent.p, F.parent.p, (size_t)F.J * sizeof(int));
        } else {
            cp.resize(F.J + 1);
            ci.resize(F.J);
            for (int i = 0; i <= F.J; ++i) cp[i] = i;
            for (int i = 0; i < F.J; ++i) ci[i] = i;
            F.eqcptr.alloc(cp.size() * sizeof(int));
            F.eqcidx.alloc(ci.size() * sizeof(int));
            F.eqparent.alloc(ci.size() * sizeof(int));
            ctx->upload(F.eqcptr.p, cp.data(), cp.size() * sizeof(int));
            ctx->upload(F.eqcidx.p, ci.data(), ci.size() * sizeof(int));
            ctx->upload(F.eqparent.p, ci.data(), ci.size() * sizeof(int));
        }
    }
    // level matrices in the transformed basis (the finest as a copy: it belongs to the caller)
    for (int l = 0; l + 1 < nl; ++l) {
        MgLevel& L = lv[l];
        if (l == 0) {
            bh_bsr* A0 = make_bsr(ctx, L.A->nrows, L.A->bs, L.A->h_rowptr, L.A->h_col);
            copy_bytes(ctx, A0->vals.p, L.A->vals.p, (size_t)L.A->nnzb * L.A->bs * L.A->bs * sizeof(double));
            eqA0 = A0;
            L.A = A0;
        }
        block_sandwich(L.A, L.eqL64.as<double>());
        if (prec == BH_FP32) L.A->ensure_fp32();
        const size_t nb = (size_t)L.J * L.bs * L.bs;
        if (prec == BH_FP32) {
            L.eqL32.alloc(nb * sizeof(float));
            convert_f64_f32(ctx, L.eqL64.as<double>(), L.eqL32.as<float>(), (int64_t)nb);
        }
    }
    // level 0: L^T for the output transform
    {
        MgLevel& L = lv[0];
        const int bs = L.bs;
        const size_t nb = (size_t)L.J * bs * bs;
        std::vector<double> h(nb), t(nb);
        ctx->download(h.data(), L.eqL64.p, nb * sizeof(double));
        for (int j = 0; j < L.J; ++j)
            for (int c = 0; c < bs; ++c)
                for (int r = 0; r < bs; ++r) t[(size_t)j * bs * bs + (size_t)c * bs + r] = h[(size_t)j * bs * bs + (size_t)r * bs + c];
        L.eqLt64.alloc(nb * sizeof(double));
        ctx->upload(L.eqLt64.p, t.data(), nb * sizeof(double));
        if (prec == BH_FP32) {
            L.eqLt32.alloc(nb * sizeof(float));
            convert_f64_f32(ctx, L.eqLt64.as<double>(), L.eqLt32.as<float>(), (int64_t)nb);
        }
    }
    ctx->sync();
    equilib = true;
}

// Schwarz blocks: the cells of every aggregate of an aggregation level (AS/Schwarz.cs,
// multigrid-aggregate blocking), grown by `overlap` layers of face neighbours on the
// level's grid and sorted (Schwarz.cs:683-730); each block's dense matrix is inverted
// on the device.  Without overlap every DOF is in one block and the correction is
// written directly; with overlap the local solutions are summed per DOF (additive,
// Schwarz.cs:1026-1085) and, with `scale` (EnableOverlapScaling, :518-524, :791-815),
// divided by the number of blocks that hold the DOF.
void PMultigrid::set_schwarz(bool on, int overlap, bool scale) {
    BH_CHECK(overlap >= 0 && overlap <= 2, "multigrid: Schwarz overlap must be 0, 1 or 2");
    drop_graphs();
    const int D = dim;
    for (size_t l = 0; l + 1 < lv.size(); ++l) {
        MgLevel& L = lv[l];
        L.sw_cmax = 0;
        L.sw_overlap = 0;
        L.sw_scale = false;
        for (DevBuf* b : {&L.sw64, &L.sw32, &L.sw_bptr, &L.sw_bidx, &L.sw_iptr, &L.sw_iidx, &L.sw_y64, &L.sw_y32}) b->release();
        if (!on || L.kind != 1) continue;
        const int nagg = lv[l + 1].J;
        if (overlap == 0) {
            const int cmax = 1 << D;
            const int BA = cmax * L.bs;
            if (BA > 64) continue;   // block inversion kernel limit: such levels keep block-Jacobi
            DevBuf dense((size_t)nagg * BA * BA * sizeof(double));
            gather_agg_blocks(ctx, nagg, L.bs, cmax, L.cptr.as<int>(), L.cidx.as<int>(), L.A, dense.as<double>());
            L.sw64.alloc((size_t)nagg * BA * BA * sizeof(double));
            if (prec == BH_FP32) L.sw32.alloc((size_t)nagg * BA * BA * sizeof(float));
            invert_dense_blocks(ctx, nagg, BA, dense.as<double>(), L.sw64.as<double>(), prec == BH_FP32 ? L.sw32.as<float>() : nullptr);
            if (prec == BH_FP32) L.sw64.release();
            L.sw_cmax = cmax;
            L.sw_nagg = nagg;
            continue;
        }
        // the aggregates' children, then the face-neighbour layers
        std::vector<int> cptr(nagg + 1), cidx(L.J);
        ctx->download(cptr.data(), L.cptr.p, (nagg + 1) * sizeof(int));
        ctx->download(cidx.data(), L.cidx.p, L.J * sizeof(int));
        std::vector<int> bptr(nagg + 1, 0), bidx;
        std::vector<char> mark(L.J, 0);
        int bmax = 0;
        for (int I = 0; I < nagg; ++I) {
            std::vector<int> blk(cidx.begin() + cptr[I], cidx.begin() + cptr[I + 1]);
            for (int c : blk) mark[c] = 1;
            for (int ov = 0; ov < overlap; ++ov) {
                const size_t nb0 = blk.size();
                for (size_t q = 0; q < nb0; ++q) {
                    int idx[3];
                    ijk_of(D, L.n, blk[q], idx);
                    for (int d = 0; d < D; ++d)
                        for (int sd = -1; sd <= 1; sd += 2) {
                            int nb[3] = {idx[0], idx[1], idx[2]};
                            nb[d] += sd;
                            if (nb[d] < 0 || nb[d] >= L.n[d]) continue;
                            const int cn = cell_of(D, L.n, nb[0], nb[1], nb[2]);
                            if (!mark[cn]) { mark[cn] = 1; blk.push_back(cn); }
                        }
                }
            }
            std::sort(blk.begin(), blk.end());
            for (int c : blk) mark[c] = 0;
            bidx.insert(bidx.end(), blk.begin(), blk.end());
            bptr[I + 1] = (int)bidx.size();
            bmax = std::max(bmax, (int)blk.size());
        }
        const int BA = bmax * L.bs;
        if (BA > 128) continue;   // the in-LDS inversion's limit: such levels keep block-Jacobi
        // per cell its flat positions in bidx, in block order
        std::vector<int> iptr(L.J + 1, 0), iidx(bidx.size());
        for (int c : bidx) ++iptr[c + 1];
        for (int c = 0; c < L.J; ++c) iptr[c + 1] += iptr[c];
        {
            std::vector<int> fillp(iptr.begin(), iptr.end() - 1);
            for (size_t pq = 0; pq < bidx.size(); ++pq) iidx[fillp[bidx[pq]]++] = (int)pq;
        }
        L.sw_bptr.alloc(bptr.size() * sizeof(int));
        L.sw_bidx.alloc(std::max<size_t>(bidx.size(), 1) * sizeof(int));
        L.sw_iptr.alloc(iptr.size() * sizeof(int));
        L.sw_iidx.alloc(std::max<size_t>(iidx.size(), 1) * sizeof(int));
        ctx->upload(L.sw_bptr.p, bptr.data(), bptr.size() * sizeof(int));
        ctx->upload(L.sw_bidx.p, bidx.data(), bidx.size() * sizeof(int));
        ctx->upload(L.sw_iptr.p, iptr.data(), iptr.size() * sizeof(int));
        ctx->upload(L.sw_iidx.p, iidx.data(), iidx.size() * sizeof(int));
        DevBuf dense((size_t)nagg * BA * BA * sizeof(double));
        gather_agg_blocks(ctx, nagg, L.bs, bmax, L.sw_bptr.as<int>(), L.sw_bidx.as<int>(), L.A, dense.as<double>());
        L.sw64.alloc((size_t)nagg * BA * BA * sizeof(double));
        if (prec == BH_FP32) L.sw32.alloc((size_t)nagg * BA * BA * sizeof(float));
        invert_dense_blocks(ctx, nagg, BA, dense.as<double>(), L.sw64.as<double>(), prec == BH_FP32 ? L.sw32.as<float>() : nullptr);
        if (prec == BH_FP32) L.sw64.release();
        if (prec == BH_FP32) L.sw_y32.alloc(std::max<size_t>(bidx.size(), 1) * L.bs * sizeof(float));
        else L.sw_y64.alloc(std::max<size_t>(bidx.size(), 1) * L.bs * sizeof(double));
        L.sw_cmax = bmax;
        L.sw_nagg = nagg;
        L.sw_overlap = overlap;
        L.sw_scale = scale;
    }
}

template <typename T>
static void axpby_t(bh_ctx* ctx, double a, const T* x, double b, T* y, int64_t n) {
    if (sizeof(T) == 4) axpby32(ctx, (float)a, (const float*)x, (float)b, (float*)y, n);
    else axpby(ctx, a, (const double*)x, b, (double*)y, n);
}

// K-cycle on level l: `kin` steps of flexible GMRES on A_l x = b_l from a zero
// guess, each preconditioned by the V-cycle from level l -- the Krylov-
// accelerated coarse solve of the reference's K-cycle / OrthonormalizationMultigrid
// configurations (AS/OrthonormalizationMultigrid.cs:797-1100, AS/PmgConfig.cs:80-200),
// standing in for the direct low-order solve of the p-two-grid (AS/LevelPmg.cs:411).
// Classical Gram-Schmidt applied twice; every scalar (norms, Hessenberg entries,
// Givens least squares, the combination weights) stays in device memory, so the
// K-cycle has no host synchronisation and is captured with the rest of the cycle.
template <typename T>
void PMultigrid::kcycle(int l) {
    MgLevel& L = lv[l];
    const int64_t n = L.N;
    T* V = vec<T>(L.kv);
    T* Z = vec<T>(L.kz);
    T* w = vec<T>(L.kw);
    double* S = L.ks.as<double>();
    const KcycleScalars K(kin);
    multi_dot_to_dev<T>(ctx, nullptr, vec<T>(L.b), 0, 1, n, S + K.B2);
    scale_inv_norm<T>(ctx, vec<T>(L.b), V, n, S + K.B2, 0.0);
    for (int j = 0; j < kin; ++j) {
        vcycle<T>(l, V + (size_t)j * n, Z + (size_t)j * n);   // Z_j = (V-cycle from this level) V_j
        apply<T>(l, Z + (size_t)j * n, w);
        // classical Gram-Schmidt twice, passes fused: H1 = V^T w; w <- w - V H1 with H2 = V^T w and
        // |w|^2 (in slot H2[j+1]) from the same pass; V_{j+1} = (w - V H2) / |.| (norm by Pythagoras)
        multi_dot_to_dev<T>(ctx, w, V, n, j + 1, n, S + K.H1);
        kc_project_dots<T>(ctx, w, V, n, j + 1, n, S + K.H1, S + K.H2);
        kc_project_scale<T>(ctx, w, V, n, j + 1, n, S + K.H2, S + K.H2 + (j + 1), S + K.W2, V + (size_t)(j + 1) * n);
        kcycle_column(ctx, S, K, j);
    }
    kcycle_lsq(ctx, S, K);
    if (sizeof(T) == 4) fill32(ctx, (float*)vec<T>(L.x), 0.f, n);
    else fill(ctx, (double*)vec<T>(L.x), 0.0, n);
    multi_axpy_from_dev<T>(ctx, vec<T>(L.x), Z, n, kin, n, S + K.Y, nullptr, 1.0);
}

void PMultigrid::attach_learned(Mlp* m) {
    BH_CHECK(!m || !equilib, "learned correction and DiagBlockEquilib are exclusive");
    learned = m;
    drop_graphs();   // the captured cycle includes (or not) the learned step
    if (!m) return;
    BH_CHECK(fine_op, "learned correction needs the fine-level operator");
    BH_CHECK(m->out == fine_op->Np && m->in == fine_op->Np + 4 * fine_op->D,
             "learned correction: network widths do not match the operator (Np + 4 D inputs, Np outputs)");
    const int J = fine_op->grid->J;
    const size_t es = prec == BH_FP32 ? sizeof(float) : sizeof(double);
    lX.alloc((size_t)J * m->in_p * sizeof(float));
    ls.alloc((size_t)J * sizeof(float));
    lY.alloc((size_t)J * m->out_p * sizeof(float));
    le.alloc(N * es);
    lrq.alloc(3 * N * es);
    lstat.alloc(8 * sizeof(double));
    fill(ctx, lstat.as<double>(), 0.0, 8);   // also zeroes the two int counters in the eighth double
    learned_calls = 0;
    learned_off = false;
    learned_mean_gain = 0.0;
    learned_t_cycle = learned_t_net = learned_l_cycle = learned_l_net = 0.0;
}

// `steps` iterations of flexible GMRES (FP64, no restart) on the finest level from
// x = 0, preconditioned by this multigrid with (net) or without the learned step, on
// the last call's right-hand side.  lred: log-reduction of |b - A x| per iteration;
// ms: wall time per iteration (the solver synchronises with the host in every
// iteration, with or without the network alike).  The calls run eagerly, outside the
// probe count, and leave the learned step's statistics as they were.
void PMultigrid::krylov_rate(bool net, int steps, double& lred, double& ms) {
    MgLevel& L = lv[0];
    double keep[8];
    ctx->download(keep, lstat.p, sizeof(keep));
    DevBuf b64(N * sizeof(double)), x64(N * sizeof(double));
    if (prec == BH_FP32) convert_f32_f64(ctx, vec<float>(L.b), b64.as<double>(), N);
    else copy(ctx, b64.as<double>(), vec<double>(L.b), N);
    fill(ctx, x64.as<double>(), 0.0, N);
    LinOp A;
    if (L.op) A.op = L.op;
    else A.M = L.A;
    A.N = N;
    const bool graphs = use_graphs, off = learned_off;
    const int its = iterations;
    use_graphs = false;   // the captured cycle has the learned step or not: run eagerly
    learned_off = !net;
    in_cost_rule = true;
    double t = 0.0;
    double r0 = 0.0, r1 = 0.0;
    int done = 0;
    try {
        GmresSolver gm(A, this, steps + 1, 1e-300, steps, true);
        ctx->sync();
        const auto t0 = std::chrono::steady_clock::now();
        gm.solve_dev(x64.as<double>(), b64.as<double>());
        ctx->sync();
        t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        r0 = gm.res0;
        r1 = gm.res;
        done = gm.iterations;
    } catch (...) {
        use_graphs = graphs;
        learned_off = off;
        in_cost_rule = false;
        iterations = its;
        throw;
    }
    use_graphs = graphs;
    learned_off = off;
    in_cost_rule = false;
    iterations = its;
    ctx->upload(lstat.p, keep, sizeof(keep));
    lred = (done > 0 && r0 > 0 && r1 > 0) ? std::log(r0 / r1) / done : 0.0;
    ms = done > 0 ? t / done : 0.0;
}

// After the cycle (x = z1, the cycle's correction from zero, in lv[0].x):
// r' = b - A z1, q1 = A z1 = b - r', e = network(features of r'), q2 = A e; then
// x = g1 z1 + g2 e with the residual-minimising g (orth2_step).  The network sees
// the residual the cycle leaves, and only a direction that reduces it is added.
template <typename T>
void PMultigrid::learned_correction() {
    MgLevel& L = lv[0];
    const int p = sizeof(T) == 4 ? BH_FP32 : BH_FP64;
    T* rq = vec<T>(lrq);
    T* rp = rq;
    T* q1 = rq + N;
    T* q2 = rq + 2 * N;
    T* e = vec<T>(le);
    const T* b = vec<T>(L.b);
    residual<T>(0, b, vec<T>(L.x), rp);
    sub_vec<T>(ctx, b, rp, q1, N);
    mlp_features(fine_op, rp, p, lX.as<float>(), ls.as<float>(), learned->in_p);
    learned->infer(fine_op->grid->J, lX.as<float>(), lY.as<float>());
    mlp_correction(fine_op, ls.as<float>(), lY.as<float>(), learned->out_p, e, p);
    sip_apply(fine_op, 1.0, e, 0.0, q2, false, op_prec(0, p == BH_FP32));
    double* st = lstat.as<double>();
    multi_dot_to_dev<T>(ctx, q1, q1, N, 2, N, st);       // q1.q1, q1.q2
    multi_dot_to_dev<T>(ctx, q2, q2, 0, 1, N, st + 2);   // q2.q2
    multi_dot_to_dev<T>(ctx, b, q1, N, 2, N, st + 3);    // b.q1, b.q2
    multi_dot_to_dev<T>(ctx, b, b, 0, 1, N, st + 5);     // b.b
    orth2_step<T>(ctx, e, vec<T>(L.x), N, st, (int*)(st + 7));
}

// The cycle from the finest level, the K-cycle inside it and the guarded learned
// step contain no host synchronisation: one graph per preconditioner call.
void PMultigrid::precondition(double* x, const double* b) {
    MgLevel& L = lv[0];
    if (prec == BH_FP32) {
        if (equilib) {   // b~ = L b and x = L^T x~ through the level's residual buffer
            convert_f64_f32(ctx, b, vec<float>(L.r), N);
            block_gemv<float>(ctx, L.J, L.bs, L.eqL32.as<float>(), vec<float>(L.r), vec<float>(L.b), 1.0, false);
            vcycle_g<float>(0);
            block_gemv<float>(ctx, L.J, L.bs, L.eqLt32.as<float>(), vec<float>(L.x), vec<float>(L.r), 1.0, false);
            convert_f32_f64(ctx, vec<float>(L.r), x, N);
        } else {
            convert_f64_f32(ctx, b, vec<float>(L.b), N);
            vcycle_g<float>(0);
            convert_f32_f64(ctx, vec<float>(L.x), x, N);
        }
    } else {
        if (equilib) block_gemv<double>(ctx, L.J, L.bs, L.eqL64.as<double>(), b, vec<double>(L.b), 1.0, false);
        else copy(ctx, vec<double>(L.b), b, N);
        vcycle_g<double>(0);
        if (equilib) block_gemv<double>(ctx, L.J, L.bs, L.eqLt64.as<double>(), vec<double>(L.x), x, 1.0, false);
        else copy(ctx, x, vec<double>(L.x), N);
    }
    iterations++;
    if (learned && !learned_off && !in_cost_rule && learned_probe > 0 && ++learned_calls == learned_probe) {
        double st[8];
        ctx->download(st, lstat.p, sizeof(st));
        learned_mean_gain = st[6] / learned_probe;
        bool keep = learned_mean_gain >= learned_min_gain;
        if (keep && learned_cost_rule) {
            // what the network removes may be residual the Krylov iteration removes
            // anyway: compare flexible GMRES's residual reduction per millisecond
            // (each run twice, the first as a warm-up of both paths)
            constexpr int steps = 12;
            for (int rep = 0; rep < 2; ++rep) {
                krylov_rate(false, steps, learned_l_cycle, learned_t_cycle);
                krylov_rate(true, steps, learned_l_net, learned_t_net);
            }
            keep = learned_t_cycle > 0 && learned_t_net > 0 &&
                   learned_l_net / learned_t_net >= learned_l_cycle / learned_t_cycle;
        }
        if (!keep) {
            learned_off = true;   // the plain multigrid takes over
            drop_graphs();
        }
    }
}

void PMultigrid::learned_stats(int* acc, int* rej) {
    int c[2] = {0, 0};
    if (learned && lstat.p) ctx->download(c, lstat.as<double>() + 7, sizeof(c));
    if (acc) *acc = c[0];
    if (rej) *rej = c[1];
}

void PMultigrid::solve_dev(double* x, const double* b) {
    // x <- x + V(b - A x) in FP64
    DevBuf r(N * sizeof(double)), c(N * sizeof(double));
    if (lv[0].op) {
        copy(ctx, r.as<double>(), b, N);
        sip_apply(lv[0].op, -1.0, x, 1.0, r.p, false, BH_FP64);
    } else {
        bsr_residual(lv[0].A, b, x, r.p, BH_FP64);
    }
    precondition(c.as<double>(), r.as<double>());
    axpby(ctx, 1.0, c.as<double>(), 1.0, x, N);
    converged = 1;
}

template void PMultigrid::restrict_to<float>(int, const float*);
template void PMultigrid::restrict_to<double>(int, const double*);
template void PMultigrid::prolong_add<float>(int, float*);
template void PMultigrid::prolong_add<double>(int, double*);
template void PMultigrid::residual<float>(int, const float*, const float*, float*);
template void PMultigrid::residual<double>(int, const double*, const double*, double*);
template void PMultigrid::vcycle<float>(int, const float*, float*);
template void PMultigrid::vcycle<double>(int, const double*, double*);
template void PMultigrid::vcycle_g<float>(int);
template void PMultigrid::vcycle_g<double>(int);
template void PMultigrid::cycle_body<float>(int);
template void PMultigrid::smooth<float>(int, const float*, float*, bool);
template void PMultigrid::smooth<double>(int, const double*, double*, bool);
template void PMultigrid::cycle_body<double>(int);
template void PMultigrid::kcycle<float>(int);
template void PMultigrid::kcycle<double>(int);

}  // namespace bh

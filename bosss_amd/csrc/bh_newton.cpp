// Newton-Krylov solve of the implicit (backward) Euler step of the DG Euler operator,
// F(U) = U - U_n - dt R(U) = 0, after the reference's Newton solver
// (AS/Newton.cs:195-370 GlobalizedNewton, :690-705 the matrix-free GMRES option,
// :1546-1600 dirder, :880-960 LineSearch):
//  * the Jacobian is applied matrix-free by the directional difference
//    J w = (F(U + eps w) - F(U)) / eps, eps = 1e-7 max(|<U, w>| / |w|, 1) sign(<U, w>) / |w|;
//  * each Newton step solves J s = -F(U) by restarted GMRES (no preconditioner) to the
//    inexact-Newton threshold 1e-5 |F(U)|;
//  * globalisation by the Armijo line search with safeguarded parabolic backtracking
//    (sigma = 1/2 for the first cut, then the three-point parabola clamped to
//    [0.1, 0.5] of the last step, Kelley's parab3p), at most maxStep = 30 cuts;
//  * termination when |F| <= ConvCrit |F_0| + ConvCrit, or after MaxIter steps.
#include <cmath>
#include <memory>

#include "bh_euler.hpp"
#include "bh_solvers.hpp"

namespace bh {

namespace {

// Kelley's parab3p: the minimiser of the parabola through (0, ff0), (lamc, ffc),
// (lamm, ffm), kept in [0.1 lamc, 0.5 lamc]
double parab3p(double lamc, double lamm, double ff0, double ffc, double ffm) {
    const double sigma0 = 0.1, sigma1 = 0.5;
    const double c2 = lamm * (ffc - ff0) - lamc * (ffm - ff0);
    if (c2 >= 0) return sigma1 * lamc;
    const double c1 = lamc * lamc * (ffm - ff0) - lamm * lamm * (ffc - ff0);
    double lamp = -c1 * 0.5 / c2;
    if (lamp < sigma0 * lamc) lamp = sigma0 * lamc;
    if (lamp > sigma1 * lamc) lamp = sigma1 * lamc;
    return lamp;
}

}  // namespace

struct NewtonEuler {
    EulerOp* E;
    bh_ctx* ctx;
    int64_t N;
    double dt;
    DevBuf U0, Fc, Ft, Ut, step, tmp, Fd;
    explicit NewtonEuler(EulerOp* e) : E(e), ctx(e->ctx), N(e->N) {
        for (DevBuf* b : {&U0, &Fc, &Ft, &Ut, &step, &tmp, &Fd}) b->alloc(N * sizeof(double));
    }
    // out = U - U0 - dt R(U)
    void F(const double* U, double* out) {
        E->rhs(U, out, BH_FP64);
        axpby(ctx, 1.0, U, -dt, out, N);
        axpby(ctx, -1.0, U0.as<double>(), 1.0, out, N);
    }
};

// U: initial guess on entry (typically U_n), the solution on exit; info[0] Newton
// steps, info[1] GMRES iterations in total, info[2] line-search cuts; res[0] |F(U_start)|, res[1] |F(U_end)|
void newton_euler(EulerOp* E, const double* Un, double* U, double dt, double conv, int maxit, int restart, int* info,
                  double* res) {
    NewtonEuler ne(E);
    bh_ctx* ctx = ne.ctx;
    const int64_t N = ne.N;
    ne.dt = dt;
    copy(ctx, ne.U0.as<double>(), Un, N);
    double* Fc = ne.Fc.as<double>();
    double* Ft = ne.Ft.as<double>();
    double* Ut = ne.Ut.as<double>();
    double* st = ne.step.as<double>();
    double* tmp = ne.tmp.as<double>();
    double* Fd = ne.Fd.as<double>();
    ne.F(U, Fc);
    double nf = nrm2(ctx, Fc, N);
    const double nf0 = nf;
    int itc = 0, ngm = 0, ncut = 0;
    // matrix-free Jacobian at the current U (Fc = F(U))
    LinOp J;
    J.fctx = ctx;
    J.N = N;
    J.fn = [&](const double* w, double* y) {
        const double nw = nrm2(ctx, w, N);
        if (nw == 0) { fill(ctx, y, 0.0, N); return; }
        const double xs = dot(ctx, U, w, N) / nw;
        double eps = 1e-7;
        if (xs != 0) eps = eps * std::max(std::fabs(xs), 1.0) * (xs > 0 ? 1.0 : -1.0);
        eps /= nw;
        copy(ctx, tmp, U, N);
        axpby(ctx, eps, w, 1.0, tmp, N);
        ne.F(tmp, Fd);
        axpby(ctx, -1.0 / eps, Fc, 1.0 / eps, Fd, N);   // (F(U + eps w) - F(U)) / eps
        copy(ctx, y, Fd, N);
    };
    while (nf > conv * nf0 + conv && itc < maxit) {
        ++itc;
        // inexact Newton step: J s = -F to 1e-5 |F|, solved for the normalised right-hand
        // side -F / |F| (so the GMRES termination |r| < tol |r0| + tol is relative) and scaled back
        GmresSolver gm(J, nullptr, restart, 1e-5, 100, false);
        fill(ctx, st, 0.0, N);
        copy(ctx, Ft, Fc, N);
        axpby(ctx, 0.0, Ft, -1.0 / nf, Ft, N);   // Ft = -F / |F|
        gm.solve_dev(st, Ft);
        axpby(ctx, 0.0, st, nf, st, N);
        ngm += gm.iterations;
        // Armijo line search
        const double alpha = 1e-4, sigma1 = 0.5;
        double lambda = 1.0, lamm = 1.0, lamc = 1.0;
        int iarm = 0;
        copy(ctx, Ut, U, N);
        axpby(ctx, lambda, st, 1.0, Ut, N);
        ne.F(Ut, Ft);
        double nft = nrm2(ctx, Ft, N);
        const double ff0 = nf * nf;
        double ffc = nft * nft, ffm = nft * nft;
        while (nft >= (1 - alpha * lambda) * nf && iarm < 30) {
            lambda = iarm == 0 ? sigma1 * lambda : parab3p(lamc, lamm, ff0, ffc, ffm);
            copy(ctx, Ut, U, N);
            axpby(ctx, lambda, st, 1.0, Ut, N);
            lamm = lamc;
            lamc = lambda;
            ne.F(Ut, Ft);
            nft = nrm2(ctx, Ft, N);
            ffm = ffc;
            ffc = nft * nft;
            ++iarm;
        }
        ncut += iarm;
        copy(ctx, U, Ut, N);
        copy(ctx, Fc, Ft, N);
        nf = nft;
    }
    ctx->sync();
    if (info) { info[0] = itc; info[1] = ngm; info[2] = ncut; }
    if (res) { res[0] = nf0; res[1] = nf; }
}

}  // namespace bh

// Device-resident solvers behind ISolverSmootherTemplate (AS/ISmootherTemplate.cs):
// BlockJacobi, SoftGMRES (left preconditioned) and FlexGMRES (right,
// flexible).  All vectors live in HBM; only the small Hessenberg system and
// the reduction results cross to the host.
#include <algorithm>
#include <cmath>

#include "bh_internal.hpp"
#include "bh_solvers.hpp"
#include "bh_xdgdev.hpp"

namespace bh {

void bsr_residual(bh_bsr* A, const void* b, const void* x, void* r, int prec);

void bh_solver_precondition_default(bh_solver* s, double* x, const double* b) {
    fill(s->ctx, x, 0.0, s->N);
    s->solve_dev(x, b);
}

// ------------------------------------------------------------ LinOp
void LinOp::apply(const double* x, double* y) const {
    if (fn) fn(x, y);
    else if (V) V->spmv(1.0, x, 0.0, y, BH_FP64);
    else if (M) bsr_spmv(M, 1.0, x, 0.0, y, BH_FP64);
    else sip_apply(op, 1.0, x, 0.0, y, false, BH_FP64);
}
void LinOp::residual(const double* b, const double* x, double* r) const {
    if (fn) {
        fn(x, r);
        axpby(ctx(), 1.0, b, -1.0, r, N);
    } else if (V) V->spmv(-1.0, x, 0.0, r, BH_FP64, b);
    else if (M) bsr_residual(M, b, x, r, BH_FP64);
    else {
        copy(ctx(), r, b, N);
        sip_apply(op, -1.0, x, 1.0, r, false, BH_FP64);
    }
}
bh_ctx* LinOp::ctx() const { return fn ? fctx : V ? V->ctx : M ? M->ctx : op->ctx; }

// ------------------------------------------------------------ BlockJacobi
// AS/BlockJacobi.cs:100-255.  NoOfIterations + 1 sweeps (the reference's loop
// runs while iter <= NoOfIterations), each x <- x + omega*Dinv*(b - A x).
BlockJacobiSolver::BlockJacobiSolver(bh_bsr* A_, double omega_, int iters_, int prec_)
    : A(A_), omega(omega_), iters(iters_), prec(prec_) {
    ctx = A->ctx;
    N = (int64_t)A->nrows * A->bs;
    bsr_invert_diag(A, dinv64, prec == BH_FP32 ? &dinv32 : nullptr);
    if (prec == BH_FP32) {
        A->ensure_fp32();
        work32.alloc(3 * N * sizeof(float));
    } else {
        work64.alloc(N * sizeof(double));
    }
}

void BlockJacobiSolver::solve_dev(double* x, const double* b) {
    const int sweeps = iters + 1;
    if (prec == BH_FP64) {
        double* t = work64.as<double>();
        double* cur = x;
        double* nxt = t;
        for (int s = 0; s < sweeps; ++s) {
            bjacobi_sweep(A, dinv64.p, omega, b, cur, nxt, BH_FP64);
            std::swap(cur, nxt);
        }
        if (cur != x) copy(ctx, x, cur, N);
    } else {
        float* b32 = work32.as<float>();
        float* xa = b32 + N;
        float* xb = xa + N;
        convert_f64_f32(ctx, b, b32, N);
        convert_f64_f32(ctx, x, xa, N);
        sweeps_f32(b32, xa, xb, sweeps);
        convert_f32_f64(ctx, (sweeps % 2) ? xb : xa, x, N);
    }
    iterations += sweeps;
    converged = 1;
}

void BlockJacobiSolver::sweeps_f32(const float* b, float* xa, float* xb, int sweeps) {
    float* cur = xa;
    float* nxt = xb;
    for (int s = 0; s < sweeps; ++s) {
        bjacobi_sweep(A, dinv32.p, omega, b, cur, nxt, BH_FP32);
        std::swap(cur, nxt);
    }
}

// ------------------------------------------------------------ GMRES helpers
static void rotmat(double& c, double& s, double a, double b) {
    // AS/SoftGMRES.cs:208-222
    if (b == 0.0) { c = 1.0; s = 0.0; }
    else if (std::fabs(b) > std::fabs(a)) { double t = a / b; s = 1.0 / std::sqrt(1.0 + t * t); c = t * s; }
    else { double t = b / a; c = 1.0 / std::sqrt(1.0 + t * t); s = t * c; }
}

// IterativeSolverConfig.DefaultTermination (AS/IterativeSolverConfig.cs:48-58)
bool Termination::keep_running(int iter, double R0, double R, bool& success) const {
    success = false;
    if (iter <= MinIter) return true;
    if (R < R0 * tol + tol) { success = true; return false; }
    if (iter > MaxIter) return false;
    if (R0 < MachineEpsilon || R < MachineEpsilon) { success = true; return false; }
    return true;
}

// Arnoldi step on V[:, 0..i]: orthogonalise w against V_0..V_{i-1} by
// classical Gram-Schmidt applied twice (CGS2: two batched multi-dot passes
// instead of i dependent dot products of the reference's MGS,
// AS/SoftGMRES.cs:330-345 -- same Krylov basis in exact arithmetic, two
// host synchronisations per step instead of i).
static void cgs2(bh_ctx* ctx, double* w, const double* V, int64_t N, int i, double* hcol) {
    std::vector<double> h1(i), h2(i);
    multi_dot(ctx, w, V, N, i, N, h1.data());
    multi_axpy(ctx, w, V, N, i, N, h1.data(), -1.0);
    multi_dot(ctx, w, V, N, i, N, h2.data());
    multi_axpy(ctx, w, V, N, i, N, h2.data(), -1.0);
    for (int k = 0; k < i; ++k) hcol[k] = h1[k] + h2[k];
}

// Classical Gram-Schmidt with selective reorthogonalisation (the Daniel-Gragg-
// Kaufman-Stewart criterion): w is stored in slot i of V, so one multi-dot pass
// gives <w, V_k> (k < i) and |w|^2.  After the projection, |w'|^2 = |w|^2 - sum h^2;
// only when that drops below |w|^2 / 2 (cancellation: w nearly in span V) is a
// second pass made.  The criterion is known before the projection, so the
// projection, the second pass's multi-dot and the normalisation are fused:
//   no second pass:  w <- (w - V h1) / sqrt(|w|^2 - |h1|^2)          (one pass)
//   second pass:     w <- w - V h1 with h2 = V^T w', |w'|^2 in the same pass;
//                    w <- (w' - V h2) / sqrt(|w'|^2 - |h2|^2)          (two passes)
// (the norms by Pythagoras, well conditioned in both cases).  Returns the norm;
// w is left normalised.  More than 31 vectors: separate passes, explicit norm.
static double cgs_selective(bh_ctx* ctx, double* V, int64_t N, int i, double* hcol, int* passes) {
    double* w = V + (size_t)i * N;
    std::vector<double> h1(i + 1), h2(i);
    multi_dot(ctx, w, V, N, i + 1, N, h1.data());
    const double w2 = h1[i];
    double proj = 0;
    for (int k = 0; k < i; ++k) { hcol[k] = h1[k]; proj += h1[k] * h1[k]; }
    const bool second = !(w2 - proj > 0.5 * w2);
    *passes = second ? 2 : 1;
    if (i >= 1 && i <= 31) {
        if (!second) {
            const double nw = std::sqrt(w2 - proj);
            axpy_scale_dots(ctx, w, V, N, i, N, h1.data(), nw > 0 ? 1.0 / nw : 1.0, nullptr, nullptr);
            return nw;
        }
        double n2 = 0;
        axpy_scale_dots(ctx, w, V, N, i, N, h1.data(), 1.0, h2.data(), &n2);
        double p2 = 0;
        for (int k = 0; k < i; ++k) { hcol[k] += h2[k]; p2 += h2[k] * h2[k]; }
        const double nn = n2 - p2;
        if (nn > 1e-4 * n2) {
            const double nw = std::sqrt(nn);
            axpy_scale_dots(ctx, w, V, N, i, N, h2.data(), 1.0 / nw, nullptr, nullptr);
            return nw;
        }
        // w' still nearly in span V: project again, explicit norm
        axpy_scale_dots(ctx, w, V, N, i, N, h2.data(), 1.0, nullptr, nullptr);
    } else {
        multi_axpy(ctx, w, V, N, i, N, h1.data(), -1.0);
        if (second) {
            multi_dot(ctx, w, V, N, i, N, h2.data());
            multi_axpy(ctx, w, V, N, i, N, h2.data(), -1.0);
            for (int k = 0; k < i; ++k) hcol[k] += h2[k];
        }
    }
    const double nw = nrm2(ctx, w, N);
    if (nw > 0) axpby(ctx, 1.0 / nw, w, 0.0, w, N);
    return nw;
}

static void back_solve(const std::vector<double>& H, int ld, int i, const double* s, double* y) {
    for (int r = i - 1; r >= 0; --r) {
        double a = s[r];
        for (int c = r + 1; c < i; ++c) a -= H[r * ld + c] * y[c];
        y[r] = a / H[r * ld + r];
    }
}

// ------------------------------------------------------------ SoftGMRES
// AS/SoftGMRES.cs:227-455: left preconditioned restarted GMRES(m).
GmresSolver::GmresSolver(LinOp A_, bh_solver* P_, int m_, double tol, int maxit, bool flexible_)
    : A(A_), P(P_), m(m_), flexible(flexible_) {
    ctx = A.ctx();
    N = A.N;
    term.tol = tol;
    term.MaxIter = maxit;
    V.alloc((size_t)(m + 1) * N * sizeof(double));
    if (flexible) Z.alloc((size_t)m * N * sizeof(double));
    r.alloc(N * sizeof(double));
    z.alloc(N * sizeof(double));
    w.alloc(N * sizeof(double));
}

void GmresSolver::precond(const double* in, double* out) {
    if (P) {
        P->precondition(out, in);
    } else {
        copy(ctx, out, in, N);
    }
}

void GmresSolver::solve_dev(double* X, const double* B) {
    if (flexible) return solve_flexible(X, B);
    double* Vp = V.as<double>();
    double* rp = r.as<double>();
    double* zp = z.as<double>();
    double* wp = w.as<double>();
    iterations = 0;
    converged = 0;
    A.residual(B, X, zp);
    precond(zp, rp);
    double error2 = nrm2(ctx, zp, N);
    const double iter0 = error2;
    res0 = iter0;
    bool succ;
    if (!term.keep_running(0, error2, error2, succ)) { converged = succ; res = error2; return; }
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), s(m + 2), y, hcol(m + 1);
    int tot = 0;
    while (true) {
        A.residual(B, X, zp);
        error2 = nrm2(ctx, zp, N);
        precond(zp, rp);
        const double nr = nrm2(ctx, rp, N);
        axpby(ctx, 1.0 / nr, rp, 0.0, Vp, N);
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = nr;
        std::fill(H.begin(), H.end(), 0.0);
        bool done = false;
        int i;
        for (i = 1; i <= m; ++i) {
            iterations++;
            A.apply(Vp + (size_t)(i - 1) * N, zp);
            precond(zp, wp);
            cgs2(ctx, wp, Vp, N, i, hcol.data());
            for (int k = 0; k < i; ++k) H[k * m + i - 1] = hcol[k];
            const double nw = nrm2(ctx, wp, N);
            H[i * m + i - 1] = nw;
            axpby(ctx, 1.0 / nw, wp, 0.0, Vp + (size_t)i * N, N);
            for (int k = 1; k <= i - 1; ++k) {
                double t = cs[k - 1] * H[(k - 1) * m + i - 1] + sn[k - 1] * H[k * m + i - 1];
                H[k * m + i - 1] = -sn[k - 1] * H[(k - 1) * m + i - 1] + cs[k - 1] * H[k * m + i - 1];
                H[(k - 1) * m + i - 1] = t;
            }
            rotmat(cs[i - 1], sn[i - 1], H[(i - 1) * m + i - 1], H[i * m + i - 1]);
            double t = cs[i - 1] * s[i - 1];
            H[(i - 1) * m + i - 1] = cs[i - 1] * H[(i - 1) * m + i - 1] + sn[i - 1] * H[i * m + i - 1];
            H[i * m + i - 1] = 0.0;
            s[i] = -sn[i - 1] * s[i - 1];
            s[i - 1] = t;
            error2 = std::fabs(s[i]);
            if (!term.keep_running(tot, iter0, error2, succ)) {
                y.assign(i, 0.0);
                back_solve(H, m, i, s.data(), y.data());
                multi_axpy(ctx, X, Vp, N, i, N, y.data(), 1.0);
                converged = succ;
                done = true;
                break;
            }
            tot++;
        }
        if (done) break;
        if (!term.keep_running(tot, iter0, error2, succ)) { converged = succ; break; }
        y.assign(m, 0.0);
        back_solve(H, m, m, s.data(), y.data());
        multi_axpy(ctx, X, Vp, N, m, N, y.data(), 1.0);
    }
    A.residual(B, X, zp);
    res = nrm2(ctx, zp, N);
}

// ------------------------------------------------------------ FlexGMRES
// AS/FlexGMRES.cs: right preconditioned, Z_j = P(V_j) kept, x += Z y.
// Convergence: true FP64 residual ||b - A x|| <= tol * ||b|| (checked at
// every restart and whenever the Arnoldi estimate meets the tolerance).
void GmresSolver::solve_flexible(double* X, const double* B) {
    double* Vp = V.as<double>();
    double* Zp = Z.as<double>();
    double* rp = r.as<double>();
    iterations = 0;
    converged = 0;
    reorth_passes = 0;
    const double bn = nrm2(ctx, B, N);
    const double target = term.tol * (bn > 0 ? bn : 1.0);
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), s(m + 2), y, hcol(m + 1);
    A.residual(B, X, rp);
    double beta = nrm2(ctx, rp, N);
    res0 = beta;
    res = beta;
    while (true) {
        if (beta <= target) { converged = 1; break; }
        if (iterations >= term.MaxIter) break;
        axpby(ctx, 1.0 / beta, rp, 0.0, Vp, N);
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = beta;
        std::fill(H.begin(), H.end(), 0.0);
        int i;
        int used = 0;
        for (i = 1; i <= m; ++i) {
            iterations++;
            precond(Vp + (size_t)(i - 1) * N, Zp + (size_t)(i - 1) * N);
            double* vi = Vp + (size_t)i * N;   // w = A z_{i-1}, orthogonalised in place in slot i
            A.apply(Zp + (size_t)(i - 1) * N, vi);
            int passes = 1;
            const double nw = cgs_selective(ctx, Vp, N, i, hcol.data(), &passes);
            reorth_passes += passes - 1;
            for (int k = 0; k < i; ++k) H[k * m + i - 1] = hcol[k];
            H[i * m + i - 1] = nw;   // vi is normalised
            for (int k = 1; k <= i - 1; ++k) {
                double t = cs[k - 1] * H[(k - 1) * m + i - 1] + sn[k - 1] * H[k * m + i - 1];
                H[k * m + i - 1] = -sn[k - 1] * H[(k - 1) * m + i - 1] + cs[k - 1] * H[k * m + i - 1];
                H[(k - 1) * m + i - 1] = t;
            }
            rotmat(cs[i - 1], sn[i - 1], H[(i - 1) * m + i - 1], H[i * m + i - 1]);
            double t = cs[i - 1] * s[i - 1];
            H[(i - 1) * m + i - 1] = cs[i - 1] * H[(i - 1) * m + i - 1] + sn[i - 1] * H[i * m + i - 1];
            H[i * m + i - 1] = 0.0;
            s[i] = -sn[i - 1] * s[i - 1];
            s[i - 1] = t;
            used = i;
            if (std::fabs(s[i]) <= target || iterations >= term.MaxIter || nw == 0) break;
        }
        y.assign(used, 0.0);
        back_solve(H, m, used, s.data(), y.data());
        multi_axpy(ctx, X, Zp, N, used, N, y.data(), 1.0);
        A.residual(B, X, rp);
        beta = nrm2(ctx, rp, N);
        res = beta;
        if (history) history->push_back(beta / (bn > 0 ? bn : 1.0));
    }
}

}  // namespace bh

void bh_solver::precondition(double* x, const double* b) { bh::bh_solver_precondition_default(this, x, b); }

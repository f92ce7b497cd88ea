// OrthonormalizationMultigrid on the finest level of a p/h multigrid hierarchy
// (AS/OrthonormalizationMultigrid.cs): every correction of an iteration -- the
// pre-smoother's, the coarse-grid correction, each post-smoother sweep's and,
// with a network attached, the learned one -- is orthonormalised in the A-image
// (A c against the earlier A c_i) and added with the residual-minimising
// coefficient, so the iterate minimises |b - A x| over the span of all
// corrections of the solve (a GCR whose directions are the multigrid's parts).
// Vectors and coefficients are FP64; the smoother and the coarse levels run in
// the multigrid's precision.
#include <cmath>
#include <string>

#include "bh_mg.hpp"
#include "bh_mlp.hpp"
#include "bh_ortho.hpp"

namespace bh {

OrthoMG::OrthoMG(PMultigrid* m, LinOp a, bh_op* fine, int pre_, int post_sweeps_, int coarse_reps_, int max_hist_, double tol,
                 int maxit)
    : mg(m), A(a), fop(fine), pre_on(pre_), post_sweeps(post_sweeps_), coarse_reps(coarse_reps_), max_hist(max_hist_) {
    ctx = A.ctx();
    N = A.N;
    term.tol = tol;
    term.MaxIter = maxit;
    X.alloc((size_t)max_hist * N * sizeof(double));
    Q.alloc((size_t)max_hist * N * sizeof(double));
    for (DevBuf* b : {&vsol0, &vres0, &vres, &vc, &vq}) b->alloc(N * sizeof(double));
}

double* OrthoMG::Xi(int i) { return X.as<double>() + (size_t)i * N; }
double* OrthoMG::Qi(int i) { return Q.as<double>() + (size_t)i * N; }

// c = S r: the level-0 smoother of the multigrid from a zero guess, `sweeps` sweeps
// (BlockJacobi.Solve with NoOfIterations = sweeps - 1, AS/BlockJacobi.cs:196-255)
template <typename T>
void OrthoMG::smoother_t(const double* r, double* out, int sweeps) {
    MgLevel& L = mg->lv[0];
    T* bT = (T*)L.b.p;
    T* xT = (T*)L.x.p;
    T* rT = (T*)L.r.p;
    if constexpr (sizeof(T) == 4) convert_f64_f32(ctx, r, (float*)bT, N);
    else copy(ctx, (double*)bT, r, N);
    mg->smooth<T>(0, bT, xT, false);
    for (int s = 1; s < sweeps; ++s) {
        mg->residual<T>(0, bT, xT, rT);
        mg->smooth<T>(0, rT, xT, true);
    }
    if constexpr (sizeof(T) == 4) convert_f32_f64(ctx, (const float*)xT, out, N);
    else copy(ctx, out, (const double*)xT, N);
}

// c = P_0 (coarse solve) R_0 r: restriction, the multigrid from level 1 (V- or K-cycle),
// prolongation (the CoarserLevelSolver with CoarseOnLowerLevel, :931-957)
template <typename T>
void OrthoMG::coarse_t(const double* r, double* out) {
    MgLevel& L = mg->lv[0];
    T* xT = (T*)L.x.p;
    T* rT = (T*)L.r.p;
    if constexpr (sizeof(T) == 4) convert_f64_f32(ctx, r, (float*)rT, N);
    else copy(ctx, (double*)rT, r, N);
    mg->restrict_to<T>(0, rT);
    mg->solve_level<T>(1);
    if constexpr (sizeof(T) == 4) fill32(ctx, (float*)xT, 0.0f, N);
    else fill(ctx, (double*)xT, 0.0, N);
    mg->prolong_add<T>(0, xT);
    if constexpr (sizeof(T) == 4) convert_f32_f64(ctx, (const float*)xT, out, N);
    else copy(ctx, out, (const double*)xT, N);
}

void OrthoMG::smoother(const double* r, double* out, int sweeps) {
    if (mg->prec == BH_FP64) smoother_t<double>(r, out, sweeps);
    else smoother_t<float>(r, out, sweeps);
}

void OrthoMG::coarse(const double* r, double* out) {
    if (mg->prec == BH_FP64) coarse_t<double>(r, out);
    else coarse_t<float>(r, out);
}

void OrthoMG::learned_corr(const double* r, double* out) {
    const int J = fop->grid->J;
    mlp_features(fop, r, BH_FP64, nX.as<float>(), ns.as<float>(), net->in_p);
    net->infer(J, nX.as<float>(), nY.as<float>());
    mlp_correction(fop, ns.as<float>(), nY.as<float>(), net->out_p, out, BH_FP64);
}

void OrthoMG::attach_learned(Mlp* m) {
    net = m;
    if (!m) return;
    BH_CHECK(fop, "orthonormalization multigrid: the learned correction needs the fine operator");
    BH_CHECK(m->in >= fop->Np + 4 * fop->D && m->out == fop->Np,
             "orthonormalization multigrid: the network must take Np + 4D inputs and give Np outputs");
    const int J = fop->grid->J;
    nX.alloc((size_t)J * m->in_p * sizeof(float));
    ns.alloc((size_t)J * sizeof(float));
    nY.alloc((size_t)J * m->out_p * sizeof(float));
}

// X = random, q = A X (FillXwithRandom, :224-240); returns |q|
double OrthoMG::fill_random_dir(double* x, double* qq) {
    ++n_random;
    fill_random(ctx, x, N, 0x5DEECE66Dull + 7919ull * (uint64_t)n_random);
    A.apply(x, qq);
    const double nq = nrm2(ctx, qq, N);
    BH_CHECK(nq > 0 && std::isfinite(nq), "orthonormalization multigrid: numerical breakdown: |A x| of a random vector is " +
                                               std::to_string(nq));
    return nq;
}

// AddSol (:219-340): normalise A c, orthogonalise against the history (classical
// Gram-Schmidt, two passes), renormalise; severe cancellation (|A c| <= 1e-5 after
// the passes) restarts from the rescaled remainder, a zero remainder from a random
// vector.  Returns whether cancellation occurred.
bool OrthoMG::add_sol() {
    double* cc = vc.as<double>();
    double* qq = vq.as<double>();
    A.apply(cc, qq);
    double nq = nrm2(ctx, qq, N);
    if (!std::isfinite(nq)) throw Error(-5, "orthonormalization multigrid: numerical breakdown: |A x| = " + std::to_string(nq));
    if (nq == 0) nq = fill_random_dir(cc, qq);
    axpby(ctx, 0.0, cc, 1.0 / nq, cc, N);
    axpby(ctx, 0.0, qq, 1.0 / nq, qq, N);
    constexpr int MaxOrtho = 10;
    bool cancel = false;
    used_more = false;
    std::vector<double> beta(nh);
    for (int jj = 0; jj <= MaxOrtho; ++jj) {
        used_more = jj > 0;
        for (int pass = 0; pass < 2 && nh > 0; ++pass) {
            multi_dot(ctx, qq, Qi(0), N, nh, N, beta.data());
            multi_axpy(ctx, cc, Xi(0), N, nh, N, beta.data(), -1.0);
            multi_axpy(ctx, qq, Qi(0), N, nh, N, beta.data(), -1.0);
        }
        const double nn = nrm2(ctx, qq, N);
        if (nn <= 1e-5) {
            cancel = true;
            ++n_cancel;
            if (jj == MaxOrtho) throw Error(-5, "orthonormalization multigrid: numerical breakdown");
            const double xn = nrm2(ctx, cc, N);
            double nm;
            if (xn < 1e-200 || jj == MaxOrtho - 1) {
                nm = fill_random_dir(cc, qq);
            } else {
                axpby(ctx, 0.0, cc, 1.0 / xn, cc, N);
                A.apply(cc, qq);
                nm = nrm2(ctx, qq, N);
                if (nm == 0) nm = fill_random_dir(cc, qq);
            }
            axpby(ctx, 0.0, cc, 1.0 / nm, cc, N);
            axpby(ctx, 0.0, qq, 1.0 / nm, qq, N);
        } else {
            axpby(ctx, 0.0, cc, 1.0 / nn, cc, N);
            axpby(ctx, 0.0, qq, 1.0 / nn, qq, N);
            break;
        }
    }
    if (nh == max_hist) {
        // history full: restart the minimisation from the current iterate
        ++n_restart;
        copy(ctx, vsol0.as<double>(), xcur, N);
        copy(ctx, vres0.as<double>(), vres.as<double>(), N);
        nh = 0;
    }
    copy(ctx, Xi(nh), cc, N);
    copy(ctx, Qi(nh), qq, N);
    ++nh;
    ++n_added;
    return cancel;
}

// MinimizeResidual (:149-215): x += alpha X_i, r -= alpha A X_i with alpha = <A X_i, r_0>;
// the residual is then recomputed as b - A x (the reference recomputes it at every
// step by default, RecalcResiPeriod = 1); a residual increase beyond 20% throws.
double OrthoMG::minimize_residual(const double* b) {
    double* rr = vres.as<double>();
    const double old = nrm2(ctx, rr, N);
    const double alpha = dot(ctx, Qi(nh - 1), vres0.as<double>(), N);
    axpby(ctx, alpha, Xi(nh - 1), 1.0, xcur, N);
    axpby(ctx, -alpha, Qi(nh - 1), 1.0, rr, N);
    if (used_more || recalc_period <= 1 || nh % recalc_period == 0) A.residual(b, xcur, rr);
    const double rn = nrm2(ctx, rr, N);
    const double tol = std::max(std::max(rn, old) * 0.2, 1.0e-7);
    if (!(rn <= old + tol))
        throw Error(-5, "orthonormalization multigrid: residual increase: old norm " + std::to_string(old) + ", new norm " +
                            std::to_string(rn));
    return rn;
}

double OrthoMG::add_and_minimize(const double* b, bool& cancel) {
    cancel = add_sol();
    return minimize_residual(b);
}

void OrthoMG::solve_dev(double* x, const double* b) {
    xcur = x;
    double* rr = vres.as<double>();
    copy(ctx, vsol0.as<double>(), x, N);
    A.residual(b, x, vres0.as<double>());
    copy(ctx, rr, vres0.as<double>(), N);
    nh = 0;
    const double r0 = nrm2(ctx, rr, N);
    double rn = r0;
    bool skip_pre = false, ok = false, cancel = false;
    int it, completed = 0;
    double* cc = vc.as<double>();
    auto done = [&](int i) { return !term.keep_running(i, r0, rn, ok); };
    for (it = 1;; ++it) {
        if (done(it)) break;
        completed = it;   // an iteration that starts counts
        if (pre_on && !skip_pre) {
            smoother(rr, cc, mg->pre);
            rn = add_and_minimize(b, cancel);
            if (done(it)) break;
        }
        skip_pre = false;
        for (int i = 0; i < coarse_reps; ++i) {
            coarse(rr, cc);
            rn = add_and_minimize(b, cancel);
        }
        if (done(it)) break;
        bool stop = false;
        for (int g = 0; g < post_sweeps; ++g) {
            smoother(rr, cc, mg->post);
            rn = add_and_minimize(b, cancel);
            if (done(it)) { stop = true; break; }
            if (cancel) {
                // the post-smoother adds nothing new: skip the pre-smoother next time (:1047-1055)
                skip_pre = true;
                break;
            }
        }
        if (stop) break;
        if (net) {
            learned_corr(rr, cc);
            rn = add_and_minimize(b, cancel);
        }
    }
    iterations = completed;
    converged = ok ? 1 : 0;
    res0 = r0;
    res = rn;
    xcur = nullptr;
}

}  // namespace bh

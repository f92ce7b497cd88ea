// OrthonormalizationMultigrid solver of libbosss_hip.so (see bh_ortho.cpp).
#pragma once
#include "bh_solvers.hpp"

namespace bh {

struct PMultigrid;
struct Mlp;

struct OrthoMG : bh_solver {
    PMultigrid* mg;                // provides the level-0 smoother and the coarse levels (not owned)
    LinOp A;                       // the FP64 operator of the solve (matrix or matrix-free)
    bh_op* fop;                    // the fine SIP operator (grid and diffusion of the learned features)
    Mlp* net = nullptr;            // optional learned correction (not owned)
    int pre_on, post_sweeps, coarse_reps, max_hist;
    int recalc_period = 1;         // residual recomputed as b - A x every recalc_period corrections
    Termination term;
    DevBuf X, Q;                   // correction history X_i and A X_i (orthonormal), max_hist vectors each
    DevBuf vsol0, vres0, vres, vc, vq;
    DevBuf nX, ns, nY;             // learned correction: features, scales, outputs
    int nh = 0;                    // corrections in the history
    bool used_more = false;        // the last orthonormalisation needed more than one cycle
    double* xcur = nullptr;        // the caller's iterate during a solve
    int n_added = 0, n_cancel = 0, n_random = 0, n_restart = 0;
    OrthoMG(PMultigrid* m, LinOp a, bh_op* fine, int pre, int post_sweeps, int coarse_reps, int max_hist, double tol, int maxit);
    void attach_learned(Mlp* m);
    void solve_dev(double* x, const double* b) override;
    double* Xi(int i);
    double* Qi(int i);
    template <typename T> void smoother_t(const double* r, double* out, int sweeps);
    template <typename T> void coarse_t(const double* r, double* out);
    void smoother(const double* r, double* out, int sweeps);
    void coarse(const double* r, double* out);
    void learned_corr(const double* r, double* out);
    double fill_random_dir(double* x, double* q);
    bool add_sol();
    double minimize_residual(const double* b);
    double add_and_minimize(const double* b, bool& cancel);
};

}  // namespace bh

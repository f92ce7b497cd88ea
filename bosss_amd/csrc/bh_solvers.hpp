// Solver classes of libbosss_hip.so (device resident).
#pragma once
#include <functional>
#include <vector>

#include "bh_internal.hpp"

namespace bh {

struct Vbsr;

// The operator a Krylov solver works on: an assembled BlockMsrMatrix, the
// matrix-free SIP operator (linear part) or a variable-block (XDG) matrix.
struct LinOp {
    bh_bsr* M = nullptr;
    bh_op* op = nullptr;
    Vbsr* V = nullptr;
    // or any operator given as y = A x on device vectors (e.g. a finite-difference Jacobian)
    std::function<void(const double*, double*)> fn;
    bh_ctx* fctx = nullptr;
    int64_t N = 0;
    void apply(const double* x, double* y) const;                  // y = A x
    void residual(const double* b, const double* x, double* r) const;  // r = b - A x
    bh_ctx* ctx() const;
};

struct Termination {
    double tol = 1e-10;
    int MinIter = 0, MaxIter = 10000;
    double MachineEpsilon = 2.2204460492503131e-16 * 1000;
    bool keep_running(int iter, double R0, double R, bool& success) const;
};

struct BlockJacobiSolver : bh_solver {
    bh_bsr* A;
    double omega;
    int iters;
    int prec;
    DevBuf dinv64, dinv32, work64, work32;
    BlockJacobiSolver(bh_bsr* A, double omega, int iters, int prec);
    void solve_dev(double* x, const double* b) override;
    void sweeps_f32(const float* b, float* xa, float* xb, int sweeps);
};

struct GmresSolver : bh_solver {
    LinOp A;
    bh_solver* P;
    int m;
    bool flexible;
    Termination term;
    DevBuf V, Z, r, z, w;
    std::vector<double>* history = nullptr;
    int reorth_passes = 0;   // flexible solve: Arnoldi steps that needed the second Gram-Schmidt pass
    GmresSolver(LinOp A, bh_solver* P, int m, double tol, int maxit, bool flexible);
    void solve_dev(double* x, const double* b) override;
    void solve_flexible(double* x, const double* b);
    void precond(const double* in, double* out);
};

}  // namespace bh

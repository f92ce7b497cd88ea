// Explicit DG Euler / HLLC operator (see bh_euler.hip).
#pragma once
#include "bh_internal.hpp"

namespace bh {

struct EulerOp;
// MFMA right-hand side (bh_eulermfma.hip); false if (D, k, nodes) has no MFMA kernel
bool euler_rhs_mfma(EulerOp* e, const void* U, void* out, int prec);

struct EulerOp {
    bh_ctx* ctx;
    bh_grid* g;
    int D, k, nq1, Np, nv;
    int64_t N;
    double gamma, MS;
    DevBuf work;
    DevBuf tabm64, tabm32;    // MFMA kernel tables (built on first use)
    EulerOp(bh_grid* g, int k, int nq1, double gamma, double mach);
    void rhs(const void* U, void* out, int prec);
    void tvd3(void* U, double dt, int steps, int prec);
    template <typename R> void tvd3_t(R* U, double dt, int steps);
};

// Newton-Krylov solve of the backward-Euler step U - Un - dt R(U) = 0 (bh_newton.cpp)
void newton_euler(EulerOp* E, const double* Un, double* U, double dt, double conv, int maxit, int restart, int* info,
                  double* res);

}  // namespace bh

struct bh_euler {
    bh::EulerOp* e = nullptr;
};

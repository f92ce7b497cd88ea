// Explicit DG Euler / HLLC operator (see bh_euler.hip).
#pragma once
#include "bh_internal.hpp"

namespace bh {

struct EulerOp {
    bh_ctx* ctx;
    bh_grid* g;
    int D, k, nq1, Np, nv;
    int64_t N;
    double gamma, MS;
    DevBuf tab64, tab32, work;
    EulerOp(bh_grid* g, int k, int nq1, double gamma, double mach);
    void rhs(const void* U, void* out, int prec);
    void tvd3(void* U, double dt, int steps, int prec);
    template <typename R> void tvd3_t(R* U, double dt, int steps);
};

}  // namespace bh

struct bh_euler {
    bh::EulerOp* e = nullptr;
};

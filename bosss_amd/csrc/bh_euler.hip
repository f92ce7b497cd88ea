// Explicit DG for the compressible Euler equations with the HLLC flux, on a
// periodic Cartesian grid (gfx950).
//
// What: dU/dt = R(U) = int F(U).grad(phi) - oint Fhat(U-, U+, n) phi, with the
// fluxes of OptimizedHLLC{Density,Momentum,Energy}Flux (Flux() and
// InnerEdgeFlux(), L3-solution/BoSSS.Solution.CompressibleFlowCommon/
// Convection/OptimizedHLLCDensityFlux.cs:45-190 and the Momentum/Energy
// variants) and the non-dimensionalisation MachScaling = gamma*Mach^2 used
// there; time stepping with RungeKuttaScheme.TVD3 (L3-solution/
// BoSSS.Solution/TimeStepping/RungeKuttaScheme.cs:145-167).
//
// How (MI355X): the right-hand side is one launch of euler_rhs_mfma_kernel
// (bh_eulermfma.hip: 16 cells per wave, all modal contractions on the matrix
// cores); this file holds the operator set-up and the TVD3 stage chaining.
#include <cmath>

#include "bh_internal.hpp"
#include "bh_euler.hpp"

namespace bh {

// out = sum_i c_i * v_i  (up to 4 vectors)
template <typename R>
__global__ void lincomb_k(int64_t n, R* __restrict__ out, int m, R c0, const R* __restrict__ v0, R c1, const R* __restrict__ v1,
                          R c2, const R* __restrict__ v2, R c3, const R* __restrict__ v3) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        R s = c0 * v0[i];
        if (m > 1) s += c1 * v1[i];
        if (m > 2) s += c2 * v2[i];
        if (m > 3) s += c3 * v3[i];
        out[i] = s;
    }
}

template <typename R>
static void lincomb(bh_ctx* ctx, int64_t n, R* out, int m, R c0, const R* v0, R c1 = 0, const R* v1 = nullptr, R c2 = 0,
                    const R* v2 = nullptr, R c3 = 0, const R* v3 = nullptr) {
    int64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    hipLaunchKernelGGL((lincomb_k<R>), dim3((unsigned)b), dim3(256), 0, ctx->stream, n, out, m, c0, v0, c1, v1, c2, v2, c3, v3);
}

// ------------------------------------------------------------------ host
EulerOp::EulerOp(bh_grid* g_, int k_, int nq1_, double gamma_, double mach) : g(g_), k(k_), nq1(nq1_), gamma(gamma_) {
    ctx = g->ctx;
    D = g->D;
    MS = gamma * mach * mach;
    Np = basis_count(D, k);
    nv = D + 2;
    N = (int64_t)g->J * nv * Np;
}

void EulerOp::rhs(const void* U, void* out, int prec) {
    if (!euler_rhs_mfma(this, U, out, prec)) throw Error(-3, "Euler: unsupported (D, k, quadrature points) combination");
}

template <typename R>
void EulerOp::tvd3_t(R* U, double dt, int steps) {
    const size_t bytes = N * sizeof(R);
    if (work.bytes < 4 * bytes) work.alloc(4 * bytes);
    R* k1 = (R*)work.p;
    R* k2 = k1 + N;
    R* k3 = k2 + N;
    R* t = k3 + N;
    const int p = sizeof(R) == 4 ? BH_FP32 : BH_FP64;
    const R h = (R)dt;
    for (int s = 0; s < steps; ++s) {
        rhs(U, k1, p);
        lincomb<R>(ctx, N, t, 2, R(1), U, h, k1);
        rhs(t, k2, p);
        lincomb<R>(ctx, N, t, 3, R(1), U, R(0.25) * h, k1, R(0.25) * h, k2);
        rhs(t, k3, p);
        lincomb<R>(ctx, N, U, 4, R(1), U, h / R(6), k1, h / R(6), k2, R(2) * h / R(3), k3);
    }
    BH_HIP(hipGetLastError());
}

void EulerOp::tvd3(void* U, double dt, int steps, int prec) {
    if (prec == BH_FP32) tvd3_t<float>((float*)U, dt, steps);
    else tvd3_t<double>((double*)U, dt, steps);
}

}  // namespace bh

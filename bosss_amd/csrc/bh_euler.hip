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
// How (MI355X): one wave per cell, one lane per quadrature node (4^3 = 64
// volume nodes for k = 2 fill exactly one wave64).  The cell's modes are
// wave-uniform (scalar loads); each lane evaluates the state at its node, the
// physical flux (volume) or the HLLC flux (one face node per lane), and the
// test-function contractions are done by re-mapping lanes to (variable, mode)
// outputs through a per-wave LDS tile.  The result row of a cell is written
// by 50 consecutive lanes (coalesced); no atomics.
#include <cmath>

#include "bh_internal.hpp"
#include "bh_euler.hpp"

namespace bh {

template <typename R>
struct EulerArgs {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ width;
    const R* __restrict__ tab;   // V[nqv][Np] | G[nqv][D][Np] | T[2D][nqf][Np] | vw[nqv] | fw[nqf]
    double gamma, MS;
    const R* __restrict__ U;
    R* __restrict__ out;
};

constexpr int ebinom(int n, int k) {
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
constexpr int epow(int b, int e) { return e == 0 ? 1 : b * epow(b, e - 1); }

template <typename R>
__device__ __forceinline__ void hllc(int D, R gamma, R MS, const R* Ui, const R* Uo, int dn, R sgn, R* F) {
    // normal = sgn * e_dn
    const R rIn = Ui[0], rOut = Uo[0];
    R m2In = 0, m2Out = 0;
    for (int d = 0; d < D; ++d) { m2In += Ui[d + 1] * Ui[d + 1]; m2Out += Uo[d + 1] * Uo[d + 1]; }
    const R vnIn = sgn * Ui[dn + 1] / rIn, vnOut = sgn * Uo[dn + 1] / rOut;
    const R EIn = Ui[D + 1], EOut = Uo[D + 1];
    const R pIn = (gamma - R(1)) * (EIn - MS * R(0.5) * m2In / rIn);
    const R pOut = (gamma - R(1)) * (EOut - MS * R(0.5) * m2Out / rOut);
    const R Mach = sqrt(MS / gamma);
    const R aIn = sqrt(pIn / rIn) / Mach, aOut = sqrt(pOut / rOut) / Mach;
    const R rMean = R(0.5) * (rIn + rOut), pMean = R(0.5) * (pIn + pOut), aMean = R(0.5) * (aIn + aOut);
    const R pStar = fmax(R(0), pMean - R(0.5) * MS * (vnOut - vnIn) * aMean * rMean);
    const R qIn = pStar > pIn ? sqrt(R(1) + R(0.5) * (gamma + R(1)) * (pStar / pIn - R(1)) / gamma) : R(1);
    const R qOut = pStar > pOut ? sqrt(R(1) + R(0.5) * (gamma + R(1)) * (pStar / pOut - R(1)) / gamma) : R(1);
    const R sIn = vnIn - aIn * qIn, sOut = vnOut + aOut * qOut;
    const R cIn = rIn * (sIn - vnIn), cOut = rOut * (sOut - vnOut);
    const R sStar = (cOut * vnOut - cIn * vnIn + (pIn - pOut) / MS) / (cOut - cIn);
    const bool left = sStar > R(0);
    const R* Us = left ? Ui : Uo;
    const R rS = left ? rIn : rOut, vnS = left ? vnIn : vnOut, pS = left ? pIn : pOut, ES = left ? EIn : EOut;
    const R sK = left ? sIn : sOut;
    F[0] = vnS * rS;
    for (int c = 0; c < D; ++c) F[c + 1] = vnS * Us[c + 1] + (c == dn ? sgn * pS / MS : R(0));
    F[D + 1] = vnS * (ES + pS);
    if (left ? (sIn <= R(0)) : (sOut >= R(0))) {
        const R f = (sK - vnS) / (sK - sStar);
        F[0] += sK * (rS * f - rS);
        for (int c = 0; c < D; ++c)
            F[c + 1] += sK * (rS * f * (Us[c + 1] / rS + (c == dn ? sgn * (sStar - vnS) : R(0))) - Us[c + 1]);
        const R factor = rS * sStar * MS + pS / (sK - vnS);
        F[D + 1] += sK * (f * (ES + factor * (sStar - vnS)) - ES);
    }
}

template <int D, int K, int NQ1, typename R>
__global__ __launch_bounds__(256) void euler_rhs_kernel(EulerArgs<R> a) {
    constexpr int Np = ebinom(K + D, D), NV = D + 2, NQV = epow(NQ1, D), NQF = epow(NQ1, D - 1);
    constexpr int NOUT = NV * Np, NSLOT = 2 * D * NQF;
    static_assert(NQV <= 64 && NOUT <= 64, "one wave per cell needs <= 64 nodes and outputs");
    constexpr int OFF_G = NQV * Np, OFF_T = OFF_G + NQV * D * Np, OFF_VW = OFF_T + 2 * D * NQF * Np;
    constexpr int OFF_FW = OFF_VW + NQV, TABN = OFF_FW + NQF;
    constexpr int SCR = (NV * D * NQV > NV * NSLOT) ? NV * D * NQV : NV * NSLOT;
    __shared__ R tab[TABN];
    __shared__ R scr[4][SCR];
    for (int i = threadIdx.x; i < TABN; i += blockDim.x) tab[i] = a.tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    R* S = scr[wave];
    const int nn[3] = {a.n0, a.n1, a.n2};
    for (int c = blockIdx.x * 4 + wave; c < a.J; c += gridDim.x * 4) {
        int idx[3];
        if (D == 2) { idx[0] = c / a.n1; idx[1] = c - idx[0] * a.n1; idx[2] = 0; }
        else { int sl = a.n0 * a.n1; idx[2] = c / sl; int rr = c - idx[2] * sl; idx[1] = rr / a.n0; idx[0] = rr - idx[1] * a.n0; }
        R h[3];
        R detJ = 1;
        for (int d = 0; d < D; ++d) { h[d] = (R)a.width[d * a.nmax + idx[d]]; detJ *= R(0.5) * h[d]; }
        const R sc = R(1) / sqrt(detJ);
        const R* __restrict__ Uc = a.U + (size_t)c * NOUT;
        // ---- volume: lane = node
        if (lane < NQV) {
            R Uq[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                R s = 0;
#pragma unroll
                for (int n = 0; n < Np; ++n) s += tab[lane * Np + n] * Uc[v * Np + n];
                Uq[v] = s * sc;
            }
            const R r = Uq[0], E = Uq[D + 1];
            R m2 = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) m2 += Uq[d + 1] * Uq[d + 1];
            const R p = (R(a.gamma) - R(1)) * (E - R(a.MS) * R(0.5) * m2 / r);
            const R w = tab[OFF_VW + lane] * detJ * sc;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const R wd = w * R(2) / h[d];
                S[(0 * D + d) * NQV + lane] = wd * Uq[d + 1];
#pragma unroll
                for (int cc = 0; cc < D; ++cc)
                    S[((cc + 1) * D + d) * NQV + lane] = wd * (Uq[cc + 1] * Uq[d + 1] / r + (cc == d ? p / R(a.MS) : R(0)));
                S[((D + 1) * D + d) * NQV + lane] = wd * (Uq[d + 1] / r * (E + p));
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        R acc = 0;
        const int ov = lane / Np, om = lane - ov * Np;   // output (variable, mode)
        if (lane < NOUT) {
            for (int q = 0; q < NQV; ++q)
#pragma unroll
                for (int d = 0; d < D; ++d) acc += S[(ov * D + d) * NQV + q] * tab[OFF_G + (q * D + d) * Np + om];
        }
        __builtin_amdgcn_wave_barrier();
        // ---- faces: lane = (face, node) slot
        for (int s0 = 0; s0 < NSLOT; s0 += 64) {
            const int s = s0 + lane;
            if (s < NSLOT) {
                const int f = s / NQF, q = s - f * NQF;
                const int d = f >> 1, sd = f & 1, fo = 2 * d + (1 - sd);
                int nidx[3] = {idx[0], idx[1], idx[2]};
                nidx[d] = (nidx[d] + (sd ? 1 : -1) + nn[d]) % nn[d];
                const int cn = D == 2 ? nidx[0] * a.n1 + nidx[1] : nidx[0] + nidx[1] * a.n0 + nidx[2] * a.n0 * a.n1;
                const R hn = (R)a.width[d * a.nmax + nidx[d]];
                const R scn = R(1) / sqrt(detJ / h[d] * hn);
                const R* __restrict__ Un = a.U + (size_t)cn * NOUT;
                R Ui[NV], Uo[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    R si = 0, so = 0;
#pragma unroll
                    for (int n = 0; n < Np; ++n) {
                        si += tab[OFF_T + (f * NQF + q) * Np + n] * Uc[v * Np + n];
                        so += tab[OFF_T + (fo * NQF + q) * Np + n] * Un[v * Np + n];
                    }
                    Ui[v] = si * sc;
                    Uo[v] = so * scn;
                }
                R F[NV];
                hllc<R>(D, R(a.gamma), R(a.MS), Ui, Uo, d, sd ? R(1) : R(-1), F);
                const R w = tab[OFF_FW + q] * (detJ / (R(0.5) * h[d])) * sc;
#pragma unroll
                for (int v = 0; v < NV; ++v) S[v * NSLOT + s] = w * F[v];
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        if (lane < NOUT) {
            for (int s = 0; s < NSLOT; ++s) acc -= S[ov * NSLOT + s] * tab[OFF_T + s * Np + om];
            a.out[(size_t)c * NOUT + lane] = acc;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

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
    SipTables t;
    t.build(D, k, nq1);
    Np = t.Np;
    nv = D + 2;
    N = (int64_t)g->J * nv * Np;
    std::vector<double> tab;
    std::vector<double> val(Np);
    for (int q = 0; q < t.nqv; ++q) {
        basis_eval(D, k, &t.vxi[q * D], val.data(), nullptr);
        tab.insert(tab.end(), val.begin(), val.end());
    }
    tab.insert(tab.end(), t.G.begin(), t.G.end());
    tab.insert(tab.end(), t.T.begin(), t.T.end());
    tab.insert(tab.end(), t.vw.begin(), t.vw.end());
    tab.insert(tab.end(), t.fw.begin(), t.fw.end());
    tab64.alloc(tab.size() * sizeof(double));
    ctx->upload(tab64.p, tab.data(), tab.size() * sizeof(double));
    std::vector<float> t32(tab.begin(), tab.end());
    tab32.alloc(t32.size() * sizeof(float));
    ctx->upload(tab32.p, t32.data(), t32.size() * sizeof(float));
}

template <typename R, int D, int K, int NQ1>
static void launch_rhs(EulerOp* e, const R* U, R* out) {
    EulerArgs<R> a;
    a.J = e->g->J;
    a.n0 = e->g->n[0]; a.n1 = e->g->n[1]; a.n2 = e->g->n[2];
    a.nmax = e->g->nmax;
    a.width = e->g->d_width.as<double>();
    a.tab = (const R*)(sizeof(R) == 4 ? e->tab32.p : e->tab64.p);
    a.gamma = e->gamma;
    a.MS = e->MS;
    a.U = U;
    a.out = out;
    int blocks = (a.J + 3) / 4;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL((euler_rhs_kernel<D, K, NQ1, R>), dim3(blocks), dim3(256), 0, e->ctx->stream, a);
    BH_HIP(hipGetLastError());
}

void EulerOp::rhs(const void* U, void* out, int prec) {
    if (use_mfma && euler_rhs_mfma(this, U, out, prec)) return;
#define BH_CASE(DD, KK, QQ)                                                                                 \
    if (D == DD && k == KK && nq1 == QQ) {                                                                  \
        if (prec == BH_FP32) launch_rhs<float, DD, KK, QQ>(this, (const float*)U, (float*)out);             \
        else launch_rhs<double, DD, KK, QQ>(this, (const double*)U, (double*)out);                           \
        return;                                                                                             \
    }
    BH_CASE(3, 1, 3) BH_CASE(3, 2, 4) BH_CASE(3, 1, 2) BH_CASE(2, 2, 4) BH_CASE(2, 3, 5) BH_CASE(2, 1, 3)
#undef BH_CASE
    throw Error(-3, "Euler: unsupported (D, k, quadrature points) combination");
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

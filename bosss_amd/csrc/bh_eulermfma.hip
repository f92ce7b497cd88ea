// Explicit DG Euler right-hand side with the HLLC flux on the matrix cores
// (gfx950): the same operator as euler_rhs_kernel (OptimizedHLLC{Density,
// Momentum,Energy}Flux, L3-solution/BoSSS.Solution.CompressibleFlowCommon/
// Convection/OptimizedHLLCDensityFlux.cs:45-190, periodic Cartesian grid),
// with every modal contraction as a 16x16x4 MFMA (FP64, or FP32 for the FP32
// run).
//
// One wave = a tile of 16 consecutive cells.  Per tile:
//   states at volume nodes  U_v[q][c] = sum_n psi[q][n] u_c[v][n]        (A = table rows q, B = modes)
//   volume flux to modes    Y_v[c][m] += sum_{q,d} (w_q 2/h_d F_vd[c][q]) dpsi_m/dxi_d[q]
//   face traces             U_v[q][c] = sum_n T_f[q][n] u_c[v][n]  (own cell and the neighbour across f)
//   face flux to modes      Y_v[c][m] -= sum_q w_q HLLC_v[c][q] T_f[q][m]
// The accumulator of a trace product holds, per lane, one cell and 4 nodes: the
// A-operand layout of the flux product (K slots = those 4 nodes), so the
// pointwise physical and HLLC fluxes are formed in registers between the two
// products.  Tables (psi, gradients, face traces, weights; node rows padded to
// 16 with zero weights) are staged in LDS once per workgroup.
#include <cmath>
#include <vector>

#include "bh_internal.hpp"
#include "bh_euler.hpp"

namespace bh {

typedef double ed4_t __attribute__((ext_vector_type(4)));
typedef float ef4_t __attribute__((ext_vector_type(4)));

namespace {
constexpr int eb(int n, int k) {
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
constexpr int ep(int b, int e) { return e == 0 ? 1 : b * ep(b, e - 1); }

// 16x16x4 MFMA per precision.  Accumulator register r of lane l holds row
// row(l >> 4, r), column l & 15 (FP64: interleaved rows; FP32: blocked rows).
template <typename R> struct Mma;
template <> struct Mma<double> {
    typedef ed4_t acc;
    static __device__ __forceinline__ acc mma(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};
template <> struct Mma<float> {
    typedef ef4_t acc;
    static __device__ __forceinline__ acc mma(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
}  // namespace

template <int D, int K, int NQ1>
struct EulerMfmaLayout {
    static constexpr int Np = eb(K + D, D), NV = D + 2;
    static constexpr int NQV = ep(NQ1, D), NQF = ep(NQ1, D - 1);
    static constexpr int NK = (Np + 3) / 4, NPK = NK * 4;      // K steps over the modes
    static constexpr int NVT = (NQV + 15) / 16;                // 16-node volume tiles
    static constexpr int PSI = 0;                              // [NVT*16][NPK]   psi_n at volume node q
    static constexpr int GT = PSI + NVT * 16 * NPK;            // [D][NVT*16][16] dpsi_m/dxi_d at node q (m padded)
    static constexpr int VW = GT + D * NVT * 16 * 16;          // [NVT*16] volume weights (0 on padding)
    static constexpr int TA = VW + NVT * 16;                   // [2D][16][NPK]   trace rows q, modes n
    static constexpr int TB = TA + 2 * D * 16 * NPK;           // [2D][16][16]    trace T_f[q][m] (m padded)
    static constexpr int FW = TB + 2 * D * 16 * 16;            // [16] face weights (0 on padding)
    static constexpr int TOTAL = FW + 16;
    static_assert(Np <= 16 && NQF <= 16, "Euler MFMA kernel: Np <= 16, face nodes <= 16");
};

template <int D, int K, int NQ1>
void build_euler_mfma_tables(const SipTables& t, std::vector<double>& h) {
    using L = EulerMfmaLayout<D, K, NQ1>;
    h.assign(L::TOTAL, 0.0);
    std::vector<double> val(L::Np);
    for (int q = 0; q < L::NQV; ++q) {
        basis_eval(D, K, &t.vxi[(size_t)q * D], val.data(), nullptr);
        for (int n = 0; n < L::Np; ++n) h[L::PSI + (size_t)q * L::NPK + n] = val[n];
        for (int d = 0; d < D; ++d)
            for (int m = 0; m < L::Np; ++m)
                h[L::GT + ((size_t)d * L::NVT * 16 + q) * 16 + m] = t.G[((size_t)q * D + d) * L::Np + m];
        h[L::VW + q] = t.vw[q];
    }
    for (int f = 0; f < 2 * D; ++f)
        for (int q = 0; q < L::NQF; ++q)
            for (int n = 0; n < L::Np; ++n) {
                const double v = t.T[((size_t)f * L::NQF + q) * L::Np + n];
                h[L::TA + ((size_t)f * 16 + q) * L::NPK + n] = v;
                h[L::TB + ((size_t)f * 16 + q) * 16 + n] = v;
            }
    for (int q = 0; q < L::NQF; ++q) h[L::FW + q] = t.fw[q];
}

template <typename R>
struct EulerMfmaArgs {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ width;
    const R* __restrict__ tab;
    R gamma, MS;
    const R* __restrict__ U;
    R* __restrict__ out;
};

// HLLC as OptimizedHLLC*Flux.InnerEdgeFlux (normal = sgn * e_dn); same algebra as hllc() of the per-node kernel,
// with each denominator inverted once (7 divisions per node instead of 20; rounding-level differences)
template <int D, typename R>
__device__ __forceinline__ void hllc_m(R gamma, R MS, const R* Ui, const R* Uo, int dn, R sgn, R* F) {
    const R rIn = Ui[0], rOut = Uo[0];
    const R irIn = R(1) / rIn, irOut = R(1) / rOut;
    const R iMS = R(1) / MS, igam = R(1) / gamma, iMach = sqrt(gamma * iMS);   // uniform: hoisted by the compiler
    R m2In = 0, m2Out = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) { m2In += Ui[d + 1] * Ui[d + 1]; m2Out += Uo[d + 1] * Uo[d + 1]; }
    const R uIn = dn == 0 ? Ui[1] : (dn == 1 ? Ui[2] : Ui[D == 3 ? 3 : 2]);
    const R uOut = dn == 0 ? Uo[1] : (dn == 1 ? Uo[2] : Uo[D == 3 ? 3 : 2]);
    const R vnIn = sgn * uIn * irIn, vnOut = sgn * uOut * irOut;
    const R EIn = Ui[D + 1], EOut = Uo[D + 1];
    const R pIn = (gamma - R(1)) * (EIn - MS * R(0.5) * m2In * irIn);
    const R pOut = (gamma - R(1)) * (EOut - MS * R(0.5) * m2Out * irOut);
    const R aIn = sqrt(pIn * irIn) * iMach, aOut = sqrt(pOut * irOut) * iMach;
    const R rMean = R(0.5) * (rIn + rOut), pMean = R(0.5) * (pIn + pOut), aMean = R(0.5) * (aIn + aOut);
    const R pStar = fmax(R(0), pMean - R(0.5) * MS * (vnOut - vnIn) * aMean * rMean);
    const R hg = R(0.5) * (gamma + R(1)) * igam;
    const R qIn = pStar > pIn ? sqrt(R(1) + hg * (pStar / pIn - R(1))) : R(1);
    const R qOut = pStar > pOut ? sqrt(R(1) + hg * (pStar / pOut - R(1))) : R(1);
    const R sIn = vnIn - aIn * qIn, sOut = vnOut + aOut * qOut;
    const R cIn = rIn * (sIn - vnIn), cOut = rOut * (sOut - vnOut);
    const R sStar = (cOut * vnOut - cIn * vnIn + (pIn - pOut) * iMS) / (cOut - cIn);
    const bool left = sStar > R(0);
    const R rS = left ? rIn : rOut, irS = left ? irIn : irOut, vnS = left ? vnIn : vnOut, pS = left ? pIn : pOut;
    const R ES = left ? EIn : EOut;
    const R sK = left ? sIn : sOut;
    R Us[D];
#pragma unroll
    for (int c = 0; c < D; ++c) Us[c] = left ? Ui[c + 1] : Uo[c + 1];
    F[0] = vnS * rS;
#pragma unroll
    for (int c = 0; c < D; ++c) F[c + 1] = vnS * Us[c] + (c == dn ? sgn * pS * iMS : R(0));
    F[D + 1] = vnS * (ES + pS);
    if (left ? (sIn <= R(0)) : (sOut >= R(0))) {
        const R f = (sK - vnS) / (sK - sStar);
        F[0] += sK * (rS * f - rS);
#pragma unroll
        for (int c = 0; c < D; ++c)
            F[c + 1] += sK * (rS * f * (Us[c] * irS + (c == dn ? sgn * (sStar - vnS) : R(0))) - Us[c]);
        const R factor = rS * sStar * MS + pS / (sK - vnS);
        F[D + 1] += sK * (f * (ES + factor * (sStar - vnS)) - ES);
    }
}

template <int D, int K, int NQ1, typename R, int W>
__global__ __launch_bounds__(64 * W, 2) void euler_rhs_mfma_kernel(EulerMfmaArgs<R> a) {
    using L = EulerMfmaLayout<D, K, NQ1>;
    using M = Mma<R>;
    typedef typename M::acc acc_t;
    constexpr int Np = L::Np, NV = L::NV, NK = L::NK, NPK = L::NPK;
    __shared__ R tab[L::TOTAL];
    for (int i = threadIdx.x; i < L::TOTAL; i += blockDim.x) tab[i] = a.tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int ntiles = (a.J + 15) / 16;
    const acc_t z = {0, 0, 0, 0};
    for (int tile = blockIdx.x * W + wave; tile < ntiles; tile += gridDim.x * W) {
        const int c = tile * 16 + col;
        const bool valid = c < a.J;
        const int cc = valid ? c : a.J - 1;
        int idx[3];
        if (D == 2) { idx[0] = cc / a.n1; idx[1] = cc - idx[0] * a.n1; idx[2] = 0; }
        else { const int sl = a.n0 * a.n1; idx[2] = cc / sl; const int rr = cc - idx[2] * sl; idx[1] = rr / a.n0; idx[0] = rr - idx[1] * a.n0; }
        R h[3] = {1, 1, 1}, detJ = 1;
        R ih2[3] = {2, 2, 2};    // 2 / h_d from the grid's reciprocal widths
        for (int d = 0; d < D; ++d) {
            h[d] = (R)a.width[d * a.nmax + idx[d]];
            ih2[d] = R(2) * (R)a.width[(3 + d) * a.nmax + idx[d]];
            detJ *= R(0.5) * h[d];
        }
        const R sc = R(1) / sqrt(detJ);
        // own modes in the B-operand layout: variable v, mode 4s + grp
        R bU[NV][NK];
        const R* __restrict__ Uc = a.U + (size_t)cc * NV * Np;
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
            for (int s = 0; s < NK; ++s) {
                const int n = 4 * s + grp;
                bU[v][s] = (valid && n < Np) ? Uc[v * Np + n] : R(0);
            }
        acc_t y[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) y[v] = z;
        // ---------------- volume
#pragma unroll 1
        for (int t = 0; t < L::NVT; ++t) {
            acc_t Ut[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                Ut[v] = z;
#pragma unroll
                for (int s = 0; s < NK; ++s) Ut[v] = M::mma(tab[L::PSI + (t * 16 + col) * NPK + 4 * s + grp], bU[v][s], Ut[v]);
            }
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = t * 16 + M::row(grp, r);          // node of register r for this lane's cell
                R Uq[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) Uq[v] = Ut[v][r] * sc;
                const bool live = valid && q < L::NQV;
                const R rho = live ? Uq[0] : R(1), E = Uq[D + 1];
                const R irho = R(1) / rho;
                R m2 = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) m2 += Uq[d + 1] * Uq[d + 1];
                const R p = (a.gamma - R(1)) * (E - a.MS * R(0.5) * m2 * irho);
                const R pMS = p / a.MS;
                const R w = live ? tab[L::VW + q] * detJ * sc : R(0);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const R wd = w * ih2[d];
                    const R ud = Uq[d + 1] * irho;
                    const R Bg = tab[L::GT + (d * L::NVT * 16 + q) * 16 + col];
                    y[0] = M::mma(wd * Uq[d + 1], Bg, y[0]);
#pragma unroll
                    for (int cc2 = 0; cc2 < D; ++cc2)
                        y[cc2 + 1] = M::mma(wd * (Uq[cc2 + 1] * ud + (cc2 == d ? pMS : R(0))), Bg, y[cc2 + 1]);
                    y[D + 1] = M::mma(wd * ud * (E + p), Bg, y[D + 1]);
                }
            }
        }
        // ---------------- faces (periodic)
#pragma unroll 1
        for (int f = 0; f < 2 * D; ++f) {
            const int d = f >> 1, sd = f & 1, fo = 2 * d + (1 - sd);
            const R sgn = sd ? R(1) : R(-1);
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] = (nidx[d] + (sd ? 1 : -1) + nn[d]) % nn[d];
            const int cn = D == 2 ? nidx[0] * a.n1 + nidx[1] : nidx[0] + nidx[1] * a.n0 + nidx[2] * a.n0 * a.n1;
            const R hd = d == 0 ? h[0] : (d == 1 ? h[1] : h[2]);
            const R hn = (R)a.width[d * a.nmax + nidx[d]];
            const R scn = R(1) / sqrt(detJ / hd * hn);
            const R* __restrict__ Un = a.U + (size_t)cn * NV * Np;
            acc_t Ti[NV], To[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                Ti[v] = z;
                To[v] = z;
#pragma unroll
                for (int s = 0; s < NK; ++s) {
                    const int n = 4 * s + grp;
                    const R bq = n < Np ? Un[v * Np + n] : R(0);
                    Ti[v] = M::mma(tab[L::TA + (f * 16 + col) * NPK + 4 * s + grp], bU[v][s], Ti[v]);
                    To[v] = M::mma(tab[L::TA + (fo * 16 + col) * NPK + 4 * s + grp], bq, To[v]);
                }
            }
            const R wf = detJ / (R(0.5) * hd) * sc;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = M::row(grp, r);                  // face node of register r
                const bool live = valid && q < L::NQF;
                R Ui[NV], Uo[NV], F[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    Ui[v] = live ? Ti[v][r] * sc : (v == 0 ? R(1) : (v == D + 1 ? R(1) : R(0)));
                    Uo[v] = live ? To[v][r] * scn : (v == 0 ? R(1) : (v == D + 1 ? R(1) : R(0)));
                }
                hllc_m<D, R>(a.gamma, a.MS, Ui, Uo, d, sgn, F);
                const R w = live ? -tab[L::FW + q] * wf : R(0);
                const R Bt = tab[L::TB + (f * 16 + q) * 16 + col];
#pragma unroll
                for (int v = 0; v < NV; ++v) y[v] = M::mma(w * F[v], Bt, y[v]);
            }
        }
        // ---------------- output: lane holds Y_v[cell tile*16 + row(grp, r)][mode col]
        if (col < Np) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = tile * 16 + M::row(grp, r);
                if (co < a.J) {
#pragma unroll
                    for (int v = 0; v < NV; ++v) a.out[((size_t)co * NV + v) * Np + col] = y[v][r];
                }
            }
        }
    }
}

template <int D, int K, int NQ1>
static void ensure_tables(EulerOp* e) {
    if (e->tabm64.p) return;
    SipTables t;
    t.build(D, K, NQ1);
    std::vector<double> h;
    build_euler_mfma_tables<D, K, NQ1>(t, h);
    e->tabm64.alloc(h.size() * sizeof(double));
    e->ctx->upload(e->tabm64.p, h.data(), h.size() * sizeof(double));
    std::vector<float> h32(h.begin(), h.end());
    e->tabm32.alloc(h32.size() * sizeof(float));
    e->ctx->upload(e->tabm32.p, h32.data(), h32.size() * sizeof(float));
}

template <int D, int K, int NQ1, typename R>
static void launch_mfma(EulerOp* e, const R* U, R* out) {
    ensure_tables<D, K, NQ1>(e);
    EulerMfmaArgs<R> a;
    a.J = e->g->J;
    a.n0 = e->g->n[0]; a.n1 = e->g->n[1]; a.n2 = e->g->n[2];
    a.nmax = e->g->nmax;
    a.width = e->g->d_width.as<double>();
    a.tab = (const R*)(sizeof(R) == 4 ? e->tabm32.p : e->tabm64.p);
    a.gamma = (R)e->gamma;
    a.MS = (R)e->MS;
    a.U = U;
    a.out = out;
    constexpr int W = 4;
    const int ntiles = (a.J + 15) / 16;
    int blocks = (ntiles + W - 1) / W;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL((euler_rhs_mfma_kernel<D, K, NQ1, R, W>), dim3(blocks), dim3(64 * W), 0, e->ctx->stream, a);
    BH_HIP(hipGetLastError());
}

bool euler_rhs_mfma(EulerOp* e, const void* U, void* out, int prec) {
#define BH_CASE(DD, KK, QQ)                                                                                  \
    if (e->D == DD && e->k == KK && e->nq1 == QQ) {                                                          \
        if (prec == BH_FP32) launch_mfma<DD, KK, QQ, float>(e, (const float*)U, (float*)out);                \
        else launch_mfma<DD, KK, QQ, double>(e, (const double*)U, (double*)out);                             \
        return true;                                                                                         \
    }
    BH_CASE(3, 1, 3) BH_CASE(3, 2, 4) BH_CASE(3, 1, 2) BH_CASE(2, 2, 4) BH_CASE(2, 3, 5) BH_CASE(2, 1, 3)
#undef BH_CASE
    return false;
}

}  // namespace bh

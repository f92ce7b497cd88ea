// FP64 matrix-free SIP apply on the matrix cores (gfx950): the drop-in
// operator (SIPLaplace, L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250,
// quadrature as F/NECQuadratureVolume.cs:377-870, F/NECQuadratureEdge.cs:561-1400)
// with every contraction as v_mfma_f64_16x16x4_f64.  Same results as the
// vector kernel up to summation order.
//
// One wave = 16 consecutive cells.  Lane layouts of the f64 16x16x4 product:
// A[l & 15][l >> 4], B[l >> 4][l & 15], D[(l >> 4) + 4r][l & 15] (r = 0..3).
//   volume          Y[c][m] += sum_n (-nu_c 4/h_d^2 u_c[n]) Kv[d][n][m]  (reference stiffness
//                   per direction: the volume rule of an axis-aligned box factorises)
//   face traces     D[q][c] = sum_n T_f[q][n] u_c[n]   (rows q = face nodes, K = modes in steps of 4)
//   flux to modes   Y[c][m] += sum_q F[c][q] T_f[q][m]  (K = the 4 nodes (l >> 4) + 4r of register r)
// Register r of a trace product is, per lane, the A operand of the flux product
// whose K slots are the nodes {g + 4r}: the pointwise flux is formed in registers.
// FP64 tables (stiffness rows; face trace / normal-derivative rows with the row
// stride padded to an odd number of doubles against LDS bank conflicts) are
// staged once per workgroup; workgroups loop over tiles, XCD by XCD.
#include <cmath>
#include <vector>

#include "bh_internal.hpp"

namespace bh {

typedef double d4_t __attribute__((ext_vector_type(4)));

namespace {
constexpr int bnm(int n, int k) {
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
constexpr int pw(int b, int e) { return e == 0 ? 1 : b * pw(b, e - 1); }
}  // namespace

template <int D, int K>
struct M64Layout {
    static constexpr int Np = bnm(K + D, D);
    static constexpr int NQ1 = K + 1;
    static constexpr int NQV = pw(NQ1, D);
    static constexpr int NQF = pw(NQ1, D - 1);
    static constexpr int NK = (Np + 3) / 4;                 // K steps of the first product
    static constexpr int NPK = NK * 4;                      // modes padded to the K step
    static constexpr int LD = NPK | 1;                      // row stride (doubles): odd, >= Np
    static constexpr int MT = (Np + 15) / 16;               // 16-column output tiles
    static constexpr int LDK = MT * 16;                     // row stride of the stiffness tables
    static constexpr int KV = 0;                            // [D][NPK][LDK] reference stiffness Kv[d][n][m]
    static constexpr int F = KV + D * NPK * LDK;            // [2D][2][16][LD] trace / normal derivative at face node q
    static constexpr int FW = F + 2 * D * 2 * 16 * LD;      // [16] face weights
    static constexpr int TOTAL = FW + 16;
    static_assert(NQF <= 16 && Np <= 32, "FP64 MFMA apply: face nodes <= 16, Np <= 32");
};

template <int D, int K>
static void build_tables64(const SipTables& t, std::vector<double>& h) {
    using L = M64Layout<D, K>;
    h.assign(L::TOTAL, 0.0);
    for (int d = 0; d < D; ++d)
        for (int n = 0; n < L::Np; ++n)
            for (int m = 0; m < L::Np; ++m)
                h[L::KV + ((size_t)d * L::NPK + n) * L::LDK + m] = t.Kv[((size_t)d * L::Np + n) * L::Np + m];
    for (int f = 0; f < 2 * D; ++f)
        for (int q = 0; q < L::NQF; ++q)
            for (int n = 0; n < L::Np; ++n) {
                h[L::F + (((size_t)f * 2 + 0) * 16 + q) * L::LD + n] = t.T[((size_t)f * L::NQF + q) * L::Np + n];
                h[L::F + (((size_t)f * 2 + 1) * 16 + q) * L::LD + n] = t.Nn[((size_t)f * L::NQF + q) * L::Np + n];
            }
    for (int q = 0; q < L::NQF; ++q) h[L::FW + q] = t.fw[q];
}

struct M64Args {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ width;
    const double* __restrict__ nu;
    const double* __restrict__ tab;
    double pen;
    int bc_neumann;
    const double* __restrict__ u;
    double* __restrict__ y;
    const double* __restrict__ affine;
    double alpha, beta;
    int tiles_per_xcd;
};

template <int D>
__device__ __forceinline__ void dec64(int c, int n0, int n1, int* idx) {
    if (D == 2) { idx[0] = c / n1; idx[1] = c - idx[0] * n1; idx[2] = 0; }
    else { int s = n0 * n1; idx[2] = c / s; int r = c - idx[2] * s; idx[1] = r / n0; idx[0] = r - idx[1] * n0; }
}
template <int D>
__device__ __forceinline__ int enc64(const int* idx, int n0, int n1) {
    return D == 2 ? idx[0] * n1 + idx[1] : idx[0] + idx[1] * n0 + idx[2] * n0 * n1;
}

#define BH_MFMA64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)
// 4 blocks of 4x4x4: lane 16k + 4b + i holds A_b[i][k] (the same lanes as the 16x16x4 A
// operand of row 4b + i, slot k), lane 16k + 4b + j holds B_b[k][j], and lane 16g + 4b + j
// receives D_b[g][j].  Used for output modes 16..19 (cells 4b..4b+3 per block): a quarter of
// the work of a 16-column tile that would be three-quarters padding.
#define BH_MFMA4(a, b, c) __builtin_amdgcn_mfma_f64_4x4x4f64((a), (b), (c), 0, 0, 0)

template <int D, int K, int W>
__global__ __launch_bounds__(64 * W) void sip_apply_mfma64_kernel(M64Args a) {
    using L = M64Layout<D, K>;
    constexpr int Np = L::Np, NK = L::NK, LD = L::LD;
    constexpr int TAIL = (Np > 16 && Np <= 20) ? Np - 16 : 0;   // modes 16.. on the 4x4x4 blocks
    constexpr int MTF = TAIL ? 1 : L::MT;                         // full 16-column tiles
    __shared__ double tab[L::TOTAL];
    for (int i = threadIdx.x; i < L::TOTAL; i += blockDim.x) tab[i] = a.tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const int tcol = 16 + (col & 3);           // tail mode of this lane's B operand and output
    const bool tlive = TAIL > 0 && tcol < Np;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int ntiles = (a.J + 15) / 16;
    const int xcd = blockIdx.x & 7, nb_x = gridDim.x >> 3;
    const int t0 = xcd * a.tiles_per_xcd, t1 = min(ntiles, t0 + a.tiles_per_xcd);
    const d4_t z = {0, 0, 0, 0};
    for (int tile = t0 + (blockIdx.x >> 3) * W + wave; tile < t1; tile += nb_x * W) {
        const int c = tile * 16 + col;
        const bool valid = c < a.J;
        const int cc = valid ? c : a.J - 1;
        int idx[3];
        dec64<D>(cc, a.n0, a.n1, idx);
        double h[3] = {1, 1, 1}, detJ = 1;
        for (int d = 0; d < D; ++d) { h[d] = a.width[d * a.nmax + idx[d]]; detJ *= 0.5 * h[d]; }
        const double nuP = a.nu[cc];
        double bU[NK];
#pragma unroll
        for (int s = 0; s < NK; ++s) {
            const int n = 4 * s + grp;
            bU[s] = (valid && n < Np) ? a.u[(size_t)cc * Np + n] : 0.0;
        }
        d4_t yacc[MTF];
#pragma unroll
        for (int mt = 0; mt < MTF; ++mt) yacc[mt] = z;
        double ytail = 0;
        // ---------------- volume: Y += sum_d diag(-nu 4/h_d^2) U Kv[d]  (A = scaled modes, B = Kv[d])
#pragma unroll 1
        for (int d = 0; d < D; ++d) {
            const double hd = d == 0 ? h[0] : (d == 1 ? h[1] : h[2]);
            const double cd = -nuP * 4.0 / (hd * hd);
#pragma unroll
            for (int s = 0; s < NK; ++s) {
                const double av = cd * bU[s];
                const int krow = L::KV + (d * L::NPK + 4 * s + grp) * L::LDK;
#pragma unroll
                for (int mt = 0; mt < MTF; ++mt) yacc[mt] = BH_MFMA64(av, tab[krow + col + 16 * mt], yacc[mt]);
                if (TAIL) ytail = BH_MFMA4(av, tlive ? tab[krow + tcol] : 0.0, ytail);
            }
        }
        // ---------------- faces
        const double sdPinv = 1.0 / sqrt(detJ);
        double cjP = 0;
        for (int d = 0; d < D; ++d) cjP += (1.0 + 0.5 * ((idx[d] == 0) + (idx[d] == nn[d] - 1))) / h[d];
#pragma unroll 1
        for (int f = 0; f < 2 * D; ++f) {
            const int d = f >> 1, s = f & 1;
            const double sgn = s ? 1.0 : -1.0;
            const int fQ = 2 * d + (1 - s);
            const bool isBnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
            const bool skip = isBnd && ((a.bc_neumann >> f) & 1);
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] += isBnd ? 0 : (s ? 1 : -1);
            const int cq = isBnd ? cc : enc64<D>(nidx, a.n0, a.n1);
            const double hq = isBnd ? h[d] : a.width[d * a.nmax + nidx[d]];
            const double nuQ = isBnd ? nuP : a.nu[cq];
            double bQ[NK];
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                const int n = 4 * t + grp;
                bQ[t] = (!isBnd && n < Np) ? a.u[(size_t)cq * Np + n] : 0.0;
            }
            d4_t dT = z, dN = z, eT = z, eN = z;
#pragma unroll
            for (int t = 0; t < NK; ++t) {
                dT = BH_MFMA64(tab[L::F + ((f * 2 + 0) * 16 + col) * LD + 4 * t + grp], bU[t], dT);
                dN = BH_MFMA64(tab[L::F + ((f * 2 + 1) * 16 + col) * LD + 4 * t + grp], bU[t], dN);
                eT = BH_MFMA64(tab[L::F + ((fQ * 2 + 0) * 16 + col) * LD + 4 * t + grp], bQ[t], eT);
                eN = BH_MFMA64(tab[L::F + ((fQ * 2 + 1) * 16 + col) * LD + 4 * t + grp], bQ[t], eN);
            }
            const double gP = 2.0 / h[d], gQ = 2.0 / hq;
            const double sdQinv = sdPinv * sqrt(h[d] / hq);
            const double wf = gP / sdPinv;                 // dA * sdPinv
            const double nuMax = fabs(nuP) > fabs(nuQ) ? nuP : nuQ;
            const double bP = 1.0 + 0.5 * ((idx[d] == 0) + (idx[d] == nn[d] - 1));
            const double bQf = 1.0 + 0.5 * ((nidx[d] == 0) + (nidx[d] == nn[d] - 1));
            const double cjQ = cjP - bP / h[d] + bQf / hq;
            const double mu = isBnd ? a.pen * cjP : a.pen * fmax(cjP, cjQ);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = grp + 4 * r;                 // face node of register r
                const double vP = sdPinv * dT[r];
                const double dP = sgn * gP * sdPinv * dN[r];
                const double vQ = sdQinv * eT[r];
                const double dQ = sgn * gQ * sdQinv * eN[r];
                double cv, cd;
                if (isBnd) {
                    cv = nuP * dP - 2.0 * mu * nuP * vP;
                    cd = nuP * vP;
                } else {
                    const double jump = vP - vQ;
                    cv = 0.5 * (nuP * dP + nuQ * dQ) - mu * nuMax * jump;
                    cd = 0.5 * nuP * jump;
                }
                const double wq = (skip || !valid) ? 0.0 : tab[L::FW + q] * wf;
                const double cvw = wq * cv, cdw = wq * sgn * gP * cd;
                const int trow = L::F + ((f * 2 + 0) * 16 + q) * LD, nrow = L::F + ((f * 2 + 1) * 16 + q) * LD;
#pragma unroll
                for (int mt = 0; mt < MTF; ++mt) {
                    const int m = col + 16 * mt;
                    const double Tv = m < Np ? tab[trow + m] : 0.0;
                    const double Nv = m < Np ? tab[nrow + m] : 0.0;
                    yacc[mt] = BH_MFMA64(cvw, Tv, yacc[mt]);
                    yacc[mt] = BH_MFMA64(cdw, Nv, yacc[mt]);
                }
                if (TAIL) {
                    ytail = BH_MFMA4(cvw, tlive ? tab[trow + tcol] : 0.0, ytail);
                    ytail = BH_MFMA4(cdw, tlive ? tab[nrow + tcol] : 0.0, ytail);
                }
            }
        }
        // ---------------- output: lane holds Y[cell tile*16 + grp + 4r][m = col + 16*mt]
        // and, for the tail, Y[cell tile*16 + 4*(col/4) + grp][m = 16 + col%4]
        if (TAIL && tlive) {
            const int co = tile * 16 + 4 * (col >> 2) + grp;
            if (co < a.J) {
                const size_t o = (size_t)co * Np + tcol;
                double v = ytail;
                if (a.affine) v += a.affine[o];
                a.y[o] = a.beta == 0.0 ? a.alpha * v : a.beta * a.y[o] + a.alpha * v;
            }
        }
#pragma unroll
        for (int mt = 0; mt < MTF; ++mt) {
            const int m = col + 16 * mt;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int co = tile * 16 + grp + 4 * r;
                if (co < a.J && m < Np) {
                    const size_t o = (size_t)co * Np + m;
                    double v = yacc[mt][r];
                    if (a.affine) v += a.affine[o];
                    a.y[o] = a.beta == 0.0 ? a.alpha * v : a.beta * a.y[o] + a.alpha * v;
                }
            }
        }
    }
}

template <int D, int K, int W>
static void m64_run(bh_op* op, const M64Args& a, int per_xcd) {
    hipLaunchKernelGGL((sip_apply_mfma64_kernel<D, K, W>), dim3(8 * per_xcd), dim3(64 * W), 0, op->ctx->stream, a);
    BH_HIP(hipGetLastError());
}

template <int D, int K>
static void m64_launch(bh_op* op, const M64Args& a0, int waves, int per_xcd) {
    if (!op->tab64m.p) {
        std::vector<double> h;
        build_tables64<D, K>(op->tab, h);
        op->tab64m.alloc(h.size() * sizeof(double));
        op->ctx->upload(op->tab64m.p, h.data(), h.size() * sizeof(double));
    }
    M64Args a = a0;
    a.tab = op->tab64m.as<double>();
    const int ntiles = (a.J + 15) / 16;
    a.tiles_per_xcd = (ntiles + 7) / 8;
    const int need = (a.tiles_per_xcd + waves - 1) / waves;
    if (per_xcd > need) per_xcd = need > 0 ? need : 1;
    if (waves == 4) m64_run<D, K, 4>(op, a, per_xcd);
    else m64_run<D, K, 8>(op, a, per_xcd);
}

bool sip_apply_mfma64_supported(int D, int K) { return (D == 3 && (K == 2 || K == 3)) || (D == 2 && (K == 2 || K == 3)); }

void sip_apply_mfma64(bh_op* op, double alpha, const double* u, double beta, double* y, const double* affine, int waves,
                      int per_xcd) {
    bh_grid* g = op->grid;
    M64Args a{};
    a.J = g->J;
    a.n0 = g->n[0]; a.n1 = g->n[1]; a.n2 = g->n[2];
    a.nmax = g->nmax;
    a.width = g->d_width.as<double>();
    a.nu = op->nu.as<double>();
    a.pen = op->penalty_base * op->penalty_deg;
    a.bc_neumann = 0;
    for (int f = 0; f < 2 * op->D; ++f) if (op->bc[f] == 1) a.bc_neumann |= 1 << f;
    a.u = u;
    a.y = y;
    a.affine = affine;
    a.alpha = alpha;
    a.beta = beta;
    const int D = op->D, K = op->k;
    if (D == 3 && K == 3) return m64_launch<3, 3>(op, a, waves, per_xcd);
    if (D == 3 && K == 2) return m64_launch<3, 2>(op, a, waves, per_xcd);
    if (D == 2 && K == 3) return m64_launch<2, 3>(op, a, waves, per_xcd);
    if (D == 2 && K == 2) return m64_launch<2, 2>(op, a, waves, per_xcd);
    throw Error(-3, "sip_apply (FP64 MFMA): unsupported (D, k)");
}

}  // namespace bh

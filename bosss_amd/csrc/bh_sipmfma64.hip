// FP64 matrix-free SIP apply on the matrix cores (gfx950): the drop-in
// operator (SIPLaplace, L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250,
// quadrature as F/NECQuadratureVolume.cs:377-870, F/NECQuadratureEdge.cs:561-1400)
// with every contraction as an FP64 MFMA.  Same results as the vector kernel up
// to summation order.
//
// One wave = 16 consecutive cells.  Lane layouts of the f64 16x16x4 product:
// A[l & 15][l >> 4], B[l >> 4][l & 15], D[(l >> 4) + 4r][l & 15] (r = 0..3).
//   volume          Y[c][m] += sum_n (-nu_c 4/h_d^2 u_c[n]) Kv[d][n][m]  (reference stiffness
//                   per direction: the volume rule of an axis-aligned box factorises)
//   face traces     D[q][c] = sum_n T_f[q][n] u_c[n]   (rows q = face nodes, K = modes in steps of 4)
//   flux to modes   Y[c][m] += sum_q F[c][q] T_f[q][m]  (K = the 4 nodes (l >> 4) + 4r of register r)
// Register r of a trace product is, per lane, the A operand of the flux product
// whose K slots are the nodes {g + 4r}: the pointwise flux is formed in registers.
//
// Work saved against a product per face side:
//  * the K slots of the trace and volume products hold the modes grouped by the
//    parity of their exponents (M64Order).  On the Legendre tensor basis the
//    traces on the two faces of direction d differ only by (-1)^{e_d} (values)
//    and -(-1)^{e_d} (normal derivatives), so one even and one odd product give
//    both faces of a cell; a K step whose modes are all even or all odd in d
//    costs one product instead of two;
//  * along the direction of consecutive cells the neighbour's trace is the
//    adjacent lane's own trace (DPP row shifts within each 16-lane row); only
//    the two cells beyond the tile are traced, on v_mfma_f64_4x4x4 blocks;
//  * widths, their reciprocals and square roots come from a table: no division
//    or square root per face.
// FP64 tables (stiffness rows; face trace / normal-derivative rows with the row
// stride padded to an odd number of doubles against LDS bank conflicts) are
// staged once per workgroup; workgroups loop over tiles, XCD by XCD.
#include <cmath>
#include <utility>
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

// K-slot order of the modes: exponent triples grouped by parity class
// (e_0, e_1, e_2 mod 2), -1 = padding slot.  Groups of four slots are one K step.
template <int D, int K> struct M64Order;
template <> struct M64Order<3, 3> {
    static constexpr int NS = 20;
    static constexpr int x[NS][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},    // e e e
                                     {1, 0, 0}, {3, 0, 0}, {1, 2, 0}, {1, 0, 2},    // o e e
                                     {0, 1, 0}, {0, 3, 0}, {2, 1, 0}, {0, 1, 2},    // e o e
                                     {0, 0, 1}, {0, 0, 3}, {2, 0, 1}, {0, 2, 1},    // e e o
                                     {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};   // singletons
};
template <> struct M64Order<3, 2> {
    static constexpr int NS = 12;
    static constexpr int x[NS][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {0, 0, 2},
                                     {1, 0, 0}, {1, 1, 0}, {1, 0, 1}, {-1, -1, -1},
                                     {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {-1, -1, -1}};
};
template <> struct M64Order<2, 3> {
    static constexpr int NS = 12;
    static constexpr int x[NS][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {-1, -1, -1},
                                     {1, 0, 0}, {3, 0, 0}, {1, 2, 0}, {-1, -1, -1},
                                     {0, 1, 0}, {2, 1, 0}, {0, 3, 0}, {1, 1, 0}};
};
template <> struct M64Order<2, 2> {
    static constexpr int NS = 8;
    static constexpr int x[NS][3] = {{0, 0, 0}, {2, 0, 0}, {0, 2, 0}, {0, 1, 0},
                                     {1, 0, 0}, {1, 1, 0}, {-1, -1, -1}, {-1, -1, -1}};
};

template <int D, int K>
struct M64Slots {
    using O = M64Order<D, K>;
    static constexpr int Np = bnm(K + D, D);
    static constexpr int NK = (Np + 3) / 4;
    static_assert(O::NS == 4 * NK, "one K step per group of four slots");
    // basis index of the mode in a slot (order of basis_exponents: degree, then e_0, then e_1), -1: padding
    static constexpr int mode(int slot) {
        if (O::x[slot][0] < 0) return -1;
        const int a = O::x[slot][0], b = O::x[slot][1], c = D == 3 ? O::x[slot][2] : 0, deg = a + b + c;
        if (D == 2) return deg * (deg + 1) / 2 + a;
        int idx = 0;
        for (int e = 0; e < deg; ++e) idx += (e + 1) * (e + 2) / 2;
        for (int ap = 0; ap < a; ++ap) idx += deg - ap + 1;
        return idx + b;
    }
    static constexpr int par(int slot, int d) { return O::x[slot][0] < 0 ? -1 : (O::x[slot][d] & 1); }
    // K step s in direction d: 0 every mode even, 1 every mode odd, 2 mixed, 3 padding only
    static constexpr int cls(int d, int s) {
        int ev = 0, od = 0;
        for (int g = 0; g < 4; ++g) {
            ev += par(4 * s + g, d) == 0;
            od += par(4 * s + g, d) == 1;
        }
        return ev && od ? 2 : ev ? 0 : od ? 1 : 3;
    }
    static constexpr unsigned oddbits(int d, int s) {   // bit g: mode of slot 4s + g is odd in d
        unsigned m = 0;
        for (int g = 0; g < 4; ++g) m |= (par(4 * s + g, d) == 1 ? 1u : 0u) << g;
        return m;
    }
    static constexpr unsigned packed(int s) {           // byte g: mode of slot 4s + g (255: padding)
        unsigned v = 0;
        for (int g = 0; g < 4; ++g) v |= (unsigned)(mode(4 * s + g) < 0 ? 255 : mode(4 * s + g)) << (8 * g);
        return v;
    }
    static constexpr bool valid() {                     // the slots hold every mode exactly once
        int seen[64] = {};
        for (int s = 0; s < O::NS; ++s) {
            const int m = mode(s);
            if (m >= Np) return false;
            if (m >= 0) ++seen[m];
        }
        for (int m = 0; m < Np; ++m)
            if (seen[m] != 1) return false;
        return true;
    }
};

template <int D, int K>
struct M64Layout {
    static constexpr int Np = bnm(K + D, D);
    static constexpr int NQ1 = K + 1;
    static constexpr int NQF = pw(NQ1, D - 1);
    static constexpr int NK = (Np + 3) / 4;                 // K steps of the trace and volume products
    static constexpr int NPK = NK * 4;                      // K slots
    static constexpr int LD = NPK | 1;                      // row stride (doubles): odd, >= Np
    static constexpr int MT = (Np + 15) / 16;               // 16-column output tiles
    static constexpr int LDK = MT * 16;                     // row stride of the stiffness tables
    static constexpr int KV = 0;                            // [D][slot][LDK] reference stiffness Kv[d][mode(slot)][m]
    static constexpr int FT = KV + D * NPK * LDK;           // [D][2][16][LD] trace / normal derivative of face 2d, by slot
    static constexpr int FO = FT + D * 2 * 16 * LD;         // [2D][2][16][LD] w_q x trace / normal derivative of face f, by mode
    static constexpr int TOTAL = FO + 2 * D * 2 * 16 * LD;  // staged in LDS; the width table follows in memory
    static_assert(NQF <= 16 && Np <= 32, "FP64 MFMA apply: face nodes <= 16, Np <= 32");
    static_assert(M64Slots<D, K>::valid(), "K-slot order is not a permutation of the modes");
};

template <int D, int K>
static void build_tables64(const SipTables& t, const bh_grid* g, std::vector<double>& h) {
    using L = M64Layout<D, K>;
    using S = M64Slots<D, K>;
    const int nmax = g->nmax;
    h.assign(L::TOTAL + 4 * 3 * (size_t)nmax, 0.0);
    for (int d = 0; d < D; ++d)
        for (int sl = 0; sl < L::NPK; ++sl) {
            const int n = S::mode(sl);
            if (n < 0) continue;
            for (int m = 0; m < L::Np; ++m)
                h[L::KV + ((size_t)d * L::NPK + sl) * L::LDK + m] = t.Kv[((size_t)d * L::Np + n) * L::Np + m];
        }
    // the kernel forms the '+' face traces from the '-' face rows: check the sign relations
    std::vector<int> ex;
    basis_exponents(D, K, ex);
    for (int d = 0; d < D; ++d)
        for (int q = 0; q < L::NQF; ++q)
            for (int n = 0; n < L::Np; ++n) {
                const double sg = (ex[(size_t)n * D + d] & 1) ? -1.0 : 1.0;
                const size_t im = ((size_t)(2 * d) * L::NQF + q) * L::Np + n, ip = ((size_t)(2 * d + 1) * L::NQF + q) * L::Np + n;
                const double dt = std::fabs(t.T[ip] - sg * t.T[im]), dn = std::fabs(t.Nn[ip] + sg * t.Nn[im]);
                if (dt > 1e-13 * (1 + std::fabs(t.T[im])) || dn > 1e-13 * (1 + std::fabs(t.Nn[im])))
                    throw Error(-3, "sip_apply (FP64 MFMA): face traces do not have the parity symmetry");
            }
    for (int d = 0; d < D; ++d)
        for (int q = 0; q < L::NQF; ++q)
            for (int sl = 0; sl < L::NPK; ++sl) {
                const int n = S::mode(sl);
                if (n < 0) continue;
                h[L::FT + (((size_t)d * 2 + 0) * 16 + q) * L::LD + sl] = t.T[((size_t)(2 * d) * L::NQF + q) * L::Np + n];
                h[L::FT + (((size_t)d * 2 + 1) * 16 + q) * L::LD + sl] = t.Nn[((size_t)(2 * d) * L::NQF + q) * L::Np + n];
            }
    for (int f = 0; f < 2 * D; ++f)
        for (int q = 0; q < L::NQF; ++q)
            for (int n = 0; n < L::Np; ++n) {
                h[L::FO + (((size_t)f * 2 + 0) * 16 + q) * L::LD + n] = t.fw[q] * t.T[((size_t)f * L::NQF + q) * L::Np + n];
                h[L::FO + (((size_t)f * 2 + 1) * 16 + q) * L::LD + n] = t.fw[q] * t.Nn[((size_t)f * L::NQF + q) * L::Np + n];
            }
    // width table [4][3][nmax]: h, 1/h, sqrt(h), 1/sqrt(h)
    double* wt = h.data() + L::TOTAL;
    for (int d = 0; d < g->D; ++d)
        for (int i = 0; i < g->n[d]; ++i) {
            const double w = g->width[d][i];
            wt[(0 * 3 + d) * nmax + i] = w;
            wt[(1 * 3 + d) * nmax + i] = 1.0 / w;
            wt[(2 * 3 + d) * nmax + i] = std::sqrt(w);
            wt[(3 * 3 + d) * nmax + i] = 1.0 / std::sqrt(w);
        }
}

struct M64Args {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ wt;     // [4][3][nmax]: h, 1/h, sqrt(h), 1/sqrt(h)
    const double* __restrict__ nu;
    const double* __restrict__ tab;
    double pen;
    int bc_neumann;
    const double* __restrict__ u;
    double* __restrict__ y;
    const double* __restrict__ affine;
    double alpha, beta;
    int tiles_per_xcd;
    const double* __restrict__ b0;   // if set: y = b0 + alpha * (A u)
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

// load through a 32-bit byte offset (every array of the operator is below 4 GiB: checked on the host)
__device__ __forceinline__ double ldg(const double* __restrict__ p, int i) {
    return *(const double*)((const char*)p + (unsigned)i * 8u);
}

// Row-local lane move of a double (DPP): lanes whose source lies outside their
// 16-lane row keep `old`.  0x100 + n: lane i reads lane i + n; 0x110 + n: lane i - n.
template <int CTRL>
__device__ __forceinline__ double dpp64(double old, double src) {
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

#define BH_MFMA64(a, b, c) __builtin_amdgcn_mfma_f64_16x16x4f64((a), (b), (c), 0, 0, 0)
// 4 blocks of 4x4x4: lane 16k + 4b + i holds A_b[i][k] (the same lanes as the 16x16x4 A
// operand of row 4b + i, slot k), lane 16k + 4b + j holds B_b[k][j], and lane 16g + 4b + j
// receives D_b[g][j].  Used for output modes 16..19 (cells 4b..4b+3 per block): a quarter of
// the work of a 16-column tile that would be three-quarters padding; and for the traces of
// the two cells beyond a tile (columns j = 0 and 3).
#define BH_MFMA4(a, b, c) __builtin_amdgcn_mfma_f64_4x4x4f64((a), (b), (c), 0, 0, 0)

template <int D, int K, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(W == 8 ? 4 : 2))) void sip_apply_mfma64_kernel(M64Args a) {
    using L = M64Layout<D, K>;
    using S = M64Slots<D, K>;
    constexpr int Np = L::Np, NK = L::NK, LD = L::LD;
    constexpr int TAIL = (Np > 16 && Np <= 20) ? Np - 16 : 0;   // modes 16.. on the 4x4x4 blocks
    constexpr int MTF = TAIL ? 1 : L::MT;                         // full 16-column tiles
    constexpr int DF = D == 2 ? 1 : 0;                            // direction of consecutive cells
    __shared__ double tab[L::TOTAL];
    for (int i = threadIdx.x; i < L::TOTAL; i += blockDim.x) tab[i] = a.tab[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4, jj = col & 3;
    const int tcol = 16 + jj;                  // tail mode of this lane's B operand and output
    const bool tlive = TAIL > 0 && tcol < Np;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int nmax = a.nmax;
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
        // widths on demand from the table (L1-resident): t = 0 h, 1 1/h, 2 sqrt(h), 3 1/sqrt(h)
        auto wtab = [&](int t, int d, int i) { return ldg(a.wt, (t * 3 + d) * nmax + i); };
        double cjP = 0;
        for (int d = 0; d < D; ++d) {
            cjP += (1.0 + 0.5 * ((idx[d] == 0) + (idx[d] == nn[d] - 1))) * wtab(1, d, idx[d]);
        }
        const double nuP = ldg(a.nu, cc);
        double bU[NK];
        static_for<NK>([&](auto SS) {
            constexpr int s = SS;
            const int n = (S::packed(s) >> (8 * grp)) & 255;
            bU[s] = (valid && n < Np) ? ldg(a.u, cc * Np + n) : 0.0;
        });
        d4_t yacc[MTF];
#pragma unroll
        for (int mt = 0; mt < MTF; ++mt) yacc[mt] = z;
        double ytail = 0;
        // ---------------- volume: Y += sum_d diag(-nu 4/h_d^2) U Kv[d]  (A = scaled modes, B = Kv[d])
#pragma unroll 1
        for (int d = 0; d < D; ++d) {
            const double ihd = wtab(1, d, idx[d]);
            const double cd = -nuP * 4.0 * ihd * ihd;
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

        // flux of face f = 2d + s of this lane's cell into Y; tr(r, dT, dN, eT, eN) gives at node
        // grp + 4r the own trace / normal derivative and the neighbour's on the same face
        auto face_flux = [&](const int d, const int s, const int cq, const bool isBnd, const int iq, auto&& tr) {
            const int f = 2 * d + s;
            const double sgn = s ? 1.0 : -1.0;
            const bool skip = isBnd && ((a.bc_neumann >> f) & 1);
            const int ip = idx[d];
            const double ihd = wtab(1, d, ip), shd = wtab(2, d, ip);
            const double ihq = wtab(1, d, iq), ishq = wtab(3, d, iq);   // iq = ip on the boundary
            const double nuQ = isBnd ? nuP : ldg(a.nu, cq);
            const double gP = 2.0 * ihd, gQ = 2.0 * ihq;
            const double nuMax = fabs(nuP) > fabs(nuQ) ? nuP : nuQ;
            const int nd = nn[d];
            const double bP = 1.0 + 0.5 * ((ip == 0) + (ip == nd - 1));
            const double bQf = 1.0 + 0.5 * ((iq == 0) + (iq == nd - 1));
            const double cjQ = cjP - bP * ihd + bQf * ihq;
            const double mu = isBnd ? a.pen * cjP : a.pen * fmax(cjP, cjQ);
            // The flux is linear in the four traces at a node (per side, with sdinv = 1/sqrt(det J)
            // of the side's cell: v = sdinv T, d = sgn g sdinv N):
            //   interior  cv = nuP/2 dP + nuQ/2 dQ - mu nuMax (vP - vQ),  cd = nuP/2 (vP - vQ)
            //   boundary  cv = nuP dP - 2 mu nuP vP,                      cd = nuP vP
            // weighted by dA sdPinv (face weights are in the flux rows):
            //   cvw = c1 dN + c2 eN + c3 dT + c4 eT,  cdw = sgn gP cd = c1 dT + c5 eT
            const double G = (skip || !valid) ? 0.0 : gP;      // dA * sdPinv * sdPinv = 2 / h_d
            const double rq = shd * ishq;                       // sdQinv / sdPinv
            const double A = G * nuP * sgn * gP;
            const double c1 = (isBnd ? 1.0 : 0.5) * A;
            const double c3 = -G * mu * (isBnd ? 2.0 * nuP : nuMax);
            const double c2 = isBnd ? 0.0 : 0.5 * G * nuQ * sgn * gQ * rq;
            const double c4 = isBnd ? 0.0 : -c3 * rq;
            const double c5 = isBnd ? 0.0 : -0.5 * A * rq;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int q = grp + 4 * r;                 // face node of register r
                double dT, dN, eT, eN;
                tr(r, dT, dN, eT, eN);
                const double cvw = c1 * dN + c2 * eN + c3 * dT + c4 * eT;
                const double cdw = c1 * dT + c5 * eT;
                const int trow = L::FO + ((f * 2 + 0) * 16 + q) * LD, nrow = L::FO + ((f * 2 + 1) * 16 + q) * LD;
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
        };

        static_for<D>([&](auto DD) {
            constexpr int d = DD;
            __builtin_amdgcn_sched_barrier(0);      // one direction at a time: bounds the live traces
            const int ip = idx[d];
            const bool bndM = ip == 0, bndP = ip == nn[d] - 1;
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] = bndM ? ip : ip - 1;
            const int cm = bndM ? cc : enc64<D>(nidx, a.n0, a.n1), im = nidx[d];
            nidx[d] = bndP ? ip : ip + 1;
            const int cp = bndP ? cc : enc64<D>(nidx, a.n0, a.n1), iq = nidx[d];
            // own traces on faces 2d and 2d+1 from their even and odd parts in d:
            // face 2d: T = tE + tO, N = nE + nO;  face 2d+1: T = tE - tO, N = nO - nE
            d4_t tE = z, tO = z, nE = z, nO = z;
            double hT = 0, hN = 0;   // d == DF: traces of the cells beyond the tile (4x4x4 columns 0 and 3)
            {
                double hb[NK];
                if constexpr (d == DF) {
                    const int cell = jj == 0 ? max(tile * 16 - 1, 0) : min(tile * 16 + 16, a.J - 1);
                    static_for<NK>([&](auto SS) {
                        constexpr int s = SS;
                        const int n = (S::packed(s) >> (8 * grp)) & 255;
                        const bool odd = (S::oddbits(d, s) >> grp) & 1;
                        const double v = ((jj == 0 || jj == 3) && n < Np) ? ldg(a.u, cell * Np + n) : 0.0;
                        hb[s] = (jj == 0 && odd) ? -v : v;   // column 0: the previous cell's '+' face via face 2d's rows
                    });
                }
                static_for<NK>([&](auto SS) {
                    constexpr int s = SS;
                    constexpr int cl = S::cls(d, s);
                    const double aT = tab[L::FT + ((d * 2 + 0) * 16 + col) * LD + 4 * s + grp];
                    const double aN = tab[L::FT + ((d * 2 + 1) * 16 + col) * LD + 4 * s + grp];
                    if constexpr (cl == 0) {
                        tE = BH_MFMA64(aT, bU[s], tE);
                        nE = BH_MFMA64(aN, bU[s], nE);
                    } else if constexpr (cl == 1) {
                        tO = BH_MFMA64(aT, bU[s], tO);
                        nO = BH_MFMA64(aN, bU[s], nO);
                    } else if constexpr (cl == 2) {
                        const bool odd = (S::oddbits(d, s) >> grp) & 1;
                        const double be = odd ? 0.0 : bU[s], bo = odd ? bU[s] : 0.0;
                        tE = BH_MFMA64(aT, be, tE);
                        nE = BH_MFMA64(aN, be, nE);
                        tO = BH_MFMA64(aT, bo, tO);
                        nO = BH_MFMA64(aN, bo, nO);
                    }
                    if constexpr (d == DF) {
                        hT = BH_MFMA4(aT, hb[s], hT);
                        hN = BH_MFMA4(aN, hb[s], hN);
                    }
                });
            }
            // neighbour traces of a non-DF direction: '-' side neighbour through its face 2d+1
            // (odd modes sign-flipped through face 2d's rows), '+' side through its face 2d
            auto qtrace = [&](const int cn, const bool bnd, const bool flip, d4_t& qT, d4_t& qN) {
                double bq[NK];
                static_for<NK>([&](auto SS) {
                    constexpr int s = SS;
                    const int n = (S::packed(s) >> (8 * grp)) & 255;
                    const bool odd = flip && ((S::oddbits(d, s) >> grp) & 1);
                    const double v = (!bnd && n < Np) ? ldg(a.u, cn * Np + n) : 0.0;
                    bq[s] = odd ? -v : v;
                });
                qT = z; qN = z;
                static_for<NK>([&](auto SS) {
                    constexpr int s = SS;
                    qT = BH_MFMA64(tab[L::FT + ((d * 2 + 0) * 16 + col) * LD + 4 * s + grp], bq[s], qT);
                    qN = BH_MFMA64(tab[L::FT + ((d * 2 + 1) * 16 + col) * LD + 4 * s + grp], bq[s], qN);
                });
            };
            if constexpr (d == DF) {
                // lane (grp, col) of the 4x4x4 result holds node 4*(col>>2) + grp of column col & 3:
                // node grp + 4r of the previous cell sits in lane col 4r, of the next cell in col 4r + 3.
                // Neighbour lanes: col - 1 ('+' face) for face 2d, col + 1 ('-' face) for face 2d+1.
                face_flux(d, 0, cm, bndM, im, [&](int r, double& dT, double& dN, double& eT, double& eN) {
                    dT = tE[r] + tO[r];
                    dN = nE[r] + nO[r];
                    const double hTr = r == 0 ? hT : r == 1 ? dpp64<0x104>(hT, hT) : r == 2 ? dpp64<0x108>(hT, hT) : dpp64<0x10C>(hT, hT);
                    const double hNr = r == 0 ? hN : r == 1 ? dpp64<0x104>(hN, hN) : r == 2 ? dpp64<0x108>(hN, hN) : dpp64<0x10C>(hN, hN);
                    eT = dpp64<0x111>(hTr, tE[r] - tO[r]);
                    eN = dpp64<0x111>(-hNr, nO[r] - nE[r]);
                });
                __builtin_amdgcn_sched_barrier(0);
                face_flux(d, 1, cp, bndP, iq, [&](int r, double& dT, double& dN, double& eT, double& eN) {
                    dT = tE[r] - tO[r];
                    dN = nO[r] - nE[r];
                    const double hTr = r == 3 ? hT : r == 2 ? dpp64<0x114>(hT, hT) : r == 1 ? dpp64<0x118>(hT, hT) : dpp64<0x11C>(hT, hT);
                    const double hNr = r == 3 ? hN : r == 2 ? dpp64<0x114>(hN, hN) : r == 1 ? dpp64<0x118>(hN, hN) : dpp64<0x11C>(hN, hN);
                    eT = dpp64<0x101>(hTr, tE[r] + tO[r]);
                    eN = dpp64<0x101>(hNr, nE[r] + nO[r]);
                });
            } else {
                d4_t qT, qN;
                qtrace(cm, bndM, true, qT, qN);
                face_flux(d, 0, cm, bndM, im, [&](int r, double& dT, double& dN, double& eT, double& eN) {
                    dT = tE[r] + tO[r];
                    dN = nE[r] + nO[r];
                    eT = qT[r];
                    eN = -qN[r];
                });
                __builtin_amdgcn_sched_barrier(0);
                qtrace(cp, bndP, false, qT, qN);
                face_flux(d, 1, cp, bndP, iq, [&](int r, double& dT, double& dN, double& eT, double& eN) {
                    dT = tE[r] - tO[r];
                    dN = nO[r] - nE[r];
                    eT = qT[r];
                    eN = qN[r];
                });
            }
        });
        // ---------------- output: lane holds Y[cell tile*16 + grp + 4r][m = col + 16*mt]
        // and, for the tail, Y[cell tile*16 + 4*(col/4) + grp][m = 16 + col%4]
        if (TAIL && tlive) {
            const int co = tile * 16 + 4 * (col >> 2) + grp;
            if (co < a.J) {
                const size_t o = (size_t)co * Np + tcol;
                double v = ytail;
                if (a.affine) v += a.affine[o];
                a.y[o] = a.b0 ? a.b0[o] + a.alpha * v : (a.beta == 0.0 ? a.alpha * v : a.beta * a.y[o] + a.alpha * v);
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
                    a.y[o] = a.b0 ? a.b0[o] + a.alpha * v : (a.beta == 0.0 ? a.alpha * v : a.beta * a.y[o] + a.alpha * v);
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
    using L = M64Layout<D, K>;
    if (!op->tab64m.p) {
        std::vector<double> h;
        build_tables64<D, K>(op->tab, op->grid, h);
        op->tab64m.alloc(h.size() * sizeof(double));
        op->ctx->upload(op->tab64m.p, h.data(), h.size() * sizeof(double));
    }
    M64Args a = a0;
    a.tab = op->tab64m.as<double>();
    a.wt = a.tab + L::TOTAL;
    const int ntiles = (a.J + 15) / 16;
    a.tiles_per_xcd = (ntiles + 7) / 8;
    const int need = (a.tiles_per_xcd + waves - 1) / waves;
    if (per_xcd > need) per_xcd = need > 0 ? need : 1;
    if (waves == 4) m64_run<D, K, 4>(op, a, per_xcd);
    else m64_run<D, K, 8>(op, a, per_xcd);
}

bool sip_apply_mfma64_supported(int D, int K, int64_t N) {
    return ((D == 3 && (K == 2 || K == 3)) || (D == 2 && (K == 2 || K == 3))) && N * 8 < ((int64_t)1 << 32);
}

void sip_apply_mfma64(bh_op* op, double alpha, const double* u, double beta, double* y, const double* affine, int waves,
                      int per_xcd, const double* b0) {
    bh_grid* g = op->grid;
    M64Args a{};
    a.J = g->J;
    a.n0 = g->n[0]; a.n1 = g->n[1]; a.n2 = g->n[2];
    a.nmax = g->nmax;
    a.nu = op->nu.as<double>();
    a.pen = op->penalty_base * op->penalty_deg;
    a.bc_neumann = 0;
    for (int f = 0; f < 2 * op->D; ++f) if (op->bc[f] == 1) a.bc_neumann |= 1 << f;
    a.u = u;
    a.y = y;
    a.affine = affine;
    a.alpha = alpha;
    a.beta = beta;
    a.b0 = b0;
    const int D = op->D, K = op->k;
    if (D == 3 && K == 3) return m64_launch<3, 3>(op, a, waves, per_xcd);
    if (D == 3 && K == 2) return m64_launch<3, 2>(op, a, waves, per_xcd);
    if (D == 2 && K == 3) return m64_launch<2, 3>(op, a, waves, per_xcd);
    if (D == 2 && K == 2) return m64_launch<2, 2>(op, a, waves, per_xcd);
    throw Error(-3, "sip_apply (FP64 MFMA): unsupported (D, k)");
}

}  // namespace bh

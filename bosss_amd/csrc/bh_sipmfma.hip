// Matrix-free SIP apply on the matrix cores (gfx950): the same operator as
// sip_apply_kernel (SIPLaplace volume + interior/boundary face terms,
// L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250, quadrature as in
// F/NECQuadratureVolume.cs:377-870 and F/NECQuadratureEdge.cs:561-1400), with
// FP32 vectors and the basis-at-node / flux-to-modes contractions as FP16
// MFMA (v_mfma_f32_16x16x32_f16) with FP32 accumulation.
//
// One wave = a tile of 16 consecutive cells.  Every contraction is a small
// GEMM with the cells as one dimension:
//   basis at nodes:  D_t[j][c] = sum_n A_t[j][n] * u_c[n]     (A = table rows j, K = modes padded to 32)
//   flux to modes:   Y[c][m] += sum_j F[c][j] * B[j][m]        (K = 32 rows j: two 16-row tiles)
// The accumulator layout of the first product (lane = cell c, 4 rows j) is
// the A-operand layout of the second (lane = row c, 8 consecutive k), so the
// pointwise flux (scaling by nu, 4/h^2, weights; the SIP face flux from the
// four traces of one face node, which sit in the same lane and register) is
// computed in registers between the two MFMAs: no LDS round trip.
// FP16 range: each cell's modes are scaled by their max |u| before the cast,
// and each cell's flux vector by its max |flux| before the second product;
// the scales are applied to the FP32 results.
// The tables (FP16, ~48 KB for 3D k = 3) are staged once per workgroup in LDS;
// workgroups loop over tiles.
#include <cmath>
#include <cstring>
#include <vector>

#include "bh_internal.hpp"

namespace bh {

typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef float f4_t __attribute__((ext_vector_type(4)));

namespace {
constexpr int binom_(int n, int k) {
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
constexpr int ipow_(int b, int e) { return e == 0 ? 1 : b * ipow_(b, e - 1); }
}  // namespace

// Table layout (FP16, offsets in halves) and the FP32 side tables.
template <int D, int K>
struct MfmaLayout {
    static constexpr int Np = binom_(K + D, D);
    static constexpr int NQ1 = K + 1;
    static constexpr int NQV = ipow_(NQ1, D);
    static constexpr int NQF = ipow_(NQ1, D - 1);
    static constexpr int NVT = ((D * NQV + 31) / 32) * 2;   // 16-row volume tiles, even count
    static constexpr int NPAIR = NVT / 2;
    static constexpr int MT = (Np + 15) / 16;               // 16-column output tiles
    static constexpr int GA = 0;                            // [NVT*16][32]   volume rows j = q*D + d, cols n
    static constexpr int GB = GA + NVT * 16 * 32;           // [NPAIR][MT*16][32] (m, k-slot)
    static constexpr int FA = GB + NPAIR * MT * 16 * 32;    // [2D][2][16][32] face trace / normal-derivative rows
    static constexpr int FB = FA + 2 * D * 2 * 16 * 32;     // [2D][MT*16][32]
    static constexpr int TOTAL = FB + 2 * D * MT * 16 * 32;
    static constexpr int NW = NVT * 16 + 16;                // FP32: volume row weights, face node weights
    static_assert(Np <= 32, "MFMA apply: Np <= 32");
    static_assert(NQF <= 16, "MFMA apply: face nodes <= 16");
    static_assert(TOTAL % 8 == 0, "16-byte staging");
};

// k-slot s (0..7) of lane group g in a 32-deep K block built from two 16-row tiles
__host__ __device__ constexpr int slot_row(int tile_pair, int g, int s) {
    return s < 4 ? (2 * tile_pair) * 16 + 4 * g + s : (2 * tile_pair + 1) * 16 + 4 * g + (s - 4);
}

template <int D, int K>
void build_mfma_tables(const SipTables& t, std::vector<uint16_t>& h16, std::vector<float>& w32, std::vector<int>& d32) {
    using L = MfmaLayout<D, K>;
    std::vector<float> f(L::TOTAL, 0.0f);
    const int Np = L::Np, NQV = L::NQV, NQF = L::NQF;
    auto G = [&](int j, int n) -> double {   // volume row j = q*D + d
        if (j >= D * NQV || n >= Np) return 0.0;
        const int q = j / D, d = j % D;
        return t.G[((size_t)q * D + d) * Np + n];
    };
    for (int j = 0; j < L::NVT * 16; ++j)
        for (int n = 0; n < 32; ++n) f[L::GA + j * 32 + n] = (float)G(j, n);
    for (int p = 0; p < L::NPAIR; ++p)
        for (int m = 0; m < L::MT * 16; ++m)
            for (int g = 0; g < 4; ++g)
                for (int s = 0; s < 8; ++s)
                    f[L::GB + ((size_t)p * L::MT * 16 + m) * 32 + 8 * g + s] = (float)G(slot_row(p, g, s), m);
    auto T = [&](int fc, int q, int n) -> double { return (q < NQF && n < Np) ? t.T[((size_t)fc * NQF + q) * Np + n] : 0.0; };
    auto N = [&](int fc, int q, int n) -> double { return (q < NQF && n < Np) ? t.Nn[((size_t)fc * NQF + q) * Np + n] : 0.0; };
    for (int fc = 0; fc < 2 * D; ++fc) {
        for (int q = 0; q < 16; ++q)
            for (int n = 0; n < 32; ++n) {
                f[L::FA + (((size_t)fc * 2 + 0) * 16 + q) * 32 + n] = (float)T(fc, q, n);
                f[L::FA + (((size_t)fc * 2 + 1) * 16 + q) * 32 + n] = (float)N(fc, q, n);
            }
        for (int m = 0; m < L::MT * 16; ++m)
            for (int g = 0; g < 4; ++g)
                for (int s = 0; s < 8; ++s) {
                    const int q = 4 * g + (s & 3);
                    f[L::FB + ((size_t)fc * L::MT * 16 + m) * 32 + 8 * g + s] = (float)(s < 4 ? T(fc, q, m) : N(fc, q, m));
                }
    }
    h16.resize(L::TOTAL);
    for (int i = 0; i < L::TOTAL; ++i) {
        _Float16 hv = (_Float16)f[i];
        uint16_t b;
        std::memcpy(&b, &hv, 2);
        h16[i] = b;
    }
    w32.assign(L::NW, 0.0f);
    d32.assign(L::NVT * 16, 0);
    for (int j = 0; j < D * NQV; ++j) { w32[j] = (float)t.vw[j / D]; d32[j] = j % D; }
    for (int q = 0; q < NQF; ++q) w32[L::NVT * 16 + q] = (float)t.fw[q];
}

struct MfmaArgs {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ width;
    const float* __restrict__ nu;
    const _Float16* __restrict__ tab;
    const float* __restrict__ w;     // volume row weights | face node weights
    const int* __restrict__ dr;      // volume row direction
    float pen;
    int bc_neumann;
    const float* __restrict__ u;
    float* __restrict__ y;
    const double* __restrict__ affine;
    float alpha, beta;
};

__device__ __forceinline__ h8_t lds_h8(const _Float16* p) { return *(const h8_t*)p; }

__device__ __forceinline__ float lane_max4(float v) {
    v = fmaxf(v, __shfl_xor(v, 16));
    v = fmaxf(v, __shfl_xor(v, 32));
    return v;
}

template <int D>
__device__ __forceinline__ void decode_cell(int c, int n0, int n1, int* idx) {
    if (D == 2) { idx[0] = c / n1; idx[1] = c - idx[0] * n1; idx[2] = 0; }
    else { int s = n0 * n1; idx[2] = c / s; int r = c - idx[2] * s; idx[1] = r / n0; idx[0] = r - idx[1] * n0; }
}
template <int D>
__device__ __forceinline__ int encode_cell(const int* idx, int n0, int n1) {
    return D == 2 ? idx[0] * n1 + idx[1] : idx[0] + idx[1] * n0 + idx[2] * n0 * n1;
}

template <int D>
__device__ __forceinline__ float cj_of(const int* idx, const double* h, const int* nn) {
    double vol = 1, surf = 0, bnd = 0;
    for (int d = 0; d < D; ++d) vol *= h[d];
    for (int d = 0; d < D; ++d) {
        const double ar = vol / h[d];
        surf += 2 * ar;
        if (idx[d] == 0) bnd += ar;
        if (idx[d] == nn[d] - 1) bnd += ar;
    }
    return (float)(0.5 * (surf + bnd) / vol);
}

constexpr int MFMA_WAVES = 8;

template <int D, int K>
__global__ __launch_bounds__(64 * MFMA_WAVES) void sip_apply_mfma_kernel(MfmaArgs a) {
    using L = MfmaLayout<D, K>;
    constexpr int Np = L::Np;
    __shared__ __attribute__((aligned(16))) _Float16 tab[L::TOTAL];
    __shared__ float wsh[L::NW];
    __shared__ int dsh[L::NVT * 16];
    for (int i = threadIdx.x; i < L::TOTAL / 8; i += blockDim.x) ((uint4*)tab)[i] = ((const uint4*)a.tab)[i];
    for (int i = threadIdx.x; i < L::NW; i += blockDim.x) wsh[i] = a.w[i];
    for (int i = threadIdx.x; i < L::NVT * 16; i += blockDim.x) dsh[i] = a.dr[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int ntiles = (a.J + 15) / 16;
    for (int tile = blockIdx.x * MFMA_WAVES + wave; tile < ntiles; tile += gridDim.x * MFMA_WAVES) {
        // ---------------- this lane's cell (B-operand / flux phases)
        const int c = tile * 16 + col;
        const bool valid = c < a.J;
        const int cc = valid ? c : a.J - 1;
        int idx[3];
        decode_cell<D>(cc, a.n0, a.n1, idx);
        double h[3] = {1, 1, 1}, detJ = 1;
        for (int d = 0; d < D; ++d) { h[d] = a.width[d * a.nmax + idx[d]]; detJ *= 0.5 * h[d]; }
        const float nuP = a.nu[cc];
        // own modes 8*grp .. 8*grp+7, scaled to FP16
        float uf[8];
        float mx = 0.f;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const int n = 8 * grp + s;
            uf[s] = (valid && n < Np) ? a.u[(size_t)cc * Np + n] : 0.f;
            mx = fmaxf(mx, fabsf(uf[s]));
        }
        mx = lane_max4(mx);
        const float usP = mx > 0.f ? mx : 1.f;
        h8_t bU;
#pragma unroll
        for (int s = 0; s < 8; ++s) bU[s] = (_Float16)(uf[s] / usP);
        f4_t yacc[L::MT];
#pragma unroll
        for (int mt = 0; mt < L::MT; ++mt) yacc[mt] = f4_t{0.f, 0.f, 0.f, 0.f};
        // ---------------- volume: -nu * sum_q w_q (4/h_d^2) dpsi_m/dxi_d dpsi_n/dxi_d u_n
        {
            float s2[3] = {0.f, 0.f, 0.f}, s2max = 0.f;
            for (int d = 0; d < D; ++d) { s2[d] = (float)(4.0 / (h[d] * h[d])); s2max = fmaxf(s2max, s2[d]); }
            const float rs0 = -s2[0] / s2max, rs1 = -s2[1] / s2max, rs2 = D == 3 ? -s2[2] / s2max : 0.f;
            const float kap = usP * nuP * s2max;
            f4_t yv[L::MT];
#pragma unroll
            for (int mt = 0; mt < L::MT; ++mt) yv[mt] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int p = 0; p < L::NPAIR; ++p) {
                const f4_t z = {0.f, 0.f, 0.f, 0.f};
                const h8_t A0 = lds_h8(&tab[L::GA + ((2 * p) * 16 + col) * 32 + 8 * grp]);
                const h8_t A1 = lds_h8(&tab[L::GA + ((2 * p + 1) * 16 + col) * 32 + 8 * grp]);
                const f4_t d0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A0, bU, z, 0, 0, 0);
                const f4_t d1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1, bU, z, 0, 0, 0);
                h8_t F;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int j0 = (2 * p) * 16 + 4 * grp + i, j1 = j0 + 16;
                    const int e0 = dsh[j0], e1 = dsh[j1];
                    const float r0 = e0 == 0 ? rs0 : (e0 == 1 ? rs1 : rs2);
                    const float r1 = e1 == 0 ? rs0 : (e1 == 1 ? rs1 : rs2);
                    F[i] = (_Float16)(r0 * wsh[j0] * d0[i]);
                    F[4 + i] = (_Float16)(r1 * wsh[j1] * d1[i]);
                }
#pragma unroll
                for (int mt = 0; mt < L::MT; ++mt) {
                    const h8_t Bm = lds_h8(&tab[L::GB + ((size_t)p * L::MT * 16 + mt * 16 + col) * 32 + 8 * grp]);
                    yv[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(F, Bm, yv[mt], 0, 0, 0);
                }
            }
            float kr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) kr[i] = __shfl(kap, 4 * grp + i);
#pragma unroll
            for (int mt = 0; mt < L::MT; ++mt)
#pragma unroll
                for (int i = 0; i < 4; ++i) yacc[mt][i] += kr[i] * yv[mt][i];
        }
        // ---------------- faces
        const float cjP = cj_of<D>(idx, h, nn);
        const float sdPinv = (float)(1.0 / sqrt(detJ));
#pragma unroll 1
        for (int f = 0; f < 2 * D; ++f) {
            const int d = f >> 1, s = f & 1;
            const float sgn = s ? 1.f : -1.f;
            const int fQ = 2 * d + (1 - s);
            const bool isBnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
            const bool skip = isBnd && ((a.bc_neumann >> f) & 1);
            int cq = cc;
            double hq[3] = {h[0], h[1], h[2]};
            int nidx[3] = {idx[0], idx[1], idx[2]};
            if (!isBnd) {
                nidx[d] += s ? 1 : -1;
                cq = encode_cell<D>(nidx, a.n0, a.n1);
                hq[d] = a.width[d * a.nmax + nidx[d]];
            }
            float uq[8];
            float mq = 0.f;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int n = 8 * grp + t;
                uq[t] = (!isBnd && n < Np) ? a.u[(size_t)cq * Np + n] : 0.f;
                mq = fmaxf(mq, fabsf(uq[t]));
            }
            mq = lane_max4(mq);
            const float usQ = mq > 0.f ? mq : 1.f;
            h8_t bQ;
#pragma unroll
            for (int t = 0; t < 8; ++t) bQ[t] = (_Float16)(uq[t] / usQ);
            const f4_t z = {0.f, 0.f, 0.f, 0.f};
            const f4_t dT = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((f * 2 + 0) * 16 + col) * 32 + 8 * grp]), bU, z, 0, 0, 0);
            const f4_t dN = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((f * 2 + 1) * 16 + col) * 32 + 8 * grp]), bU, z, 0, 0, 0);
            const f4_t eT = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((fQ * 2 + 0) * 16 + col) * 32 + 8 * grp]), bQ, z, 0, 0, 0);
            const f4_t eN = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((fQ * 2 + 1) * 16 + col) * 32 + 8 * grp]), bQ, z, 0, 0, 0);
            // face scalars of this lane's cell
            double detQ = 1;
            for (int e = 0; e < D; ++e) detQ *= 0.5 * hq[e];
            const float sdQinv = (float)(1.0 / sqrt(detQ));
            const float gP = (float)(2.0 / h[d]), gQ = (float)(2.0 / hq[d]);
            const float dA = (float)(detJ / (0.5 * h[d]));
            const float nuQ = isBnd ? nuP : a.nu[cq];
            const float nuMax = fabsf(nuP) > fabsf(nuQ) ? nuP : nuQ;
            const float mu = isBnd ? a.pen * cjP : a.pen * fmaxf(cjP, cj_of<D>(nidx, hq, nn));
            float cvw[4], cdw[4], sm = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int q = 4 * grp + i;
                const float wq = wsh[L::NVT * 16 + q] * dA * sdPinv;   // zero for padded nodes
                const float vP = sdPinv * usP * dT[i];
                const float dP = sgn * gP * sdPinv * usP * dN[i];
                const float vQ = sdQinv * usQ * eT[i];
                const float dQ = sgn * gQ * sdQinv * usQ * eN[i];
                float cv, cd;
                if (isBnd) {
                    cv = nuP * dP - 2.f * mu * nuP * vP;
                    cd = nuP * vP;
                } else {
                    const float jump = vP - vQ;
                    cv = 0.5f * (nuP * dP + nuQ * dQ) - mu * nuMax * jump;
                    cd = 0.5f * nuP * jump;
                }
                cvw[i] = (skip || !valid) ? 0.f : wq * cv;
                cdw[i] = (skip || !valid) ? 0.f : wq * sgn * gP * cd;
                sm = fmaxf(sm, fmaxf(fabsf(cvw[i]), fabsf(cdw[i])));
            }
            sm = lane_max4(sm);
            const float sig = sm > 0.f ? sm : 1.f;
            h8_t F;
#pragma unroll
            for (int i = 0; i < 4; ++i) { F[i] = (_Float16)(cvw[i] / sig); F[4 + i] = (_Float16)(cdw[i] / sig); }
            float sr[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) sr[i] = __shfl(sig, 4 * grp + i);
#pragma unroll
            for (int mt = 0; mt < L::MT; ++mt) {
                const h8_t Bm = lds_h8(&tab[L::FB + ((size_t)f * L::MT * 16 + mt * 16 + col) * 32 + 8 * grp]);
                const f4_t yf = __builtin_amdgcn_mfma_f32_16x16x32_f16(F, Bm, z, 0, 0, 0);
#pragma unroll
                for (int i = 0; i < 4; ++i) yacc[mt][i] += sr[i] * yf[i];
            }
        }
        // ---------------- output: lane holds Y[cell tile*16 + 4*grp + i][m = col + 16*mt]
#pragma unroll
        for (int mt = 0; mt < L::MT; ++mt) {
            const int m = col + 16 * mt;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int co = tile * 16 + 4 * grp + i;
                if (co < a.J && m < Np) {
                    const size_t o = (size_t)co * Np + m;
                    float v = yacc[mt][i];
                    if (a.affine) v += (float)a.affine[o];
                    a.y[o] = a.beta == 0.f ? a.alpha * v : a.beta * a.y[o] + a.alpha * v;
                }
            }
        }
    }
}

template <int D, int K>
static void mfma_launch(bh_op* op, const MfmaArgs& a0) {
    using L = MfmaLayout<D, K>;
    if (!op->tab16.p) {
        std::vector<uint16_t> h16;
        std::vector<float> w;
        std::vector<int> dr;
        build_mfma_tables<D, K>(op->tab, h16, w, dr);
        op->tab16.alloc(h16.size() * 2 + w.size() * 4 + dr.size() * 4 + 64);
        char* base = (char*)op->tab16.p;
        op->ctx->upload(base, h16.data(), h16.size() * 2);
        op->ctx->upload(base + h16.size() * 2, w.data(), w.size() * 4);
        op->ctx->upload(base + h16.size() * 2 + w.size() * 4, dr.data(), dr.size() * 4);
    }
    MfmaArgs a = a0;
    const char* base = (const char*)op->tab16.p;
    a.tab = (const _Float16*)base;
    a.w = (const float*)(base + (size_t)L::TOTAL * 2);
    a.dr = (const int*)(base + (size_t)L::TOTAL * 2 + (size_t)L::NW * 4);
    const int ntiles = (a.J + 15) / 16;
    int blocks = (ntiles + MFMA_WAVES - 1) / MFMA_WAVES;
    if (blocks > 1024) blocks = 1024;   // workgroups loop over tiles; tables staged once per workgroup
    hipLaunchKernelGGL((sip_apply_mfma_kernel<D, K>), dim3(blocks), dim3(64 * MFMA_WAVES), 0, op->ctx->stream, a);
    BH_HIP(hipGetLastError());
}

void sip_apply_mfma(bh_op* op, double alpha, const float* u, double beta, float* y, bool add_affine) {
    if (add_affine && op->has_affine && !op->affine_valid) op->rebuild_affine();
    bh_grid* g = op->grid;
    MfmaArgs a{};
    a.J = g->J;
    a.n0 = g->n[0]; a.n1 = g->n[1]; a.n2 = g->n[2];
    a.nmax = g->nmax;
    a.width = g->d_width.as<double>();
    a.nu = op->nu32.as<float>();
    a.pen = (float)(op->penalty_base * op->penalty_deg);
    a.bc_neumann = 0;
    for (int f = 0; f < 2 * op->D; ++f) if (op->bc[f] == 1) a.bc_neumann |= 1 << f;
    a.u = u;
    a.y = y;
    a.affine = (add_affine && op->has_affine) ? op->affine.as<double>() : nullptr;
    a.alpha = (float)alpha;
    a.beta = (float)beta;
    const int D = op->D, K = op->k;
    if (D == 3 && K == 3) return mfma_launch<3, 3>(op, a);
    if (D == 3 && K == 2) return mfma_launch<3, 2>(op, a);
    if (D == 3 && K == 1) return mfma_launch<3, 1>(op, a);
    if (D == 2 && K == 3) return mfma_launch<2, 3>(op, a);
    if (D == 2 && K == 2) return mfma_launch<2, 2>(op, a);
    throw Error(-3, "sip_apply (FP16 MFMA): unsupported (D, k) = (" + std::to_string(D) + ", " + std::to_string(K) + ")");
}

}  // namespace bh

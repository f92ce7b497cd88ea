// Matrix-free SIP apply on the matrix cores (gfx950): the same operator as
// sip_apply_kernel (SIPLaplace volume + interior/boundary face terms,
// L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250, quadrature as in
// F/NECQuadratureVolume.cs:377-870 and F/NECQuadratureEdge.cs:561-1400), with
// FP32 vectors and every contraction as FP16 MFMA (v_mfma_f32_16x16x32_f16)
// with FP32 accumulation.
//
// One wave = a tile of 16 consecutive cells.  Every contraction is a small
// GEMM with the cells as one dimension:
//   volume:          Y[c][m] += sum_n (s_d(c) u_c[n]) Kv[d][n][m]   (reference stiffness per direction)
//   face traces:     D[q][c] = sum_n T_f[q][n] * u_c[n]           (rows q = face nodes, K = modes padded to 32)
//   flux to modes:   Y[c][m] += sum_k F[c][k] * B[k][m]           (K = 16 nodes x {trace, normal derivative})
// The accumulator layout of a trace product (lane = cell c, 4 nodes) is the
// A-operand layout of the flux product (lane = row c, 8 consecutive k), so the
// SIP face flux (from the four traces of one face node, which sit in the same
// lane and register) is computed in registers between the two MFMAs: no LDS
// round trip.
// FP16 range: each cell's modes are scaled by their max |u| before the cast,
// and each cell's flux vector by its max |flux| before the second product;
// the scales are applied to the FP32 results.
// The tables (FP16, ~30 KB for 3D k = 3) are staged once per workgroup in LDS;
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

// Table layout (FP16, offsets in halves): the reference stiffness per direction
// and the face tables; the face back-contraction rows carry the face weight
// (FB = w_q * trace).
template <int D, int K>
struct MfmaLayout {
    static constexpr int Np = binom_(K + D, D);
    static constexpr int NQ1 = K + 1;
    static constexpr int NQV = ipow_(NQ1, D);
    static constexpr int NQF = ipow_(NQ1, D - 1);
    static constexpr int MT = (Np + 15) / 16;               // 16-column output tiles
    static constexpr int KB = 0;                            // [D][MT*16][32] reference stiffness Kv[d] (m, k = n)
    static constexpr int FA = KB + D * MT * 16 * 32;        // [2D][2][16][32] face trace / normal-derivative rows
    static constexpr int FB = FA + 2 * D * 2 * 16 * 32;     // [2D][MT*16][32]
    static constexpr int TOTAL = FB + 2 * D * MT * 16 * 32;
    static_assert(Np <= 32, "MFMA apply: Np <= 32");
    static_assert(NQF <= 16, "MFMA apply: face nodes <= 16");
    static_assert(TOTAL % 8 == 0, "16-byte staging");
};

template <int D, int K>
void build_mfma_tables(const SipTables& t, std::vector<uint16_t>& h16) {
    using L = MfmaLayout<D, K>;
    std::vector<float> f(L::TOTAL, 0.0f);
    const int Np = L::Np, NQF = L::NQF;
    for (int d = 0; d < D; ++d)
        for (int m = 0; m < Np; ++m)
            for (int n = 0; n < Np; ++n) f[L::KB + ((size_t)d * L::MT * 16 + m) * 32 + n] = (float)t.Kv[((size_t)d * Np + n) * Np + m];
    auto T = [&](int fc, int q, int n) -> double { return (q < NQF && n < Np) ? t.T[((size_t)fc * NQF + q) * Np + n] : 0.0; };
    auto N = [&](int fc, int q, int n) -> double { return (q < NQF && n < Np) ? t.Nn[((size_t)fc * NQF + q) * Np + n] : 0.0; };
    auto FW = [&](int q) -> double { return q < NQF ? t.fw[q] : 0.0; };
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
                    f[L::FB + ((size_t)fc * L::MT * 16 + m) * 32 + 8 * g + s] =
                        (float)(FW(q) * (s < 4 ? T(fc, q, m) : N(fc, q, m)));
                }
    }
    h16.resize(L::TOTAL);
    for (int i = 0; i < L::TOTAL; ++i) {
        _Float16 hv = (_Float16)f[i];
        uint16_t b;
        std::memcpy(&b, &hv, 2);
        h16[i] = b;
    }
}

struct MfmaArgs {
    int J, n0, n1, n2, nmax;
    const double* __restrict__ width;
    const float* __restrict__ nu;
    const _Float16* __restrict__ tab;
    float pen;
    int bc_neumann;
    const float* __restrict__ u;
    float* __restrict__ y;
    const double* __restrict__ affine;
    float alpha, beta;
    int tiles_per_xcd;   // tiles are split into 8 contiguous ranges, one per XCD
    const float* __restrict__ b0;   // if set: y = b0 + alpha * (A u)
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

template <int D, int K, int MFMA_WAVES, bool PF>
__global__ __launch_bounds__(64 * MFMA_WAVES) void sip_apply_mfma_kernel(MfmaArgs a) {
    using L = MfmaLayout<D, K>;
    constexpr int Np = L::Np;
    __shared__ __attribute__((aligned(16))) _Float16 tab[L::TOTAL];
    for (int i = threadIdx.x; i < L::TOTAL / 8; i += blockDim.x) ((uint4*)tab)[i] = ((const uint4*)a.tab)[i];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, grp = lane >> 4;
    const int nn[3] = {a.n0, a.n1, a.n2};
    const int ntiles = (a.J + 15) / 16;
    // XCD-aware order: workgroup b runs on XCD b % 8; XCD x walks its own contiguous
    // range of tiles, so the y/z neighbours a wave reads sit in that XCD's L2
    const int xcd = blockIdx.x & 7, nb_x = gridDim.x >> 3;
    const int t0 = xcd * a.tiles_per_xcd, t1 = min(ntiles, t0 + a.tiles_per_xcd);
    for (int tile = t0 + (blockIdx.x >> 3) * MFMA_WAVES + wave; tile < t1; tile += nb_x * MFMA_WAVES) {
        // ---------------- this lane's cell (B-operand / flux phases)
        const int c = tile * 16 + col;
        const bool valid = c < a.J;
        const int cc = valid ? c : a.J - 1;
        int idx[3];
        decode_cell<D>(cc, a.n0, a.n1, idx);
        // widths from the grid table [4][3][nmax]: h, 1/h, sqrt(h), 1/sqrt(h)
        float ih[3] = {1.f, 1.f, 1.f};
        float cjP = 0.f;
        for (int d = 0; d < D; ++d) {
            ih[d] = (float)a.width[(3 + d) * a.nmax + idx[d]];
            cjP += (1.f + 0.5f * (float)((idx[d] == 0) + (idx[d] == nn[d] - 1))) * ih[d];
        }
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
        const float usP = mx > 0.f ? mx : 1.f, rusP = 1.f / usP;
        h8_t bU;
#pragma unroll
        for (int s = 0; s < 8; ++s) bU[s] = (_Float16)(uf[s] * rusP);
        f4_t yacc[L::MT];
#pragma unroll
        for (int mt = 0; mt < L::MT; ++mt) yacc[mt] = f4_t{0.f, 0.f, 0.f, 0.f};
        // ---------------- volume: Y += sum_d diag(-nu 4/h_d^2) U Kv[d] (reference stiffness per
        // direction: the volume rule of an axis-aligned box factorises); the FP16 A operand is
        // the cell's scaled modes times -s2_d / max_d s2_d, the rest of the factor is applied
        // to the FP32 result rows
        {
            float s2max = 0.f;
            for (int d = 0; d < D; ++d) s2max = fmaxf(s2max, 4.f * ih[d] * ih[d]);
            const float kap = usP * nuP * s2max;
            f4_t yv[L::MT];
#pragma unroll
            for (int mt = 0; mt < L::MT; ++mt) yv[mt] = f4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const float rs = -4.f * ih[d] * ih[d] / s2max;
                h8_t Ad;
#pragma unroll
                for (int s = 0; s < 8; ++s) Ad[s] = (_Float16)(rs * (float)bU[s]);
#pragma unroll
                for (int mt = 0; mt < L::MT; ++mt) {
                    const h8_t Bm = lds_h8(&tab[L::KB + ((size_t)d * L::MT * 16 + mt * 16 + col) * 32 + 8 * grp]);
                    yv[mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(Ad, Bm, yv[mt], 0, 0, 0);
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
        // ---------------- faces.  PF: all six neighbour loads are issued first so their
        // latencies overlap (more registers); otherwise each face loads when it starts.
        constexpr int NPF = PF ? 2 * D : 1;
        float uqr[NPF][8], nuQa[NPF], ihqa[NPF], usQa[NPF];
        h8_t bQa[NPF];
        auto load_face = [&](int f, float* uq, float& nuq, float& ihq) {
            const int d = f >> 1, s = f & 1;
            const bool isBnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] += isBnd ? 0 : (s ? 1 : -1);
            const int cq = isBnd ? cc : encode_cell<D>(nidx, a.n0, a.n1);
            ihq = (float)a.width[(3 + d) * a.nmax + nidx[d]];
            nuq = isBnd ? nuP : a.nu[cq];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int n = 8 * grp + t;
                uq[t] = (!isBnd && n < Np) ? a.u[(size_t)cq * Np + n] : 0.f;
            }
        };
        auto to_h8 = [&](const float* uq, h8_t& bq, float& us) {
            float mq = 0.f;
#pragma unroll
            for (int t = 0; t < 8; ++t) mq = fmaxf(mq, fabsf(uq[t]));
            mq = lane_max4(mq);
            us = mq > 0.f ? mq : 1.f;
            const float r = 1.f / us;
#pragma unroll
            for (int t = 0; t < 8; ++t) bq[t] = (_Float16)(uq[t] * r);
        };
        if constexpr (PF) {
#pragma unroll
            for (int f = 0; f < 2 * D; ++f) load_face(f, uqr[f], nuQa[f], ihqa[f]);
#pragma unroll
            for (int f = 0; f < 2 * D; ++f) to_h8(uqr[f], bQa[f], usQa[f]);
        }
#pragma unroll (PF ? 2 * D : 1)
        for (int f = 0; f < 2 * D; ++f) {
            const int fi = PF ? f : 0;
            if constexpr (PF) {
                __builtin_amdgcn_sched_barrier(0);   // one face at a time: bounds the register pressure
            } else {
                load_face(f, uqr[0], nuQa[0], ihqa[0]);
                to_h8(uqr[0], bQa[0], usQa[0]);
            }
            const int d = f >> 1, s = f & 1;
            const float sgn = s ? 1.f : -1.f;
            const int fQ = 2 * d + (1 - s);
            const bool isBnd = s ? (idx[d] == nn[d] - 1) : (idx[d] == 0);
            const bool skip = isBnd && ((a.bc_neumann >> f) & 1);
            int nidx[3] = {idx[0], idx[1], idx[2]};
            nidx[d] += isBnd ? 0 : (s ? 1 : -1);
            const float ihq = ihqa[fi];
            const float usQ = usQa[fi];
            const h8_t bQ = bQa[fi];
            const f4_t z = {0.f, 0.f, 0.f, 0.f};
            const f4_t dT = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((f * 2 + 0) * 16 + col) * 32 + 8 * grp]), bU, z, 0, 0, 0);
            const f4_t dN = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((f * 2 + 1) * 16 + col) * 32 + 8 * grp]), bU, z, 0, 0, 0);
            const f4_t eT = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((fQ * 2 + 0) * 16 + col) * 32 + 8 * grp]), bQ, z, 0, 0, 0);
            const f4_t eN = __builtin_amdgcn_mfma_f32_16x16x32_f16(lds_h8(&tab[L::FA + ((fQ * 2 + 1) * 16 + col) * 32 + 8 * grp]), bQ, z, 0, 0, 0);
            // face scalars of this lane's cell (float; only direction d differs for Q)
            const float gP = 2.f * ih[d], gQ = 2.f * ihq;
            const float rq = (float)(a.width[(6 + d) * a.nmax + idx[d]] * a.width[(9 + d) * a.nmax + nidx[d]]);   // sdQinv / sdPinv
            const float nuQ = nuQa[fi];
            const float nuMax = fabsf(nuP) > fabsf(nuQ) ? nuP : nuQ;
            const float bP = 1.f + 0.5f * (float)((idx[d] == 0) + (idx[d] == nn[d] - 1));
            const float bQf = 1.f + 0.5f * (float)((nidx[d] == 0) + (nidx[d] == nn[d] - 1));
            const float cjQ = cjP - bP * ih[d] + bQf * ihq;
            const float mu = isBnd ? a.pen * cjP : a.pen * fmaxf(cjP, cjQ);
            // the SIP flux as linear combinations of the FP16-scaled traces (scales usP, usQ and
            // the per-side 1/sqrt(det J), normal sign and 2/h folded into the coefficients;
            // dA sdPinv^2 = 2/h_d):
            //   cvw = c1 dN + c2 eN + c3 dT + c4 eT,  cdw = c1 dT + c5 eT
            const float G = (skip || !valid) ? 0.f : gP;
            const float A = G * nuP * sgn * gP;
            const float c3u = -G * mu * (isBnd ? 2.f * nuP : nuMax);
            const float c1 = (isBnd ? 1.f : 0.5f) * A * usP;
            const float c3 = c3u * usP;
            const float c2 = isBnd ? 0.f : 0.5f * G * nuQ * sgn * gQ * rq * usQ;
            const float c4 = isBnd ? 0.f : -c3u * rq * usQ;
            const float c5 = isBnd ? 0.f : -0.5f * A * rq * usQ;
            float cvw[4], cdw[4], sm = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                cvw[i] = c1 * dN[i] + c2 * eN[i] + c3 * dT[i] + c4 * eT[i];
                cdw[i] = c1 * dT[i] + c5 * eT[i];
                sm = fmaxf(sm, fmaxf(fabsf(cvw[i]), fabsf(cdw[i])));
            }
            sm = lane_max4(sm);
            const float sig = sm > 0.f ? sm : 1.f, rsig = 1.f / sig;
            h8_t F;
#pragma unroll
            for (int i = 0; i < 4; ++i) { F[i] = (_Float16)(cvw[i] * rsig); F[4 + i] = (_Float16)(cdw[i] * rsig); }
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
                    a.y[o] = a.b0 ? a.b0[o] + a.alpha * v : (a.beta == 0.f ? a.alpha * v : a.beta * a.y[o] + a.alpha * v);
                }
            }
        }
    }
}

// Launch shape: waves per workgroup (4 or 8) and workgroups per XCD.  Picked
// once per operator by timing the candidates on the operator's own grid
// (autotuner); the choice is kept in the operator.
template <int D, int K, int W, bool PF>
static void mfma_run(bh_op* op, const MfmaArgs& a, int per_xcd) {
    hipLaunchKernelGGL((sip_apply_mfma_kernel<D, K, W, PF>), dim3(8 * per_xcd), dim3(64 * W), 0, op->ctx->stream, a);
    BH_HIP(hipGetLastError());
}

// cfg: waves per workgroup (4 / 8) + 100 if the neighbour loads are prefetched
template <int D, int K>
static void mfma_run_cfg(bh_op* op, const MfmaArgs& a, int cfg, int per_xcd) {
    switch (cfg) {
        case 4: mfma_run<D, K, 4, false>(op, a, per_xcd); break;
        case 8: mfma_run<D, K, 8, false>(op, a, per_xcd); break;
        case 104: mfma_run<D, K, 4, true>(op, a, per_xcd); break;
        default: mfma_run<D, K, 8, true>(op, a, per_xcd); break;
    }
}

template <int D, int K>
static void mfma_launch(bh_op* op, const MfmaArgs& a0) {
    if (!op->tab16.p) {
        std::vector<uint16_t> h16;
        build_mfma_tables<D, K>(op->tab, h16);
        op->tab16.alloc(h16.size() * 2);
        op->ctx->upload(op->tab16.p, h16.data(), h16.size() * 2);
    }
    MfmaArgs a = a0;
    a.tab = (const _Float16*)op->tab16.p;
    const int ntiles = (a.J + 15) / 16;
    a.tiles_per_xcd = (ntiles + 7) / 8;
    if (op->mfma_waves == 0) {
        // autotune on scratch output (beta = 0, no affine): candidates (waves, workgroups per XCD)
        DevBuf ys((size_t)a.J * basis_count(D, K) * sizeof(float));
        MfmaArgs t = a;
        t.y = ys.as<float>();
        t.beta = 0.f;
        t.b0 = nullptr;
        t.affine = nullptr;
        float best = 1e30f;
        int bw = 8, bx = 64;
        hipEvent_t e0, e1;
        BH_HIP(hipEventCreate(&e0));
        BH_HIP(hipEventCreate(&e1));
        for (int cfg : {4, 8, 104, 108})
            for (int px : {32, 64, 128, 256}) {
                const int waves = cfg % 100;
                const int need = (a.tiles_per_xcd + waves - 1) / waves;   // no more workgroups than tiles
                const int pxe = px < need ? px : (need > 0 ? need : 1);
                if (pxe != px && px != 32) continue;
                mfma_run_cfg<D, K>(op, t, cfg, pxe);
                BH_HIP(hipEventRecord(e0, op->ctx->stream));
                for (int r = 0; r < 3; ++r) mfma_run_cfg<D, K>(op, t, cfg, pxe);
                BH_HIP(hipEventRecord(e1, op->ctx->stream));
                BH_HIP(hipEventSynchronize(e1));
                float ms = 0;
                BH_HIP(hipEventElapsedTime(&ms, e0, e1));
                if (ms < best) { best = ms; bw = cfg; bx = pxe; }
            }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        op->mfma_waves = bw;
        op->mfma_wg_per_xcd = bx;
        op->mfma_tuned_ms = best / 3;
    }
    mfma_run_cfg<D, K>(op, a, op->mfma_waves, op->mfma_wg_per_xcd);
}

void sip_apply_mfma(bh_op* op, double alpha, const float* u, double beta, float* y, bool add_affine, const float* b0) {
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
    a.b0 = b0;
    const int D = op->D, K = op->k;
    if (D == 3 && K == 3) return mfma_launch<3, 3>(op, a);
    if (D == 3 && K == 2) return mfma_launch<3, 2>(op, a);
    if (D == 3 && K == 1) return mfma_launch<3, 1>(op, a);
    if (D == 2 && K == 3) return mfma_launch<2, 3>(op, a);
    if (D == 2 && K == 2) return mfma_launch<2, 2>(op, a);
    throw Error(-3, "sip_apply (FP16 MFMA): unsupported (D, k) = (" + std::to_string(D) + ", " + std::to_string(K) + ")");
}

}  // namespace bh

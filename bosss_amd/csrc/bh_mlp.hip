// Learned per-cell coarse correction: a 3-layer MLP (IN -> H -> H -> OUT,
// ReLU) evaluated per cell, BF16 weights with FP32 accumulation on MFMA
// (v_mfma_f32_16x16x16bf16_1k) for inference; FP32 forward/backward and Adam
// for training, all on the device.  No reference counterpart (the reference
// has no learned components); it is an extra correction inside the
// preconditioner, guarded by residual minimisation with fallback to the plain
// multigrid (the acceptance rule of OrthonormalizationMultigrid.AddSol /
// MinimizeResidual, AS/OrthonormalizationMultigrid.cs:149-215).
//
// Feature map per cell j over its face adjacency (k-th mode residual r_jk,
// mean residual mode r_n0 of the 2D face neighbours n, diffusion nu):
//   s_j = sqrt(||r_j||^2 + sum_n r_n0^2) + tiny,
//   x_j = [ r_j / s_j , r_n0 / s_j for the 2D face neighbours (boundary: 0) ,
//           log10(nu_n / nu_j) for the same neighbours (boundary: +2) , 0-padding ]
// (Np + 4D inputs).  Output y_j maps to the correction e_j = s_j * (h_j^2 / nu_j) * y_j.
#include <cmath>
#include <mutex>
#include <string>

#include "bh_internal.hpp"

#include "bh_mlp.hpp"

namespace bh {

// ------------------------------------------------------------------ bf16
__host__ __device__ inline uint16_t f32_to_bf16(float f) {
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);                                    // round to nearest even
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

// ------------------------------------------------------------------ features
// X[j][:] as described above (fp32, row stride IN_P), scale s[j].
template <typename T, int D>
__global__ __launch_bounds__(256) void mlp_features_k(int J, int Np, int n0, int n1, int n2, const T* __restrict__ r,
                                                      const double* __restrict__ nu, int in_p, float* __restrict__ X,
                                                      float* __restrict__ s) {
    // one segment of in_p lanes (16 or 32) per cell: lane t of the segment owns
    // feature slot t (coalesced reads of the residual row and writes of the
    // feature row); the patch norm is a butterfly inside the segment
    const int segs = 256 / in_p;
    const int t = threadIdx.x & (in_p - 1);
    const int nn[3] = {n0, n1, n2};
    for (int j = blockIdx.x * segs + threadIdx.x / in_p; j < J; j += gridDim.x * segs) {
        int q[3];
        if (D == 2) { q[0] = j / n1; q[1] = j - q[0] * n1; q[2] = 0; }
        else { const int sl = n0 * n1; q[2] = j / sl; const int rr = j - q[2] * sl; q[1] = rr / n0; q[0] = rr - q[1] * n0; }
        // the neighbour across face f = 2d + side of slots Np + f and Np + 2D + f (-1: boundary)
        int nbc = -1;
        if (t >= Np && t < Np + 4 * D) {
            const int f = (t - Np) % (2 * D), d = f >> 1, side = f & 1;
            int qq[3] = {q[0], q[1], q[2]};
            qq[d] += side ? 1 : -1;
            if (qq[d] >= 0 && qq[d] < nn[d]) nbc = D == 2 ? qq[0] * n1 + qq[1] : qq[0] + qq[1] * n0 + qq[2] * n0 * n1;
        }
        double v = 0.0;
        if (t < Np) v = (double)r[(size_t)j * Np + t];
        else if (t < Np + 2 * D && nbc >= 0) v = (double)r[(size_t)nbc * Np];
        double sq = v * v;
        for (int off = in_p >> 1; off > 0; off >>= 1) sq += __shfl_xor(sq, off, 64);
        const double sc = sqrt(sq) + 1e-30;
        float out = 0.f;
        if (t < Np + 2 * D) {
            out = (float)(v / sc);
        } else if (t < Np + 4 * D) {
            out = nbc < 0 ? 2.0f : log10f((float)nu[nbc]) - log10f((float)nu[j]);
        }
        X[(size_t)j * in_p + t] = out;
        if (t == 0) s[j] = (float)sc;
    }
}

// correction e_j += w * s_j * (h_j^2/nu_j) * y_j  (h^2 = product of widths^(2/D))
template <typename T, int D>
__global__ void mlp_apply_scale_k(int J, int Np, int n0, int n1, const double* __restrict__ width, int nmax,
                                  const double* __restrict__ nu, const float* __restrict__ s, const float* __restrict__ Y,
                                  int out_p, T* __restrict__ e) {
    const int64_t tot = (int64_t)J * Np;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / Np), m = (int)(t - (int64_t)j * Np);
        int idx[3];
        if (D == 2) { idx[0] = j / n1; idx[1] = j - idx[0] * n1; idx[2] = 0; }
        else { int sl = n0 * n1; idx[2] = j / sl; int rr = j - idx[2] * sl; idx[1] = rr / n0; idx[0] = rr - idx[1] * n0; }
        double vol = 1;
        for (int d = 0; d < D; ++d) vol *= width[d * nmax + idx[d]];
        const double h2 = pow(vol, 2.0 / D);
        e[t] = (T)((double)s[j] * h2 / nu[j] * (double)Y[(size_t)j * out_p + m]);
    }
}

// targets T[j][m] = e_jm / (s_j h_j^2 / nu_j)
template <int D>
__global__ void mlp_targets_k(int J, int Np, int n0, int n1, const double* __restrict__ width, int nmax,
                              const double* __restrict__ nu, const float* __restrict__ s, const double* __restrict__ e,
                              int out_p, float* __restrict__ Tg) {
    const int64_t tot = (int64_t)J * out_p;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / out_p), m = (int)(t - (int64_t)j * out_p);
        if (m >= Np) { Tg[t] = 0.f; continue; }
        int idx[3];
        if (D == 2) { idx[0] = j / n1; idx[1] = j - idx[0] * n1; idx[2] = 0; }
        else { int sl = n0 * n1; idx[2] = j / sl; int rr = j - idx[2] * sl; idx[1] = rr / n0; idx[0] = rr - idx[1] * n0; }
        double vol = 1;
        for (int d = 0; d < D; ++d) vol *= width[d * nmax + idx[d]];
        const double h2 = pow(vol, 2.0 / D);
        Tg[t] = (float)(e[(size_t)j * Np + m] / ((double)s[j] * h2 / nu[j]));
    }
}

// upstream gradient of a loss L(e) through the correction map e_j = sigma_j y_j
// (sigma_j = s_j h_j^2 / nu_j): dY[j][m] = c * sigma_j * g[j][m], with g = dL/de / c
template <int D>
__global__ void mlp_upstream_k(int J, int Np, int n0, int n1, const double* __restrict__ width, int nmax,
                               const double* __restrict__ nu, const float* __restrict__ s, const double* __restrict__ g,
                               double c, int out_p, float* __restrict__ dY) {
    const int64_t tot = (int64_t)J * out_p;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t / out_p), m = (int)(t - (int64_t)j * out_p);
        if (m >= Np) { dY[t] = 0.f; continue; }
        int idx[3];
        if (D == 2) { idx[0] = j / n1; idx[1] = j - idx[0] * n1; idx[2] = 0; }
        else { int sl = n0 * n1; idx[2] = j / sl; int rr = j - idx[2] * sl; idx[1] = rr / n0; idx[0] = rr - idx[1] * n0; }
        double vol = 1;
        for (int d = 0; d < D; ++d) vol *= width[d * nmax + idx[d]];
        const double h2 = pow(vol, 2.0 / D);
        dY[t] = (float)(c * (double)s[j] * h2 / nu[j] * g[(size_t)j * Np + m]);
    }
}

// ------------------------------------------------------------------ inference (MFMA)
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// One wave = 16 cells.  Operand layouts of v_mfma_f32_16x16x16bf16_1k:
//   A (16 x 16): lane l holds A[l%16][4*(l/16) + i], i = 0..3
//   B (16 x 16): lane l holds B[4*(l/16) + i][l%16]
//   C (16 x 16): lane l holds C[4*(l/16) + i][l%16]
// Weights are staged in LDS transposed (Wt[n][k], bf16) so a B fragment is
// one 8-byte read; activations go through a per-wave LDS tile (bf16, row-major
// [cell][feature]) so an A fragment is one 8-byte read.
template <int INP, int H, int OUTP>
__global__ __launch_bounds__(512) void mlp_infer_k(int J, const float* __restrict__ X, const uint16_t* __restrict__ W1t,
                                                   const float* __restrict__ b1, const uint16_t* __restrict__ W2t,
                                                   const float* __restrict__ b2, const uint16_t* __restrict__ W3t,
                                                   const float* __restrict__ b3, float* __restrict__ Y) {
    constexpr int WAVES = 8;   // the weights (48 KB) are staged once per workgroup of 8 waves
    __shared__ uint16_t sW1[H * INP];
    __shared__ uint16_t sW2[H * H];
    __shared__ uint16_t sW3[OUTP * H];
    __shared__ float sb[2 * H + OUTP];
    __shared__ uint16_t act[WAVES][16 * H];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < H * INP; i += blockDim.x) sW1[i] = W1t[i];
    for (int i = tid; i < H * H; i += blockDim.x) sW2[i] = W2t[i];
    for (int i = tid; i < OUTP * H; i += blockDim.x) sW3[i] = W3t[i];
    for (int i = tid; i < H; i += blockDim.x) { sb[i] = b1[i]; sb[H + i] = b2[i]; }
    for (int i = tid; i < OUTP; i += blockDim.x) sb[2 * H + i] = b3[i];
    __syncthreads();
    uint16_t* A = act[wave];
    const int row = lane & 15, kq = lane >> 4;
    const int ngroups = (J + 15) / 16;
    for (int grp = blockIdx.x * WAVES + wave; grp < ngroups; grp += gridDim.x * WAVES) {
        const int c0 = grp * 16;
        // stage inputs (bf16) into A tile [16][INP]
        for (int t = lane; t < 16 * INP; t += 64) {
            const int rr = t / INP, cc = t - rr * INP;
            const int c = c0 + rr;
            A[rr * H + cc] = f32_to_bf16(c < J ? X[(size_t)c * INP + cc] : 0.f);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // layer 1: [16 x INP] x [INP x H]
        float4_t acc[H / 16];
#pragma unroll
        for (int nt = 0; nt < H / 16; ++nt) acc[nt] = float4_t{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < INP / 16; ++ks) {
            const short4_t a = *(const short4_t*)&A[row * H + ks * 16 + 4 * kq];
#pragma unroll
            for (int nt = 0; nt < H / 16; ++nt) {
                const short4_t b = *(const short4_t*)&sW1[(nt * 16 + row) * INP + ks * 16 + 4 * kq];
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
            }
        }
        __builtin_amdgcn_wave_barrier();
        // bias + ReLU, back to the A tile as bf16 (C layout -> row-major)
#pragma unroll
        for (int nt = 0; nt < H / 16; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * kq + i, n = nt * 16 + row;
                A[m * H + n] = f32_to_bf16(fmaxf(acc[nt][i] + sb[n], 0.f));
            }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // layer 2: [16 x H] x [H x H]
#pragma unroll
        for (int nt = 0; nt < H / 16; ++nt) acc[nt] = float4_t{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) {
            const short4_t a = *(const short4_t*)&A[row * H + ks * 16 + 4 * kq];
#pragma unroll
            for (int nt = 0; nt < H / 16; ++nt) {
                const short4_t b = *(const short4_t*)&sW2[(nt * 16 + row) * H + ks * 16 + 4 * kq];
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
            }
        }
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int nt = 0; nt < H / 16; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * kq + i, n = nt * 16 + row;
                A[m * H + n] = f32_to_bf16(fmaxf(acc[nt][i] + sb[H + n], 0.f));
            }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        // layer 3: [16 x H] x [H x OUTP]
        float4_t o[OUTP / 16];
#pragma unroll
        for (int nt = 0; nt < OUTP / 16; ++nt) o[nt] = float4_t{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < H / 16; ++ks) {
            const short4_t a = *(const short4_t*)&A[row * H + ks * 16 + 4 * kq];
#pragma unroll
            for (int nt = 0; nt < OUTP / 16; ++nt) {
                const short4_t b = *(const short4_t*)&sW3[(nt * 16 + row) * H + ks * 16 + 4 * kq];
                o[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, o[nt], 0, 0, 0);
            }
        }
#pragma unroll
        for (int nt = 0; nt < OUTP / 16; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * kq + i, n = nt * 16 + row;
                const int c = c0 + m;
                if (c < J) Y[(size_t)c * OUTP + n] = o[nt][i] + sb[2 * H + n];
            }
        __builtin_amdgcn_wave_barrier();
    }
}

// ------------------------------------------------------------------ training (FP32)
// Forward + backward for a batch of cells, one workgroup of H threads per
// chunk of cells; thread u owns hidden unit u.  Weight gradients are
// accumulated per workgroup in registers/LDS-free loops over owned entries and
// written as per-block partials (reduced by mlp_reduce_k), so there are no
// atomics.  Weights used in the forward are the BF16-rounded values (as in
// inference); the activations stay FP32.
// UPSTREAM: Tg holds dLoss/dy per cell (the loss is defined outside the network,
// e.g. the residual after the correction); the kernel back-propagates it and
// reports no loss.  Otherwise Tg holds targets of the mean-squared error.
template <int INP, int H, int OUTP, bool UPSTREAM>
__global__ __launch_bounds__(H) void mlp_train_k(int B, const int* __restrict__ idx, const float* __restrict__ X,
                                                 const float* __restrict__ Tg, int nout, const float* __restrict__ W1,
                                                 const float* __restrict__ b1, const float* __restrict__ W2,
                                                 const float* __restrict__ b2, const float* __restrict__ W3,
                                                 const float* __restrict__ b3, float* __restrict__ part, float* __restrict__ lossp) {
    // W1: [H][INP], W2: [H][H], W3: [OUTP][H] (row = output unit), bf16-rounded on use
    __shared__ float x[INP], h1[H], h2[H], y[OUTP], dy[OUTP], dh2[H], dh1[H];
    const int u = threadIdx.x;
    constexpr int NW = H * INP + H + H * H + H + OUTP * H + OUTP;
    // per-thread gradient accumulators for owned entries: W1 row u, W2 row u, W3 column u, biases
    float gW1[INP], gW2[H], gW3[OUTP];
    float gb1 = 0, gb2 = 0, gb3 = 0;
#pragma unroll
    for (int i = 0; i < INP; ++i) gW1[i] = 0;
    for (int i = 0; i < H; ++i) gW2[i] = 0;
#pragma unroll
    for (int i = 0; i < OUTP; ++i) gW3[i] = 0;
    float lacc = 0;
    const int per = (B + gridDim.x - 1) / gridDim.x;
    const int s0 = blockIdx.x * per, s1 = min(B, s0 + per);
    for (int sidx = s0; sidx < s1; ++sidx) {
        const int c = idx ? idx[sidx] : sidx;
        for (int i = u; i < INP; i += H) x[i] = X[(size_t)c * INP + i];
        __syncthreads();
        float a = b1[u];
        for (int i = 0; i < INP; ++i) a += bf16_to_f32(f32_to_bf16(W1[u * INP + i])) * x[i];
        h1[u] = fmaxf(a, 0.f);
        __syncthreads();
        a = b2[u];
        for (int i = 0; i < H; ++i) a += bf16_to_f32(f32_to_bf16(W2[u * H + i])) * h1[i];
        h2[u] = fmaxf(a, 0.f);
        __syncthreads();
        if (u < OUTP) {
            float o = b3[u];
            for (int i = 0; i < H; ++i) o += bf16_to_f32(f32_to_bf16(W3[u * H + i])) * h2[i];
            y[u] = o;
            if (UPSTREAM) {
                dy[u] = u < nout ? Tg[(size_t)c * OUTP + u] : 0.f;
            } else {
                const float t = u < nout ? Tg[(size_t)c * OUTP + u] : o;
                const float diff = o - t;
                dy[u] = 2.f * diff / (float)(B * nout);
                lacc += diff * diff;
            }
        }
        __syncthreads();
        // backward
        float g = 0;
        for (int o = 0; o < OUTP; ++o) g += bf16_to_f32(f32_to_bf16(W3[o * H + u])) * dy[o];
        dh2[u] = h2[u] > 0.f ? g : 0.f;
        __syncthreads();
        g = 0;
        for (int v = 0; v < H; ++v) g += bf16_to_f32(f32_to_bf16(W2[v * H + u])) * dh2[v];
        dh1[u] = h1[u] > 0.f ? g : 0.f;
        __syncthreads();
#pragma unroll
        for (int i = 0; i < INP; ++i) gW1[i] += dh1[u] * x[i];
        gb1 += dh1[u];
        for (int i = 0; i < H; ++i) gW2[i] += dh2[u] * h1[i];
        gb2 += dh2[u];
#pragma unroll
        for (int o = 0; o < OUTP; ++o) gW3[o] += dy[o] * h2[u];
        if (u < OUTP) gb3 += dy[u];
        __syncthreads();
    }
    float* P = part + (size_t)blockIdx.x * NW;
    for (int i = 0; i < INP; ++i) P[u * INP + i] = gW1[i];
    P[H * INP + u] = gb1;
    for (int i = 0; i < H; ++i) P[H * INP + H + u * H + i] = gW2[i];
    P[H * INP + H + H * H + u] = gb2;
    for (int o = 0; o < OUTP; ++o) P[H * INP + H + H * H + H + o * H + u] = gW3[o];
    if (u < OUTP) P[H * INP + H + H * H + H + OUTP * H + u] = gb3;
    // loss partial: sum over the threads that own outputs
    __shared__ float lsum[H];
    lsum[u] = lacc;
    __syncthreads();
    if (u == 0) {
        float t = 0;
        for (int i = 0; i < OUTP; ++i) t += lsum[i];
        lossp[blockIdx.x] = t;
    }
}

__global__ void mlp_reduce_k(int nb, int nw, const float* __restrict__ part, float* __restrict__ grad) {
    for (int w = blockIdx.x * blockDim.x + threadIdx.x; w < nw; w += gridDim.x * blockDim.x) {
        float s = 0;
        for (int b = 0; b < nb; ++b) s += part[(size_t)b * nw + w];
        grad[w] = s;
    }
}

// Adam (Kingma & Ba 2015) on the flat FP32 master parameters
__global__ void adam_k(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       float lr, float b1, float b2, float eps, float c1, float c2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.f - b1) * gi;
        const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// master FP32 (row = output unit) -> transposed BF16 inference copies
__global__ void mlp_to_bf16_k(int rows, int cols, const float* __restrict__ W, uint16_t* __restrict__ Wt) {
    // W[r][c] with r = output unit, c = input: the inference layout is also
    // [output unit][input] (B fragment = 4 consecutive inputs), i.e. the same order
    const int n = rows * cols;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) Wt[i] = f32_to_bf16(W[i]);
}

// ------------------------------------------------------------------ host
Mlp::Mlp(bh_ctx* c, int in_, int hid_, int out_, uint64_t seed) : ctx(c), in(in_), hid(hid_), out(out_) {
    BH_CHECK(hid == 128, "MLP: hidden width must be 128");
    in_p = (in + 15) / 16 * 16;
    out_p = (out + 15) / 16 * 16;
    BH_CHECK(in_p == 16 || in_p == 32, "MLP: input width must be <= 32");
    BH_CHECK(out_p == 16 || out_p == 32, "MLP: output width must be <= 32");
    nw = hid * in_p + hid + hid * hid + hid + out_p * hid + out_p;
    // He-uniform initialisation with a fixed LCG (padded rows/cols zero)
    std::vector<float> w(nw, 0.f);
    uint64_t st = seed * 6364136223846793005ULL + 1442695040888963407ULL;
    auto rnd = [&]() {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        return (double)(st >> 11) / 9007199254740992.0 * 2.0 - 1.0;
    };
    size_t o = 0;
    const double a1 = std::sqrt(6.0 / in), a2 = std::sqrt(6.0 / hid), a3 = std::sqrt(3.0 / hid) * 0.1;
    for (int r = 0; r < hid; ++r) for (int c = 0; c < in_p; ++c) w[o + r * in_p + c] = c < in ? (float)(a1 * rnd()) : 0.f;
    o += hid * in_p + hid;
    for (int r = 0; r < hid; ++r) for (int c = 0; c < hid; ++c) w[o + r * hid + c] = (float)(a2 * rnd());
    o += hid * hid + hid;
    for (int r = 0; r < out_p; ++r) for (int c = 0; c < hid; ++c) w[o + r * hid + c] = r < out ? (float)(a3 * rnd()) : 0.f;
    params.alloc(nw * sizeof(float));
    grad.alloc(nw * sizeof(float));
    m1.alloc(nw * sizeof(float));
    m2.alloc(nw * sizeof(float));
    wbf.alloc((hid * in_p + hid * hid + out_p * hid) * sizeof(uint16_t));
    set_params(w.data());
}

void Mlp::set_params(const float* w) {
    ctx->upload(params.p, w, nw * sizeof(float));
    BH_HIP(hipMemsetAsync(m1.p, 0, nw * sizeof(float), ctx->stream));
    BH_HIP(hipMemsetAsync(m2.p, 0, nw * sizeof(float), ctx->stream));
    step = 0;
    refresh_bf16();
}

void Mlp::refresh_bf16() {
    const float* P = params.as<float>();
    uint16_t* Wb = wbf.as<uint16_t>();
    hipLaunchKernelGGL(mlp_to_bf16_k, dim3(64), dim3(256), 0, ctx->stream, hid, in_p, P + off_W1(), Wb);
    hipLaunchKernelGGL(mlp_to_bf16_k, dim3(64), dim3(256), 0, ctx->stream, hid, hid, P + off_W2(), Wb + hid * in_p);
    hipLaunchKernelGGL(mlp_to_bf16_k, dim3(64), dim3(256), 0, ctx->stream, out_p, hid, P + off_W3(), Wb + hid * in_p + hid * hid);
    BH_HIP(hipGetLastError());
}

void Mlp::infer(int J, const float* X, float* Y) {
    const float* P = params.as<float>();
    const uint16_t* Wb = wbf.as<uint16_t>();
    const int groups = (J + 15) / 16;
    // persistent workgroups (2 per CU) loop over the 16-cell groups: the weight
    // staging is paid once per workgroup, not once per handful of cells
    int blocks = (groups + 7) / 8;
    if (blocks > 512) blocks = 512;
    if (blocks < 1) blocks = 1;
#define BH_CASE(I, O)                                                                                            \
    if (in_p == I && out_p == O) {                                                                               \
        hipLaunchKernelGGL((mlp_infer_k<I, 128, O>), dim3(blocks), dim3(512), 0, ctx->stream, J, X, Wb, P + off_b1(), \
                           Wb + hid * in_p, P + off_b2(), Wb + hid * in_p + hid * hid, P + off_b3(), Y);            \
        BH_HIP(hipGetLastError());                                                                               \
        return;                                                                                                  \
    }
    BH_CASE(16, 16) BH_CASE(16, 32) BH_CASE(32, 16) BH_CASE(32, 32)
#undef BH_CASE
    throw Error(-3, "MLP: unsupported widths");
}

// One training pass: mlp_train_k writes every entry of the nb per-block gradient
// partials and the nb loss partials, and mlp_reduce_k sums the partials in block
// order into grad (no atomics, so results do not depend on scheduling).
void Mlp::launch_gradient(int B, int nb, const int* idx, const float* X, const float* Tg, float* part, float* lossp,
                          bool upstream) {
    const float* P = params.as<float>();
#define BH_CASE(I, O)                                                                                                 \
    if (in_p == I && out_p == O) {                                                                                    \
        if (upstream)                                                                                                 \
            hipLaunchKernelGGL((mlp_train_k<I, 128, O, true>), dim3(nb), dim3(128), 0, ctx->stream, B, idx, X, Tg, out,  \
                               P + off_W1(), P + off_b1(), P + off_W2(), P + off_b2(), P + off_W3(), P + off_b3(), part,  \
                               lossp);                                                                                \
        else                                                                                                          \
            hipLaunchKernelGGL((mlp_train_k<I, 128, O, false>), dim3(nb), dim3(128), 0, ctx->stream, B, idx, X, Tg, out, \
                               P + off_W1(), P + off_b1(), P + off_W2(), P + off_b2(), P + off_W3(), P + off_b3(), part,  \
                               lossp);                                                                                \
    } else
    BH_CASE(16, 16) BH_CASE(16, 32) BH_CASE(32, 16) BH_CASE(32, 32) { throw Error(-3, "MLP: unsupported widths"); }
#undef BH_CASE
    BH_HIP(hipGetLastError());
    hipLaunchKernelGGL(mlp_reduce_k, dim3((nw + 255) / 256), dim3(256), 0, ctx->stream, nb, nw, part, grad.as<float>());
    BH_HIP(hipGetLastError());
}

double Mlp::gradient(int B, const int* idx, const float* X, const float* Tg) {
    const int nb = gradient_blocks(B);
    if (part.bytes < (size_t)nb * nw * sizeof(float)) part.alloc((size_t)256 * nw * sizeof(float));
    if (lossp.bytes < 256 * sizeof(float)) lossp.alloc(256 * sizeof(float));
    launch_gradient(B, nb, idx, X, Tg, part.as<float>(), lossp.as<float>(), false);
    std::vector<float> lp(nb);
    ctx->download(lp.data(), lossp.p, nb * sizeof(float));
    double l = 0;
    for (float v : lp) l += v;
    return l / ((double)B * out);
}

// Gradient of a loss defined outside the network from its upstream gradient dY
// [B][out_p] (first `out` columns used), into grad; same kernels and partials as
// gradient().
void Mlp::backward(int B, const float* X, const float* dY) {
    const int nb = gradient_blocks(B);
    if (part.bytes < (size_t)nb * nw * sizeof(float)) part.alloc((size_t)256 * nw * sizeof(float));
    if (lossp.bytes < 256 * sizeof(float)) lossp.alloc(256 * sizeof(float));
    launch_gradient(B, nb, nullptr, X, dY, part.as<float>(), lossp.as<float>(), true);
}

// ------------------------------------------------------------------ allocation-free gradient
// Fixed workspace of gradient_noalloc: the per-block gradient partials (at most
// 256 blocks of the largest parameter count the widths allow, in_p = out_p = 32)
// followed by the per-block loss partials.  One copy per GPU, 25.5 MB, held until
// the process ends.  In host code `scratch` is only a stand-in for the symbol:
// the kernels receive the device address that hipGetSymbolAddress returns for the
// current GPU (noalloc_ptr), never the array's name.
constexpr size_t kNoallocBlocks = 256;
constexpr size_t kNoallocMaxParams = 128 * 32 + 128 + 128 * 128 + 128 + 32 * 128 + 32;
constexpr size_t kNoallocFloats = kNoallocBlocks * kNoallocMaxParams + kNoallocBlocks;
static __device__ __align__(256) float scratch[kNoallocFloats];

constexpr int kMaxDevices = 64;
static float* noalloc_ptr[kMaxDevices] = {};
static std::mutex noalloc_mu;

size_t mlp_noalloc_capacity() { return kNoallocFloats; }

// Called once per context from bh_init, with the context's device current.
void mlp_noalloc_setup(int device) {
    BH_CHECK(device >= 0 && device < kMaxDevices, "bh_init: device index beyond the allocation-free workspace table");
    std::lock_guard<std::mutex> lock(noalloc_mu);
    if (noalloc_ptr[device]) return;
    void* p = nullptr;
    BH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BH_CHECK(p != nullptr, "bh_init: no device address for the allocation-free workspace");
    noalloc_ptr[device] = (float*)p;
}

__global__ void mlp_loss_k(int nb, const float* __restrict__ lossp, double denom, double* __restrict__ loss) {
    // the same additions, in the same order, as the host sum of gradient()
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double l = 0;
        for (int b = 0; b < nb; ++b) l += lossp[b];
        *loss = l / denom;
    }
}

void Mlp::gradient_noalloc(int B, const float* X, const float* Tg, float* grad_out, double* loss_dev) {
    const int nb = gradient_blocks(B);
    const size_t need = (size_t)nb * nw + (size_t)nb;
    BH_CHECK(need <= kNoallocFloats, "bh_mlp_gradient_noalloc: " + std::to_string(nb) + " blocks x " + std::to_string(nw) +
                                         " parameters exceed the fixed workspace of " + std::to_string(kNoallocFloats) +
                                         " floats");
    float* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(noalloc_mu);
        base = (ctx->device >= 0 && ctx->device < kMaxDevices) ? noalloc_ptr[ctx->device] : nullptr;
    }
    BH_CHECK(base != nullptr, "bh_mlp_gradient_noalloc: the workspace of this GPU is not set up (bh_init)");
    float* part = base;
    float* lossp = base + (size_t)nb * nw;
    launch_gradient(B, nb, nullptr, X, Tg, part, lossp, false);
    if (grad_out)
        BH_HIP(hipMemcpyAsync(grad_out, grad.p, (size_t)nw * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    if (loss_dev) {
        hipLaunchKernelGGL(mlp_loss_k, dim3(1), dim3(64), 0, ctx->stream, nb, lossp, (double)B * out, loss_dev);
        BH_HIP(hipGetLastError());
    }
}

void Mlp::adam(float lr) {
    ++step;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)step), c2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_k, dim3((nw + 255) / 256), dim3(256), 0, ctx->stream, nw, params.as<float>(), grad.as<float>(),
                       m1.as<float>(), m2.as<float>(), lr, b1, b2, 1e-8f, c1, c2);
    BH_HIP(hipGetLastError());
    refresh_bf16();
}

void mlp_features(bh_op* op, const void* r, int prec, float* X, float* s, int in_p) {
    bh_grid* g = op->grid;
    const int J = g->J;
    BH_CHECK((in_p == 16 || in_p == 32) && op->Np + 4 * op->D <= in_p,
             "mlp_features: input width must be 16 or 32 and hold Np + 4 D features");
    const int segs = 256 / in_p;
    dim3 grd((J + segs - 1) / segs > 8192 ? 8192 : (J + segs - 1) / segs), blk(256);
    if (op->D == 2) {
        if (prec == BH_FP32) hipLaunchKernelGGL((mlp_features_k<float, 2>), grd, blk, 0, op->ctx->stream, J, op->Np, g->n[0], g->n[1], g->n[2], (const float*)r, op->nu.as<double>(), in_p, X, s);
        else hipLaunchKernelGGL((mlp_features_k<double, 2>), grd, blk, 0, op->ctx->stream, J, op->Np, g->n[0], g->n[1], g->n[2], (const double*)r, op->nu.as<double>(), in_p, X, s);
    } else {
        if (prec == BH_FP32) hipLaunchKernelGGL((mlp_features_k<float, 3>), grd, blk, 0, op->ctx->stream, J, op->Np, g->n[0], g->n[1], g->n[2], (const float*)r, op->nu.as<double>(), in_p, X, s);
        else hipLaunchKernelGGL((mlp_features_k<double, 3>), grd, blk, 0, op->ctx->stream, J, op->Np, g->n[0], g->n[1], g->n[2], (const double*)r, op->nu.as<double>(), in_p, X, s);
    }
    BH_HIP(hipGetLastError());
}

void mlp_correction(bh_op* op, const float* s, const float* Y, int out_p, void* e, int prec) {
    bh_grid* g = op->grid;
    const int64_t tot = (int64_t)g->J * op->Np;
    dim3 grd((unsigned)std::min<int64_t>((tot + 255) / 256, 8192)), blk(256);
    const double* w = g->d_width.as<double>();
    if (op->D == 2) {
        if (prec == BH_FP32) hipLaunchKernelGGL((mlp_apply_scale_k<float, 2>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, Y, out_p, (float*)e);
        else hipLaunchKernelGGL((mlp_apply_scale_k<double, 2>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, Y, out_p, (double*)e);
    } else {
        if (prec == BH_FP32) hipLaunchKernelGGL((mlp_apply_scale_k<float, 3>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, Y, out_p, (float*)e);
        else hipLaunchKernelGGL((mlp_apply_scale_k<double, 3>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, Y, out_p, (double*)e);
    }
    BH_HIP(hipGetLastError());
}

void mlp_upstream(bh_op* op, const float* s, const double* g, double c, int out_p, float* dY) {
    bh_grid* gr = op->grid;
    const int64_t tot = (int64_t)gr->J * out_p;
    dim3 grd((unsigned)std::min<int64_t>((tot + 255) / 256, 8192)), blk(256);
    const double* w = gr->d_width.as<double>();
    if (op->D == 2) hipLaunchKernelGGL((mlp_upstream_k<2>), grd, blk, 0, op->ctx->stream, gr->J, op->Np, gr->n[0], gr->n[1], w, gr->nmax, op->nu.as<double>(), s, g, c, out_p, dY);
    else hipLaunchKernelGGL((mlp_upstream_k<3>), grd, blk, 0, op->ctx->stream, gr->J, op->Np, gr->n[0], gr->n[1], w, gr->nmax, op->nu.as<double>(), s, g, c, out_p, dY);
    BH_HIP(hipGetLastError());
}

void mlp_targets(bh_op* op, const float* s, const double* e, int out_p, float* Tg) {
    bh_grid* g = op->grid;
    const int64_t tot = (int64_t)g->J * out_p;
    dim3 grd((unsigned)std::min<int64_t>((tot + 255) / 256, 8192)), blk(256);
    const double* w = g->d_width.as<double>();
    if (op->D == 2) hipLaunchKernelGGL((mlp_targets_k<2>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, e, out_p, Tg);
    else hipLaunchKernelGGL((mlp_targets_k<3>), grd, blk, 0, op->ctx->stream, g->J, op->Np, g->n[0], g->n[1], w, g->nmax, op->nu.as<double>(), s, e, out_p, Tg);
    BH_HIP(hipGetLastError());
}

}  // namespace bh

// Multigrid transfer / smoothing kernels for gfx950 (FP64 and FP32).
//
// The hierarchy mirrors the reference's p-multigrid over the orthonormal
// (hierarchical) modal basis plus aggregation h-coarsening
// (AS/PmgConfig.cs:80-200, AS/MultigridOperator.cs:573-640,774-840,
// AS/AggregationGridBasis.cs): restricting to a lower degree keeps the
// leading modes of every cell (the modes are degree-graded), so R is mode
// truncation, P = R^T zero-extends, and the Galerkin operator R A P is the
// leading sub-block of every matrix block.  Aggregation of the last p-level
// merges 2^D cells; the coarse space holds the same-degree polynomials on the
// aggregate box (AggregationGridBasis), injected into each fine cell by the
// L2-projection block P_i (exact: the polynomial spaces nest).
#include "bh_internal.hpp"
#include "bh_mg.hpp"

namespace bh {

static dim3 grid1d(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

// valsc[p][j][i] = valsf[p][j][i] for i, j < bsc (column-major blocks)
__global__ void extract_subblocks_k(int64_t nnzb, int bsf, int bsc, const double* __restrict__ vf, double* __restrict__ vc) {
    const int64_t tot = nnzb * bsc * bsc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / (bsc * bsc);
        const int r = (int)(t - p * bsc * bsc);
        const int j = r / bsc, i = r - j * bsc;
        vc[t] = vf[p * bsf * bsf + (int64_t)j * bsf + i];
    }
}

// Galerkin coarse blocks by gathering (fixed summation order, no atomics):
// Ac[q](m, n) = sum over the fine blocks p of q of sum_{r,c} P_i[r][m] A_p(r, c) P_j[c][n]
// (i, j: fine row and column of p; P row-major per fine cell; blocks column-major)
__global__ __launch_bounds__(256) void galerkin_gather_k(int64_t nq, int bs, const int64_t* __restrict__ qptr,
                                                         const int* __restrict__ plist, const int* __restrict__ prow,
                                                         const int* __restrict__ fcol, const double* __restrict__ fv,
                                                         const int* __restrict__ pidx, const double* __restrict__ P,
                                                         double* __restrict__ cv) {
    const int E = bs * bs;
    const int64_t tot = nq * E;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t / E;
        const int e = (int)(t - q * E);
        const int n = e / bs, m = e - n * bs;
        double s = 0;
        for (int64_t k = qptr[q]; k < qptr[q + 1]; ++k) {
            const int64_t p = plist[k];
            const double* A = fv + p * E;
            const double* Pi = P + (int64_t)pidx[prow[k]] * E;
            const double* Pj = P + (int64_t)pidx[fcol[p]] * E;
            for (int c = 0; c < bs; ++c) {
                double t1 = 0;
                for (int r = 0; r < bs; ++r) t1 += Pi[r * bs + m] * A[c * bs + r];   // (P_i^T A_p)(m, c)
                s += t1 * Pj[c * bs + n];
            }
        }
        cv[t] = s;
    }
}

template <typename T>
__global__ void p_restrict_k(int64_t J, int bsf, int bsc, const T* __restrict__ rf, T* __restrict__ rc) {
    const int64_t tot = J * bsc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / bsc;
        const int m = (int)(t - j * bsc);
        rc[t] = rf[j * bsf + m];
    }
}

template <typename T>
__global__ void p_prolong_add_k(int64_t J, int bsf, int bsc, const T* __restrict__ xc, T* __restrict__ xf) {
    const int64_t tot = J * bsc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = t / bsc;
        const int m = (int)(t - j * bsc);
        xf[j * bsf + m] += xc[t];
    }
}

// Aggregation transfer with per-fine-cell prolongation blocks P_i (bs x bs,
// row-major [fine mode][coarse mode]):  restriction rc_I = sum_{i in I} P_i^T r_i,
// prolongation x_i += P_i xc_parent(i).  One thread per (coarse cell, mode) /
// (fine cell, mode).
template <typename T>
__global__ void agg_restrict_k(int Jc, int bs, const int* __restrict__ cptr, const int* __restrict__ cidx,
                               const int* __restrict__ pidx, const T* __restrict__ P, const T* __restrict__ rf,
                               T* __restrict__ rc) {
    const int64_t tot = (int64_t)Jc * bs;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(t / bs), n = (int)(t - (int64_t)I * bs);
        T s = 0;
        for (int p = cptr[I]; p < cptr[I + 1]; ++p) {
            const int i = cidx[p];
            const T* Pi = P + (size_t)pidx[i] * bs * bs;
            const T* ri = rf + (size_t)i * bs;
            for (int m = 0; m < bs; ++m) s += Pi[m * bs + n] * ri[m];
        }
        rc[t] = s;
    }
}

template <typename T>
__global__ void agg_prolong_add_k(int Jf, int bs, const int* __restrict__ parent, const int* __restrict__ pidx,
                                  const T* __restrict__ P, const T* __restrict__ xc, T* __restrict__ xf) {
    const int64_t tot = (int64_t)Jf * bs;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / bs), m = (int)(t - (int64_t)i * bs);
        const T* Pi = P + (size_t)pidx[i] * bs * bs + (size_t)m * bs;
        const T* xcI = xc + (size_t)parent[i] * bs;
        T s = 0;
        for (int n = 0; n < bs; ++n) s += Pi[n] * xcI[n];
        xf[t] += s;
    }
}

// The same transfers with the block size known at compile time: one thread per
// coarse cell (restriction) or fine cell (prolongation) moves whole blocks.
template <typename T, int BS>
__global__ __launch_bounds__(256) void agg_restrict_bs_k(int Jc, const int* __restrict__ cptr, const int* __restrict__ cidx,
                                                         const int* __restrict__ pidx, const T* __restrict__ P,
                                                         const T* __restrict__ rf, T* __restrict__ rc) {
    for (int I = blockIdx.x * blockDim.x + threadIdx.x; I < Jc; I += gridDim.x * blockDim.x) {
        T acc[BS] = {};
        const int p0 = cptr[I], p1 = cptr[I + 1];
        for (int p = p0; p < p1; ++p) {
            const int i = cidx[p];
            const T* __restrict__ Pi = P + (size_t)pidx[i] * BS * BS;
            T r[BS];
#pragma unroll
            for (int m = 0; m < BS; ++m) r[m] = rf[(size_t)i * BS + m];
#pragma unroll
            for (int m = 0; m < BS; ++m)
#pragma unroll
                for (int n = 0; n < BS; ++n) acc[n] += Pi[m * BS + n] * r[m];
        }
#pragma unroll
        for (int n = 0; n < BS; ++n) rc[(size_t)I * BS + n] = acc[n];
    }
}

template <typename T, int BS>
__global__ __launch_bounds__(256) void agg_prolong_add_bs_k(int Jf, const int* __restrict__ parent, const int* __restrict__ pidx,
                                                            const T* __restrict__ P, const T* __restrict__ xc,
                                                            T* __restrict__ xf) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Jf; i += gridDim.x * blockDim.x) {
        const T* __restrict__ Pi = P + (size_t)pidx[i] * BS * BS;
        const T* __restrict__ xcI = xc + (size_t)parent[i] * BS;
        T c[BS], y[BS];
#pragma unroll
        for (int n = 0; n < BS; ++n) { c[n] = xcI[n]; y[n] = xf[(size_t)i * BS + n]; }
#pragma unroll
        for (int m = 0; m < BS; ++m)
#pragma unroll
            for (int n = 0; n < BS; ++n) y[m] += Pi[m * BS + n] * c[n];
#pragma unroll
        for (int m = 0; m < BS; ++m) xf[(size_t)i * BS + m] = y[m];
    }
}

// x (+)= omega * Dinv r.  A workgroup takes R consecutive block rows: their
// inverse blocks (contiguous, column-major) and residual pieces are staged in
// LDS with coalesced loads by all 256 threads, then thread (row, i) forms
// entry i of its row's product.  R = 256 / BS, capped so that the staged
// blocks stay within 32 KB.
template <typename T, int BS>
struct BlockGemvShape {
    static constexpr int R0 = 256 / BS;
    static constexpr int RC = 32768 / (BS * BS * (int)sizeof(T));
    static constexpr int R = R0 < RC ? (R0 > 0 ? R0 : 1) : (RC > 0 ? RC : 1);
};

template <typename T, int BS, bool ADD>
__global__ __launch_bounds__(256) void block_gemv_k(int nrows, const T* __restrict__ Dinv, const T* __restrict__ r,
                                                    T* __restrict__ x, T omega) {
    constexpr int R = BlockGemvShape<T, BS>::R;
    __shared__ __attribute__((aligned(16))) T sD[R * BS * BS];
    __shared__ T sr[R * BS];
    const int row0 = blockIdx.x * R;
    const int nr = min(R, nrows - row0);
    const T* __restrict__ Dg = Dinv + (size_t)row0 * BS * BS;
    const T* __restrict__ rg = r + (size_t)row0 * BS;
    // 16-byte loads where the staged range allows them (the block start is a multiple of
    // R*BS*BS elements from an allocation start, so 16-byte aligned when that range is)
    constexpr int V16 = 16 / (int)sizeof(T);
    if ((R * BS * BS) % V16 == 0 && (nr * BS * BS) % V16 == 0) {
        const int nv = nr * BS * BS / V16;
        for (int t = threadIdx.x; t < nv; t += blockDim.x) ((uint4*)sD)[t] = ((const uint4*)Dg)[t];
    } else {
        for (int t = threadIdx.x; t < nr * BS * BS; t += blockDim.x) sD[t] = Dg[t];
    }
    for (int t = threadIdx.x; t < nr * BS; t += blockDim.x) sr[t] = rg[t];
    __syncthreads();
    const int row = threadIdx.x / BS, i = threadIdx.x - row * BS;
    if (row < nr) {
        const T* __restrict__ Di = sD + row * BS * BS;
        const T* __restrict__ ri = sr + row * BS;
        T acc = 0;
#pragma unroll 4
        for (int c = 0; c < BS; ++c) acc += Di[c * BS + i] * ri[c];
        const size_t o = (size_t)(row0 + row) * BS + i;
        x[o] = ADD ? x[o] + omega * acc : omega * acc;
    }
}

// The same smoother step with the inverse blocks stored in FP16, each block
// scaled by its largest entry (scale[row], FP32): half the bytes of the FP32
// blocks, FP32 vectors and FP32 accumulation.
template <int BS, bool ADD>
__global__ __launch_bounds__(256) void block_gemv_h_k(int nrows, const _Float16* __restrict__ Dinv, const float* __restrict__ scale,
                                                      const float* __restrict__ r, float* __restrict__ x, float omega) {
    constexpr int R = BlockGemvShape<float, BS>::R;
    __shared__ __attribute__((aligned(16))) _Float16 sD[R * BS * BS + 8];
    __shared__ float sr[R * BS];
    const int row0 = blockIdx.x * R;
    const int nr = min(R, nrows - row0);
    const _Float16* __restrict__ Dg = Dinv + (size_t)row0 * BS * BS;
    const float* __restrict__ rg = r + (size_t)row0 * BS;
    if ((R * BS * BS) % 8 == 0 && (nr * BS * BS) % 8 == 0) {
        const int nv = nr * BS * BS / 8;
        for (int t = threadIdx.x; t < nv; t += blockDim.x) ((uint4*)sD)[t] = ((const uint4*)Dg)[t];
    } else {
        for (int t = threadIdx.x; t < nr * BS * BS; t += blockDim.x) sD[t] = Dg[t];
    }
    for (int t = threadIdx.x; t < nr * BS; t += blockDim.x) sr[t] = rg[t];
    __syncthreads();
    const int row = threadIdx.x / BS, i = threadIdx.x - row * BS;
    if (row < nr) {
        const _Float16* __restrict__ Di = sD + row * BS * BS;
        const float* __restrict__ ri = sr + row * BS;
        float acc = 0;
#pragma unroll 4
        for (int c = 0; c < BS; ++c) acc += (float)Di[c * BS + i] * ri[c];
        const size_t o = (size_t)(row0 + row) * BS + i;
        const float f = omega * scale[row0 + row];
        x[o] = ADD ? x[o] + f * acc : f * acc;
    }
}

// FP16 copy of FP32 blocks: per block, the largest |entry| is the scale, entries / scale in FP16
__global__ __launch_bounds__(256) void blocks_to_f16_k(int nblocks, int bb, const float* __restrict__ src, _Float16* __restrict__ dst,
                                                       float* __restrict__ scale) {
    __shared__ float sm[4];
    for (int b = blockIdx.x; b < nblocks; b += gridDim.x) {
        const float* s = src + (size_t)b * bb;
        float m = 0.f;
        for (int t = threadIdx.x; t < bb; t += blockDim.x) m = fmaxf(m, fabsf(s[t]));
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
        __syncthreads();
        if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
        __syncthreads();
        m = fmaxf(fmaxf(sm[0], sm[1]), fmaxf(sm[2], sm[3]));
        const float sc = m > 0.f ? m : 1.f, rs = 1.f / sc;
        for (int t = threadIdx.x; t < bb; t += blockDim.x) dst[(size_t)b * bb + t] = (_Float16)(s[t] * rs);
        if (threadIdx.x == 0) scale[b] = sc;
    }
}

// ---------------------------------------------------------------- block equilibration
// Per diagonal block D_j (symmetric): D = C S C^T with C lower triangular, C_kk =
// sqrt(|pivot_k|) and S = diag(sign(pivot_k)), and L = C^-1, so that L D L^T = S
// (the reference's SymmInv, MultigridOperator_ChangeOfBasis.cs:624-700: rows and
// columns scaled by 1/sqrt(|M_ii|), diagonal +-1).  One workgroup per block, the block
// in LDS; C and L column-major, zero above the diagonal; S per row.  status: 1 + the
// first block with a zero pivot.
__global__ __launch_bounds__(64) void chol_inv_blocks_k(int nrows, int bs, const int64_t* __restrict__ diag_idx,
                                                        const double* __restrict__ vals, double* __restrict__ Cout,
                                                        double* __restrict__ Lout, double* __restrict__ Sout,
                                                        int* __restrict__ status) {
    extern __shared__ double S[];   // M: bs x bs row-major (becomes C), X: bs x bs row-major (becomes L), sg: bs signs
    double* M = S;
    double* X = S + bs * bs;
    double* sg = S + 2 * bs * bs;
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= nrows) return;
    const int64_t bi = diag_idx[r];
    for (int t = lane; t < bs * bs; t += 64) {
        const int i = t / bs, j = t - i * bs;
        M[t] = bi >= 0 ? vals[(size_t)bi * bs * bs + (size_t)j * bs + i] : 0.0;
        X[t] = 0.0;
    }
    __syncthreads();
    bool ok = true;
    for (int k = 0; k < bs; ++k) {
        const double p = M[k * bs + k];
        if (p == 0.0 || !(fabs(p) < INFINITY)) { ok = false; break; }
        const double sk = p > 0.0 ? 1.0 : -1.0, ck = sqrt(fabs(p));
        __syncthreads();
        for (int i = k + 1 + lane; i < bs; i += 64) M[i * bs + k] /= sk * ck;
        if (lane == 0) { M[k * bs + k] = ck; sg[k] = sk; }
        __syncthreads();
        for (int t = lane; t < bs * bs; t += 64) {
            const int i = t / bs, j = t - i * bs;
            if (j > k && i >= j) M[i * bs + j] -= M[i * bs + k] * sk * M[j * bs + k];
        }
        __syncthreads();
    }
    if (!ok) { if (lane == 0) atomicCAS(status, 0, r + 1); return; }
    // L = C^-1 (lower triangular): column j by lane j, forward substitution
    for (int j = lane; j < bs; j += 64) {
        for (int i = j; i < bs; ++i) {
            double s = (i == j) ? 1.0 : 0.0;
            for (int k = j; k < i; ++k) s -= M[i * bs + k] * X[k * bs + j];
            X[i * bs + j] = s / M[i * bs + i];
        }
    }
    __syncthreads();
    for (int t = lane; t < bs * bs; t += 64) {
        const int i = t / bs, j = t - i * bs;   // store column-major, upper parts zero
        Cout[(size_t)r * bs * bs + (size_t)j * bs + i] = (j <= i) ? M[t] : 0.0;
        Lout[(size_t)r * bs * bs + (size_t)j * bs + i] = (j <= i) ? X[t] : 0.0;
    }
    for (int i = lane; i < bs; i += 64) Sout[(size_t)r * bs + i] = sg[i];
}

// x (+)= omega S r, S a diagonal of signs (block-Jacobi on an equilibrated level)
template <typename T>
__global__ void sign_smooth_k(int64_t n, T omega, const T* __restrict__ S, const T* __restrict__ r, T* __restrict__ x, int add) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (add ? x[i] : T(0)) + omega * S[i] * r[i];
}

// BSR block p (row i, column j): A_p <- L_i A_p L_j^T (column-major blocks), in place
__global__ __launch_bounds__(64) void block_sandwich_k(int64_t nnzb, int bs, const int* __restrict__ prow, const int* __restrict__ col,
                                                       const double* __restrict__ Lb, double* __restrict__ vals) {
    extern __shared__ double S[];   // Li, A, Lj, T (bs x bs each, column-major)
    double* Li = S;
    double* A = S + bs * bs;
    double* Lj = S + 2 * bs * bs;
    double* T = S + 3 * bs * bs;
    const int lane = threadIdx.x;
    for (int64_t p = blockIdx.x; p < nnzb; p += gridDim.x) {
        const int i = prow[p], j = col[p];
        __syncthreads();
        for (int t = lane; t < bs * bs; t += 64) {
            Li[t] = Lb[(size_t)i * bs * bs + t];
            A[t] = vals[(size_t)p * bs * bs + t];
            Lj[t] = Lb[(size_t)j * bs * bs + t];
        }
        __syncthreads();
        // T = A Lj^T:  T(r, c) = sum_k A(r, k) Lj(c, k)
        for (int t = lane; t < bs * bs; t += 64) {
            const int c = t / bs, rr = t - c * bs;
            double s = 0;
            for (int k = 0; k <= c; ++k) s += A[k * bs + rr] * Lj[k * bs + c];
            T[t] = s;
        }
        __syncthreads();
        // A_p = Li T:  (r, c) = sum_k Li(r, k) T(k, c)
        for (int t = lane; t < bs * bs; t += 64) {
            const int c = t / bs, rr = t - c * bs;
            double s = 0;
            for (int k = 0; k <= rr; ++k) s += Li[k * bs + rr] * T[c * bs + k];
            vals[(size_t)p * bs * bs + t] = s;
        }
    }
}

// transformed transfer blocks, row-major [fine mode][coarse mode] per fine cell i:
// Pt_i = C_i^T P_i L_I^T with I = parent(i); P_i from the distinct blocks (pidx) or, when
// P == nullptr, the degree truncation [I; 0]; L == nullptr: the coarse level is not
// transformed (identity)
__global__ __launch_bounds__(256) void transfer_blocks_k(int Jf, int bsf, int bsc, const double* __restrict__ C,
                                                         const int* __restrict__ pidx, const double* __restrict__ P,
                                                         const int* __restrict__ parent, const double* __restrict__ L,
                                                         double* __restrict__ Pt, float* __restrict__ Pt32) {
    const int64_t tot = (int64_t)Jf * bsf * bsc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / (bsf * bsc));
        const int e = (int)(t - (int64_t)i * bsf * bsc);
        const int m = e / bsc, n = e - m * bsc;
        const double* Ci = C + (size_t)i * bsf * bsf;                     // column-major: C(a, m) = Ci[m * bsf + a]
        const double* Pi = P ? P + (size_t)pidx[i] * bsf * bsc : nullptr;  // row-major: P(a, b) = Pi[a * bsc + b]
        const int I = parent ? parent[i] : i;
        const double* LI = L ? L + (size_t)I * bsc * bsc : nullptr;        // column-major: L(n, b) = LI[b * bsc + n]
        double s = 0;
        for (int a = m; a < bsf; ++a) {                                    // C^T(m, a) = C(a, m), zero for a < m
            const double cam = Ci[m * bsf + a];
            double pl = 0;
            for (int b = 0; b < bsc; ++b) {
                const double pab = Pi ? Pi[a * bsc + b] : (a == b ? 1.0 : 0.0);
                const double lnb = LI ? (b <= n ? LI[b * bsc + n] : 0.0) : (b == n ? 1.0 : 0.0);
                pl += pab * lnb;
            }
            s += cam * pl;
        }
        Pt[t] = s;
        if (Pt32) Pt32[t] = (float)s;
    }
}

// transfers with one rectangular block per fine cell (row-major [fine mode][coarse mode]):
// rc_I = sum_{i in I} Pt_i^T r_i,  xf_i += Pt_i xc_{parent(i)}
template <typename T>
__global__ void rect_restrict_k(int Jc, int bsf, int bsc, const int* __restrict__ cptr, const int* __restrict__ cidx,
                                const T* __restrict__ Pt, const T* __restrict__ rf, T* __restrict__ rc) {
    const int64_t tot = (int64_t)Jc * bsc;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(t / bsc), n = (int)(t - (int64_t)I * bsc);
        T s = 0;
        for (int p = cptr[I]; p < cptr[I + 1]; ++p) {
            const int i = cidx[p];
            const T* Pi = Pt + (size_t)i * bsf * bsc;
            const T* ri = rf + (size_t)i * bsf;
            for (int m = 0; m < bsf; ++m) s += Pi[m * bsc + n] * ri[m];
        }
        rc[t] = s;
    }
}

template <typename T>
__global__ void rect_prolong_add_k(int Jf, int bsf, int bsc, const int* __restrict__ parent, const T* __restrict__ Pt,
                                   const T* __restrict__ xc, T* __restrict__ xf) {
    const int64_t tot = (int64_t)Jf * bsf;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t / bsf), m = (int)(t - (int64_t)i * bsf);
        const T* Pi = Pt + (size_t)i * bsf * bsc + (size_t)m * bsc;
        const T* xcI = xc + (size_t)parent[i] * bsc;
        T s = 0;
        for (int n = 0; n < bsc; ++n) s += Pi[n] * xcI[n];
        xf[t] += s;
    }
}

// x = Ainv b (dense, row-major n x n), one wave per row, accumulated in double.
// VEC (n a multiple of 4): each lane reads 4 consecutive entries per step, two
// steps in flight.
template <typename T, typename M, bool VEC>
__global__ __launch_bounds__(256) void dense_gemv_k(int n, const M* __restrict__ A, const T* __restrict__ b, T* __restrict__ x) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    if (row >= n) return;
    const M* __restrict__ Ar = A + (size_t)row * n;
    double s0 = 0, s1 = 0;
    if constexpr (VEC) {
        int c = 4 * lane;
        for (; c + 256 < n; c += 512) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                s0 += (double)Ar[c + e] * (double)b[c + e];
                s1 += (double)Ar[c + 256 + e] * (double)b[c + 256 + e];
            }
        }
        if (c < n) {
#pragma unroll
            for (int e = 0; e < 4; ++e) s0 += (double)Ar[c + e] * (double)b[c + e];
        }
    } else {
        for (int c = lane; c < n; c += 64) s0 += (double)Ar[c] * (double)b[c];
    }
    double s = s0 + s1;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) x[row] = (T)s;
}

template <typename T>
__global__ void copy_k(const T* __restrict__ s, T* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = s[i];
}

// ---------------------------------------------------------------- aggregate Schwarz
// Additive Schwarz over the multigrid aggregates without overlap (the MG-aggregate
// blocking of AS/Schwarz.cs): the block of aggregate I is A restricted to the DOFs
// of its children (slot s = position in the children list, DOF s*bs + i), stored
// dense and column-major with BA = cmax*bs rows; slots beyond the aggregate's child
// count are identity, so every block stays invertible.
__global__ void gather_agg_blocks_k(int nagg, int bs, int cmax, const int* __restrict__ cptr, const int* __restrict__ cidx,
                                    const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                                    const double* __restrict__ vals, double* __restrict__ out) {
    const int BA = cmax * bs;
    const int64_t tot = (int64_t)nagg * cmax * cmax;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(t / (cmax * cmax));
        const int su = (int)(t - (int64_t)I * cmax * cmax);
        const int s = su / cmax, u = su - s * cmax;
        const int nc = cptr[I + 1] - cptr[I];
        double* __restrict__ blk = out + (size_t)I * BA * BA;
        if (s < nc && u < nc) {
            const int rs = cidx[cptr[I] + s], cu = cidx[cptr[I] + u];
            int64_t p = -1;
            for (int64_t q = rowptr[rs]; q < rowptr[rs + 1]; ++q)
                if (col[q] == cu) { p = q; break; }
            for (int j = 0; j < bs; ++j)
                for (int i = 0; i < bs; ++i)
                    blk[(size_t)(u * bs + j) * BA + s * bs + i] = p >= 0 ? vals[(size_t)p * bs * bs + (size_t)j * bs + i] : 0.0;
        } else {
            for (int j = 0; j < bs; ++j)
                for (int i = 0; i < bs; ++i) blk[(size_t)(u * bs + j) * BA + s * bs + i] = (s == u && i == j) ? 1.0 : 0.0;
        }
    }
}

// x (+)= omega * sum_I R_I^T Ainv_I R_I r: thread (aggregate I, block row) with the
// aggregate's residual gathered through its children list (each DOF in one block)
template <typename T, bool ADD>
__global__ __launch_bounds__(256) void schwarz_apply_k(int nagg, int bs, int cmax, const int* __restrict__ cptr,
                                                       const int* __restrict__ cidx, const T* __restrict__ Ainv,
                                                       const T* __restrict__ r, T* __restrict__ x, T omega) {
    const int BA = cmax * bs;
    const int64_t tot = (int64_t)nagg * BA;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(t / BA), row = (int)(t - (int64_t)I * BA);
        const int c0 = cptr[I], nc = cptr[I + 1] - c0;
        const int s = row / bs, i = row - s * bs;
        if (s >= nc) continue;
        const T* __restrict__ Ai = Ainv + (size_t)I * BA * BA;
        T acc = 0;
        for (int u = 0; u < nc; ++u) {
            const T* __restrict__ ru = r + (size_t)cidx[c0 + u] * bs;
            for (int j = 0; j < bs; ++j) acc += Ai[(size_t)(u * bs + j) * BA + row] * ru[j];
        }
        const size_t o = (size_t)cidx[c0 + s] * bs + i;
        x[o] = ADD ? x[o] + omega * acc : omega * acc;
    }
}

// Overlapping blocks (Schwarz with overlap): phase 1, per block I the local solution
// Y_I = Ainv_I R_I r (thread = block row); phase 2, per DOF the sum over the blocks
// that hold it, in a fixed order (no atomics): x (+)= omega sum_I (R_I^T Y_I).
template <typename T>
__global__ __launch_bounds__(256) void schwarz_local_k(int nblk, int bs, int bmax, const int* __restrict__ bptr,
                                                       const int* __restrict__ bidx, const T* __restrict__ Ainv,
                                                       const T* __restrict__ r, T* __restrict__ Y) {
    const int BA = bmax * bs;
    const int64_t tot = (int64_t)nblk * BA;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int I = (int)(t / BA), row = (int)(t - (int64_t)I * BA);
        const int c0 = bptr[I], nc = bptr[I + 1] - c0;
        if (row >= nc * bs) continue;
        const T* __restrict__ Ai = Ainv + (size_t)I * BA * BA;
        T acc = 0;
        for (int u = 0; u < nc; ++u) {
            const T* __restrict__ ru = r + (size_t)bidx[c0 + u] * bs;
            for (int j = 0; j < bs; ++j) acc += Ai[(size_t)(u * bs + j) * BA + row] * ru[j];
        }
        Y[(size_t)c0 * bs + row] = acc;
    }
}

template <typename T, bool ADD>
__global__ __launch_bounds__(256) void schwarz_sum_k(int ncell, int bs, const int* __restrict__ iptr, const int* __restrict__ iidx,
                                                     const T* __restrict__ Y, T* __restrict__ x, T omega, int scale) {
    const int64_t tot = (int64_t)ncell * bs;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(t / bs), i = (int)(t - (int64_t)c * bs);
        T s = 0;
        for (int e = iptr[c]; e < iptr[c + 1]; ++e) s += Y[(size_t)iidx[e] * bs + i];
        if (scale) s = s / (T)(iptr[c + 1] - iptr[c]);   // one over the multiplicity
        x[t] = ADD ? x[t] + omega * s : omega * s;
    }
}

// ---------------------------------------------------------------- LevelPmg
// high-order part of every cell's diagonal block, H_c(i, j) = D_c(lo + i, lo + j)
// (blocks column-major in and out)
__global__ __launch_bounds__(256) void hi_blocks_k(int J, int bs, int lo, const int64_t* __restrict__ diag_idx,
                                                   const double* __restrict__ vals, double* __restrict__ out) {
    const int nh = bs - lo;
    const int64_t tot = (int64_t)J * nh * nh;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(t / (nh * nh));
        const int e = (int)(t - (int64_t)c * nh * nh);
        const int j = e / nh, i = e - j * nh;
        out[t] = vals[(size_t)diag_idx[c] * bs * bs + (size_t)(lo + j) * bs + lo + i];
    }
}

// x_c[lo + i] += sum_j Hinv_c(i, j) r_c[lo + j]: thread (cell, high-order row)
template <typename T>
__global__ __launch_bounds__(256) void hi_smooth_k(int J, int bs, int lo, const T* __restrict__ Hinv, const T* __restrict__ r,
                                                   T* __restrict__ x) {
    const int nh = bs - lo;
    const int64_t tot = (int64_t)J * nh;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int c = (int)(t / nh), i = (int)(t - (int64_t)c * nh);
        const T* __restrict__ H = Hinv + (size_t)c * nh * nh;
        const T* __restrict__ rc = r + (size_t)c * bs + lo;
        T acc = 0;
        for (int j = 0; j < nh; ++j) acc += H[(size_t)j * nh + i] * rc[j];
        x[(size_t)c * bs + lo + i] += acc;
    }
}

// ---------------------------------------------------------------- cell ILU
// Block ILU on a pattern (CellILU, AS/CellILU.cs): W holds the blocks of the pattern
// (column-major, E = bs^2 values each), rows CSR (prp, pcol sorted, pdiag the position
// of the diagonal).  Rows are processed level by level; a row depends only on rows of
// earlier levels, so the rows of a level run in parallel, one workgroup per row.

// W = 0 on the pattern, then A's blocks at their pattern positions
__global__ __launch_bounds__(256) void ilu_scatter_k(int64_t nnzb, int E, const int64_t* __restrict__ wpos,
                                                     const double* __restrict__ A, double* __restrict__ W) {
    const int64_t tot = nnzb * E;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / E;
        W[(size_t)wpos[p] * E + (t - p * E)] = A[t];
    }
}

__device__ inline int64_t ilu_find(const int64_t* prp, const int* pcol, int row, int c) {
    int64_t lo = prp[row], hi = prp[row + 1] - 1;
    while (lo <= hi) {
        const int64_t m = (lo + hi) >> 1;
        const int v = pcol[m];
        if (v == c) return m;
        if (v < c) lo = m + 1;
        else hi = m - 1;
    }
    return -1;
}

// Row i (IKJ order, the same operations per block as the reference's k-i-j loops): for
// every k < i of the row, ascending, W_ik <- W_ik Uinv_k and W_ij <- W_ij - W_ik W_kj for
// the row's j > k that row k holds; then Uinv_i = W_ii^-1.  LDS: the scaled W_ik, and the
// diagonal block for its inversion.
__global__ __launch_bounds__(256) void ilu_factor_k(const int* __restrict__ rows, int n, int bs, const int64_t* __restrict__ prp,
                                                    const int* __restrict__ pcol, const int64_t* __restrict__ pdiag,
                                                    double* __restrict__ W, double* __restrict__ Uinv, int* __restrict__ status) {
    extern __shared__ double sh[];   // L (E), M (E), F (bs)
    __shared__ GjScratch gs;
    if ((int)blockIdx.x >= n) return;
    const int i = rows[blockIdx.x];
    const int t = threadIdx.x, nt = blockDim.x, E = bs * bs;
    double* Ls = sh;
    double* M = sh + E;
    const int64_t p0 = prp[i], p1 = prp[i + 1], pd = pdiag[i];
    for (int64_t p = p0; p < pd; ++p) {
        const int k = pcol[p];
        const double* __restrict__ Uk = Uinv + (size_t)k * E;
        double* __restrict__ Wp = W + (size_t)p * E;
        for (int e = t; e < E; e += nt) {
            const int c = e / bs, r = e - c * bs;
            double s = 0;
            for (int m = 0; m < bs; ++m) s += Wp[(size_t)m * bs + r] * Uk[(size_t)c * bs + m];
            Ls[e] = s;
        }
        __syncthreads();
        for (int e = t; e < E; e += nt) Wp[e] = Ls[e];
        for (int64_t q = p + 1; q < p1; ++q) {
            const int64_t pk = ilu_find(prp, pcol, k, pcol[q]);
            if (pk < 0) continue;
            const double* __restrict__ Wk = W + (size_t)pk * E;
            double* __restrict__ Wq = W + (size_t)q * E;
            for (int e = t; e < E; e += nt) {
                const int c = e / bs, r = e - c * bs;
                double s = 0;
                for (int m = 0; m < bs; ++m) s += Ls[(size_t)m * bs + r] * Wk[(size_t)c * bs + m];
                Wq[e] -= s;
            }
        }
        __syncthreads();
    }
    const double* __restrict__ Wd = W + (size_t)pd * E;
    for (int e = t; e < E; e += nt) M[e] = Wd[e];
    __syncthreads();
    const bool singular = gj_invert_inplace(M, M + E, bs, gs);
    if (singular && t == 0) status[0] = i + 1;
    for (int e = t; e < E; e += nt) Uinv[(size_t)i * E + e] = M[e];
}

// forward sweep of one level: y_i = r_i - sum_{k < i} L_ik y_k, one wave per row
template <typename T>
__global__ __launch_bounds__(256) void ilu_fwd_k(const int* __restrict__ rows, int n, int bs, const int64_t* __restrict__ prp,
                                                 const int* __restrict__ pcol, const int64_t* __restrict__ pdiag,
                                                 const T* __restrict__ W, const T* __restrict__ r, T* __restrict__ y) {
    const int wv = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (wv >= n || lane >= bs) return;
    const int i = rows[wv];
    const int E = bs * bs;
    T s = r[(size_t)i * bs + lane];
    for (int64_t p = prp[i]; p < pdiag[i]; ++p) {
        const T* __restrict__ Lb = W + (size_t)p * E;
        const T* __restrict__ yk = y + (size_t)pcol[p] * bs;
        for (int m = 0; m < bs; ++m) s -= Lb[(size_t)m * bs + lane] * yk[m];
    }
    y[(size_t)i * bs + lane] = s;
}

// backward sweep of one level: z_i = Uinv_i (y_i - sum_{j > i} U_ij z_j) and
// x_i (+)= omega z_i, one wave per row
template <typename T, bool ADD>
__global__ __launch_bounds__(256) void ilu_bwd_k(const int* __restrict__ rows, int n, int bs, const int64_t* __restrict__ prp,
                                                 const int* __restrict__ pcol, const int64_t* __restrict__ pdiag,
                                                 const T* __restrict__ W, const T* __restrict__ Uinv, const T* __restrict__ y,
                                                 T* __restrict__ z, T* __restrict__ x, T omega) {
    __shared__ T sv[4][64];
    const int w = threadIdx.x >> 6, wv = blockIdx.x * 4 + w, lane = threadIdx.x & 63;
    const bool act = wv < n && lane < bs;
    const int i = wv < n ? rows[wv] : 0;
    const int E = bs * bs;
    if (act) {
        T s = y[(size_t)i * bs + lane];
        for (int64_t p = pdiag[i] + 1; p < prp[i + 1]; ++p) {
            const T* __restrict__ Ub = W + (size_t)p * E;
            const T* __restrict__ zj = z + (size_t)pcol[p] * bs;
            for (int m = 0; m < bs; ++m) s -= Ub[(size_t)m * bs + lane] * zj[m];
        }
        sv[w][lane] = s;
    }
    __syncthreads();
    if (act) {
        const T* __restrict__ Ui = Uinv + (size_t)i * E;
        T v = 0;
        for (int m = 0; m < bs; ++m) v += Ui[(size_t)m * bs + lane] * sv[w][m];
        z[(size_t)i * bs + lane] = v;
        const size_t o = (size_t)i * bs + lane;
        x[o] = ADD ? x[o] + omega * v : omega * v;
    }
}

// ---------------------------------------------------------------- K-cycle scalars// ---------------------------------------------------------------- K-cycle scalars
// The inner flexible GMRES of the K-cycle keeps its Hessenberg matrix, Gram-
// Schmidt coefficients and least-squares solution in device memory (layout
// KcycleScalars), so the whole preconditioner needs no host round trip.

// dst = src / sqrt(*s2); zero when sqrt(*s2) <= tiny (or not finite): a zero
// Krylov direction then yields zero columns, which the least squares skip
template <typename T>
__global__ __launch_bounds__(256) void scale_inv_norm_k(const T* __restrict__ src, T* __restrict__ dst, int64_t n,
                                                        const double* __restrict__ s2, double tiny) {
    const double nv = sqrt(fmax(s2[0], 0.0));
    const double f = (nv > tiny && nv < 1e300) ? 1.0 / nv : 0.0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (T)(f * (double)src[i]);
}

// H[i][j] = h1[i] + h2[i] (i <= j), H[j+1][j] = |w|
__global__ void kcycle_column_k(double* __restrict__ S, KcycleScalars L, int j) {
    const int i = threadIdx.x;
    if (i <= j) S[L.H + i * L.kin + j] = S[L.H1 + i] + S[L.H2 + i];
    if (i == 0) S[L.H + (j + 1) * L.kin + j] = sqrt(fmax(S[L.W2], 0.0));
}

// min || beta e1 - H y || by Givens rotations (one thread: kin <= 16); a zero
// diagonal after the rotations (a direction that broke down) gets y = 0
__global__ void kcycle_lsq_k(double* __restrict__ S, KcycleScalars L) {
    if (threadIdx.x != 0) return;
    const int kin = L.kin;
    double R[17 * 16], g[17];
    for (int t = 0; t < (kin + 1) * kin; ++t) R[t] = S[L.H + t];
    for (int t = 0; t <= kin; ++t) g[t] = 0.0;
    g[0] = sqrt(fmax(S[L.B2], 0.0));
    for (int j = 0; j < kin; ++j) {
        const double a = R[j * kin + j], b = R[(j + 1) * kin + j];
        const double rr = hypot(a, b);
        const double c = rr > 0 ? a / rr : 1.0, s = rr > 0 ? b / rr : 0.0;
        for (int k = j; k < kin; ++k) {
            const double t1 = R[j * kin + k], t2 = R[(j + 1) * kin + k];
            R[j * kin + k] = c * t1 + s * t2;
            R[(j + 1) * kin + k] = -s * t1 + c * t2;
        }
        const double t1 = g[j], t2 = g[j + 1];
        g[j] = c * t1 + s * t2;
        g[j + 1] = -s * t1 + c * t2;
    }
    for (int r = kin - 1; r >= 0; --r) {
        double a = g[r];
        for (int c = r + 1; c < kin; ++c) a -= R[r * kin + c] * S[L.Y + c];
        const double d = R[r * kin + r];
        S[L.Y + r] = d != 0.0 ? a / d : 0.0;
    }
}

// learned step: alpha = <q, r'> / <q, q> from st[0..2] = <q, r'>, <q, q>, <r', r'>;
// x += alpha e only when alpha > 0 and finite.  The fraction of |r'|^2 the step
// removes, <q, r'>^2 / (<q, q> <r', r'>), is summed into st[3]; counters
// cnt[0] (taken) / cnt[1] (declined).
// The learned step as the reference's orthonormalisation multigrid adds a correction
// (AS/OrthonormalizationMultigrid.cs:149-300) with a history of one: the cycle's
// correction z1 (x on entry, A z1 = q1) and the network's e (A e = q2) are combined
// as x = g1 z1 + g2 e minimising |b - A x|, i.e. A e orthonormalised against A z1
// and both added with their optimal coefficients.  st: [q1.q1, q1.q2, q2.q2, b.q1,
// b.q2, b.b, sum of gains, counters].  The gain is the fraction of the optimally
// scaled cycle's residual |r|^2 that e removes in addition.  When A e is (nearly)
// in the span of A z1 (|q2 - proj q2| <= 1e-5 |q2|, the reference's cancellation
// test) or nothing is gained, x stays the cycle's result.
template <typename T>
__global__ __launch_bounds__(256) void orth2_step_k(const T* __restrict__ e, T* __restrict__ x, int64_t n,
                                                    double* __restrict__ st, int* __restrict__ cnt) {
    const double g11 = st[0], g12 = st[1], g22 = st[2], f1 = st[3], f2 = st[4], bb = st[5];
    bool take = g11 > 0 && g22 > 0 && isfinite(g11) && isfinite(g22) && isfinite(g12);
    double c1 = 1.0, c2 = 0.0, gain = 0.0;
    if (take) {
        const double perp = g22 - g12 * g12 / g11;   // |q2 - proj_q1 q2|^2
        take = perp > 1e-10 * g22;
        if (take) {
            c2 = (f2 - g12 * f1 / g11) / perp;
            c1 = (f1 - g12 * c2) / g11;
            const double r_mg = bb - f1 * f1 / g11;  // residual^2 of the best multiple of z1
            const double r_new = bb - (c1 * f1 + c2 * f2);
            gain = r_mg > 0 ? 1.0 - r_new / r_mg : 0.0;
            take = gain > 0 && isfinite(c1) && isfinite(c2);
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&cnt[take ? 0 : 1], 1);
        if (take) st[6] += gain;
    }
    if (!take) return;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (T)(c1 * (double)x[i] + c2 * (double)e[i]);
}

// q1 = b - r (the image A z1 of the cycle's correction, from its residual)
template <typename T>
__global__ __launch_bounds__(256) void sub_k(const T* __restrict__ b, const T* __restrict__ r, T* __restrict__ q, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) q[i] = b[i] - r[i];
}

// ------------------------------------------------------------------ launchers
void ilu_scatter(bh_ctx* ctx, int64_t nnzb, int bs, const int64_t* wpos, const double* A, double* W) {
    hipLaunchKernelGGL(ilu_scatter_k, grid1d(nnzb * bs * bs), dim3(256), 0, ctx->stream, nnzb, bs * bs, wpos, A, W);
    BH_HIP(hipGetLastError());
}
void ilu_factor_level(bh_ctx* ctx, const int* rows, int n, int bs, const int64_t* prp, const int* pcol, const int64_t* pdiag,
                      double* W, double* Uinv, int* status) {
    const size_t shm = (2 * (size_t)bs * bs + bs) * sizeof(double);
    hipLaunchKernelGGL(ilu_factor_k, dim3((unsigned)n), dim3(256), shm, ctx->stream, rows, n, bs, prp, pcol, pdiag, W, Uinv, status);
    BH_HIP(hipGetLastError());
}
template <typename T>
void ilu_sweeps(bh_ctx* ctx, int nlev, const std::vector<int>& lvptr, const int* rows, int bs, const int64_t* prp,
                const int* pcol, const int64_t* pdiag, const T* W, const T* Uinv, const T* r, T* y, T* z, T* x, double omega,
                bool add) {
    for (int l = 0; l < nlev; ++l) {
        const int n = lvptr[l + 1] - lvptr[l];
        hipLaunchKernelGGL((ilu_fwd_k<T>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, rows + lvptr[l], n, bs, prp,
                           pcol, pdiag, W, r, y);
    }
    for (int l = nlev - 1; l >= 0; --l) {
        const int n = lvptr[l + 1] - lvptr[l];
        if (add)
            hipLaunchKernelGGL((ilu_bwd_k<T, true>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, rows + lvptr[l], n, bs,
                               prp, pcol, pdiag, W, Uinv, y, z, x, (T)omega);
        else
            hipLaunchKernelGGL((ilu_bwd_k<T, false>), dim3((unsigned)((n + 3) / 4)), dim3(256), 0, ctx->stream, rows + lvptr[l], n,
                               bs, prp, pcol, pdiag, W, Uinv, y, z, x, (T)omega);
    }
    BH_HIP(hipGetLastError());
}

void hi_blocks(bh_ctx* ctx, int J, int bs, int lo, const bh_bsr* A, double* out) {
    hipLaunchKernelGGL(hi_blocks_k, grid1d((int64_t)J * (bs - lo) * (bs - lo)), dim3(256), 0, ctx->stream, J, bs, lo,
                       A->diag_idx.as<int64_t>(), A->vals.as<double>(), out);
    BH_HIP(hipGetLastError());
}
template <typename T>
void hi_smooth(bh_ctx* ctx, int J, int bs, int lo, const T* Hinv, const T* r, T* x) {
    hipLaunchKernelGGL((hi_smooth_k<T>), grid1d((int64_t)J * (bs - lo)), dim3(256), 0, ctx->stream, J, bs, lo, Hinv, r, x);
    BH_HIP(hipGetLastError());
}

void gather_agg_blocks(bh_ctx* ctx, int nagg, int bs, int cmax, const int* cptr, const int* cidx, const bh_bsr* A,
                       double* out) {
    hipLaunchKernelGGL(gather_agg_blocks_k, grid1d((int64_t)nagg * cmax * cmax), dim3(256), 0, ctx->stream, nagg, bs, cmax,
                       cptr, cidx, A->rowptr.as<int64_t>(), A->col.as<int>(), A->vals.as<double>(), out);
    BH_HIP(hipGetLastError());
}
template <typename T>
void schwarz_apply(bh_ctx* ctx, int nagg, int bs, int cmax, const int* cptr, const int* cidx, const T* Ainv, const T* r,
                   T* x, double omega, bool add) {
    const int64_t tot = (int64_t)nagg * cmax * bs;
    if (add) hipLaunchKernelGGL((schwarz_apply_k<T, true>), grid1d(tot), dim3(256), 0, ctx->stream, nagg, bs, cmax, cptr, cidx, Ainv, r, x, (T)omega);
    else hipLaunchKernelGGL((schwarz_apply_k<T, false>), grid1d(tot), dim3(256), 0, ctx->stream, nagg, bs, cmax, cptr, cidx, Ainv, r, x, (T)omega);
    BH_HIP(hipGetLastError());
}

template <typename T>
void schwarz_apply_overlap(bh_ctx* ctx, int nblk, int bs, int bmax, const int* bptr, const int* bidx, const int* iptr,
                           const int* iidx, int ncell, const T* Ainv, const T* r, T* Y, T* x, double omega, bool add,
                           bool scale) {
    hipLaunchKernelGGL((schwarz_local_k<T>), grid1d((int64_t)nblk * bmax * bs), dim3(256), 0, ctx->stream, nblk, bs, bmax, bptr,
                       bidx, Ainv, r, Y);
    if (add) hipLaunchKernelGGL((schwarz_sum_k<T, true>), grid1d((int64_t)ncell * bs), dim3(256), 0, ctx->stream, ncell, bs, iptr, iidx, Y, x, (T)omega, scale ? 1 : 0);
    else hipLaunchKernelGGL((schwarz_sum_k<T, false>), grid1d((int64_t)ncell * bs), dim3(256), 0, ctx->stream, ncell, bs, iptr, iidx, Y, x, (T)omega, scale ? 1 : 0);
    BH_HIP(hipGetLastError());
}
template void schwarz_apply_overlap<float>(bh_ctx*, int, int, int, const int*, const int*, const int*, const int*, int,
                                           const float*, const float*, float*, float*, double, bool, bool);
template void schwarz_apply_overlap<double>(bh_ctx*, int, int, int, const int*, const int*, const int*, const int*, int,
                                            const double*, const double*, double*, double*, double, bool, bool);

template <typename T>
void scale_inv_norm(bh_ctx* ctx, const T* src, T* dst, int64_t n, const double* s2, double tiny) {
    hipLaunchKernelGGL((scale_inv_norm_k<T>), grid1d(n), dim3(256), 0, ctx->stream, src, dst, n, s2, tiny);
    BH_HIP(hipGetLastError());
}
void kcycle_column(bh_ctx* ctx, double* S, const KcycleScalars& L, int j) {
    hipLaunchKernelGGL(kcycle_column_k, dim3(1), dim3(64), 0, ctx->stream, S, L, j);
    BH_HIP(hipGetLastError());
}
void kcycle_lsq(bh_ctx* ctx, double* S, const KcycleScalars& L) {
    hipLaunchKernelGGL(kcycle_lsq_k, dim3(1), dim3(64), 0, ctx->stream, S, L);
    BH_HIP(hipGetLastError());
}
template <typename T>
void orth2_step(bh_ctx* ctx, const T* e, T* x, int64_t n, double* st, int* cnt) {
    hipLaunchKernelGGL((orth2_step_k<T>), grid1d(n), dim3(256), 0, ctx->stream, e, x, n, st, cnt);
    BH_HIP(hipGetLastError());
}
template <typename T>
void sub_vec(bh_ctx* ctx, const T* b, const T* r, T* q, int64_t n) {
    hipLaunchKernelGGL((sub_k<T>), grid1d(n), dim3(256), 0, ctx->stream, b, r, q, n);
    BH_HIP(hipGetLastError());
}
void galerkin_gather(bh_ctx* ctx, int64_t nq, int bs, const int64_t* qptr, const int* plist, const int* prow,
                     const int* fcol, const double* fv, const int* pidx, const double* P, double* cv) {
    hipLaunchKernelGGL(galerkin_gather_k, grid1d(nq * bs * bs), dim3(256), 0, ctx->stream, nq, bs, qptr, plist, prow, fcol, fv,
                       pidx, P, cv);
    BH_HIP(hipGetLastError());
}
void extract_subblocks(bh_ctx* ctx, int64_t nnzb, int bsf, int bsc, const double* vf, double* vc) {
    hipLaunchKernelGGL(extract_subblocks_k, grid1d(nnzb * bsc * bsc), dim3(256), 0, ctx->stream, nnzb, bsf, bsc, vf, vc);
    BH_HIP(hipGetLastError());
}

template <typename T>
void p_restrict(bh_ctx* ctx, int64_t J, int bsf, int bsc, const T* rf, T* rc) {
    hipLaunchKernelGGL((p_restrict_k<T>), grid1d(J * bsc), dim3(256), 0, ctx->stream, J, bsf, bsc, rf, rc);
}
template <typename T>
void p_prolong_add(bh_ctx* ctx, int64_t J, int bsf, int bsc, const T* xc, T* xf) {
    hipLaunchKernelGGL((p_prolong_add_k<T>), grid1d(J * bsc), dim3(256), 0, ctx->stream, J, bsf, bsc, xc, xf);
}
template <typename T>
void agg_restrict(bh_ctx* ctx, int Jc, int bs, const int* cptr, const int* cidx, const int* pidx, const T* P, const T* rf, T* rc) {
    if (bs == 4)
        hipLaunchKernelGGL((agg_restrict_bs_k<T, 4>), grid1d(Jc), dim3(256), 0, ctx->stream, Jc, cptr, cidx, pidx, P, rf, rc);
    else if (bs == 3)
        hipLaunchKernelGGL((agg_restrict_bs_k<T, 3>), grid1d(Jc), dim3(256), 0, ctx->stream, Jc, cptr, cidx, pidx, P, rf, rc);
    else
        hipLaunchKernelGGL((agg_restrict_k<T>), grid1d((int64_t)Jc * bs), dim3(256), 0, ctx->stream, Jc, bs, cptr, cidx, pidx, P,
                           rf, rc);
}
template <typename T>
void agg_prolong_add(bh_ctx* ctx, int Jf, int bs, const int* parent, const int* pidx, const T* P, const T* xc, T* xf) {
    if (bs == 4)
        hipLaunchKernelGGL((agg_prolong_add_bs_k<T, 4>), grid1d(Jf), dim3(256), 0, ctx->stream, Jf, parent, pidx, P, xc, xf);
    else if (bs == 3)
        hipLaunchKernelGGL((agg_prolong_add_bs_k<T, 3>), grid1d(Jf), dim3(256), 0, ctx->stream, Jf, parent, pidx, P, xc, xf);
    else
        hipLaunchKernelGGL((agg_prolong_add_k<T>), grid1d((int64_t)Jf * bs), dim3(256), 0, ctx->stream, Jf, bs, parent, pidx, P,
                           xc, xf);
}

template <typename T>
void block_gemv(bh_ctx* ctx, int nrows, int bs, const T* Dinv, const T* r, T* x, double omega, bool add) {
    dim3 blk(256);
#define BH_CASE(B)                                                                                                \
    case B:                                                                                                       \
        grd = dim3((nrows + BlockGemvShape<T, B>::R - 1) / BlockGemvShape<T, B>::R);                           \
        if (add) hipLaunchKernelGGL((block_gemv_k<T, B, true>), grd, blk, 0, ctx->stream, nrows, Dinv, r, x, (T)omega); \
        else hipLaunchKernelGGL((block_gemv_k<T, B, false>), grd, blk, 0, ctx->stream, nrows, Dinv, r, x, (T)omega);    \
        break;
    dim3 grd;
    switch (bs) {
        BH_CASE(1) BH_CASE(2) BH_CASE(3) BH_CASE(4) BH_CASE(5) BH_CASE(6) BH_CASE(8) BH_CASE(10)
        BH_CASE(15) BH_CASE(16) BH_CASE(20) BH_CASE(21) BH_CASE(28) BH_CASE(32) BH_CASE(35) BH_CASE(50) BH_CASE(56)
        default: throw Error(-3, "block_gemv: unsupported block size");
    }
#undef BH_CASE
    BH_HIP(hipGetLastError());
}

void blocks_to_f16(bh_ctx* ctx, int nblocks, int bs, const float* src, _Float16* dst, float* scale) {
    hipLaunchKernelGGL(blocks_to_f16_k, dim3((unsigned)std::min(nblocks, 16384)), dim3(256), 0, ctx->stream, nblocks, bs * bs, src,
                       dst, scale);
    BH_HIP(hipGetLastError());
}

void block_gemv_h(bh_ctx* ctx, int nrows, int bs, const _Float16* Dinv, const float* scale, const float* r, float* x, double omega,
                  bool add) {
    dim3 blk(256), grd;
#define BH_CASE(B)                                                                                                   \
    case B:                                                                                                          \
        grd = dim3((nrows + BlockGemvShape<float, B>::R - 1) / BlockGemvShape<float, B>::R);                        \
        if (add) hipLaunchKernelGGL((block_gemv_h_k<B, true>), grd, blk, 0, ctx->stream, nrows, Dinv, scale, r, x, (float)omega); \
        else hipLaunchKernelGGL((block_gemv_h_k<B, false>), grd, blk, 0, ctx->stream, nrows, Dinv, scale, r, x, (float)omega); \
        break;
    switch (bs) {
        BH_CASE(3) BH_CASE(4) BH_CASE(6) BH_CASE(10) BH_CASE(15) BH_CASE(20) BH_CASE(21) BH_CASE(28) BH_CASE(35)
        default: throw Error(-3, "block_gemv (FP16 blocks): unsupported block size");
    }
#undef BH_CASE
    BH_HIP(hipGetLastError());
}

void chol_inv_blocks(bh_bsr* A, DevBuf& C, DevBuf& L, DevBuf& S) {
    A->ensure_diag();
    const int bs = A->bs;
    BH_CHECK(bs <= 40, "block equilibration supports block sizes <= 40");
    C.alloc((size_t)A->nrows * bs * bs * sizeof(double));
    L.alloc((size_t)A->nrows * bs * bs * sizeof(double));
    S.alloc((size_t)A->nrows * bs * sizeof(double));
    DevBuf st(sizeof(int));
    const int zero = 0;
    A->ctx->upload(st.p, &zero, sizeof(int));
    hipLaunchKernelGGL(chol_inv_blocks_k, dim3(A->nrows), dim3(64), ((size_t)2 * bs * bs + bs) * sizeof(double), A->ctx->stream,
                       A->nrows, bs, A->diag_idx.as<int64_t>(), A->vals.as<double>(), C.as<double>(), L.as<double>(),
                       S.as<double>(), st.as<int>());
    BH_HIP(hipGetLastError());
    int h = 0;
    A->ctx->download(&h, st.p, sizeof(int));
    BH_CHECK(h == 0, "block equilibration: diagonal block " + std::to_string(h - 1) + " has a zero pivot");
}

template <typename T>
void sign_smooth(bh_ctx* ctx, int64_t n, double omega, const T* S, const T* r, T* x, bool add) {
    hipLaunchKernelGGL((sign_smooth_k<T>), grid1d(n), dim3(256), 0, ctx->stream, n, (T)omega, S, r, x, add ? 1 : 0);
}
template void sign_smooth<float>(bh_ctx*, int64_t, double, const float*, const float*, float*, bool);
template void sign_smooth<double>(bh_ctx*, int64_t, double, const double*, const double*, double*, bool);

void block_sandwich(bh_bsr* A, const double* L) {
    std::vector<int> prow((size_t)A->nnzb);
    for (int r = 0; r < A->nrows; ++r)
        for (int64_t p = A->h_rowptr[r]; p < A->h_rowptr[r + 1]; ++p) prow[p] = r;
    DevBuf dp(std::max<size_t>(prow.size(), 1) * sizeof(int));
    A->ctx->upload(dp.p, prow.data(), prow.size() * sizeof(int));
    const int bs = A->bs;
    const unsigned g = (unsigned)std::min<int64_t>(A->nnzb, 65536);
    hipLaunchKernelGGL(block_sandwich_k, dim3(g > 0 ? g : 1), dim3(64), (size_t)4 * bs * bs * sizeof(double), A->ctx->stream,
                       A->nnzb, bs, dp.as<int>(), A->col.as<int>(), L, A->vals.as<double>());
    BH_HIP(hipGetLastError());
    A->ctx->sync();
    A->vals32.release();   // stale FP32 copy
}

void transfer_blocks(bh_ctx* ctx, int Jf, int bsf, int bsc, const double* C, const int* pidx, const double* P, const int* parent,
                     const double* L, double* Pt, float* Pt32) {
    hipLaunchKernelGGL(transfer_blocks_k, grid1d((int64_t)Jf * bsf * bsc), dim3(256), 0, ctx->stream, Jf, bsf, bsc, C, pidx, P,
                       parent, L, Pt, Pt32);
    BH_HIP(hipGetLastError());
}

template <typename T>
void rect_restrict(bh_ctx* ctx, int Jc, int bsf, int bsc, const int* cptr, const int* cidx, const T* Pt, const T* rf, T* rc) {
    hipLaunchKernelGGL((rect_restrict_k<T>), grid1d((int64_t)Jc * bsc), dim3(256), 0, ctx->stream, Jc, bsf, bsc, cptr, cidx, Pt,
                       rf, rc);
}
template <typename T>
void rect_prolong_add(bh_ctx* ctx, int Jf, int bsf, int bsc, const int* parent, const T* Pt, const T* xc, T* xf) {
    hipLaunchKernelGGL((rect_prolong_add_k<T>), grid1d((int64_t)Jf * bsf), dim3(256), 0, ctx->stream, Jf, bsf, bsc, parent, Pt,
                       xc, xf);
}
template void rect_restrict<float>(bh_ctx*, int, int, int, const int*, const int*, const float*, const float*, float*);
template void rect_restrict<double>(bh_ctx*, int, int, int, const int*, const int*, const double*, const double*, double*);
template void rect_prolong_add<float>(bh_ctx*, int, int, int, const int*, const float*, const float*, float*);
template void rect_prolong_add<double>(bh_ctx*, int, int, int, const int*, const double*, const double*, double*);

template <typename T, typename M>
void dense_gemv(bh_ctx* ctx, int n, const M* A, const T* b, T* x) {
    if (n % 4 == 0)
        hipLaunchKernelGGL((dense_gemv_k<T, M, true>), dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, A, b, x);
    else
        hipLaunchKernelGGL((dense_gemv_k<T, M, false>), dim3((n + 3) / 4), dim3(256), 0, ctx->stream, n, A, b, x);
}

template <typename T>
void vcopy(bh_ctx* ctx, const T* s, T* d, int64_t n) {
    hipLaunchKernelGGL((copy_k<T>), grid1d(n), dim3(256), 0, ctx->stream, s, d, n);
}

#define BH_INST(T)                                                                                   \
    template void p_restrict<T>(bh_ctx*, int64_t, int, int, const T*, T*);                           \
    template void p_prolong_add<T>(bh_ctx*, int64_t, int, int, const T*, T*);                        \
    template void agg_restrict<T>(bh_ctx*, int, int, const int*, const int*, const int*, const T*, const T*, T*); \
    template void agg_prolong_add<T>(bh_ctx*, int, int, const int*, const int*, const T*, const T*, T*); \
    template void block_gemv<T>(bh_ctx*, int, int, const T*, const T*, T*, double, bool);             \
    template void dense_gemv<T, double>(bh_ctx*, int, const double*, const T*, T*);                       \
    template void dense_gemv<T, float>(bh_ctx*, int, const float*, const T*, T*);                         \
    template void vcopy<T>(bh_ctx*, const T*, T*, int64_t);                                          \
    template void scale_inv_norm<T>(bh_ctx*, const T*, T*, int64_t, const double*, double);           \
    template void schwarz_apply<T>(bh_ctx*, int, int, int, const int*, const int*, const T*, const T*, T*, double, bool); \
    template void orth2_step<T>(bh_ctx*, const T*, T*, int64_t, double*, int*); \
    template void sub_vec<T>(bh_ctx*, const T*, const T*, T*, int64_t);            \
    template void hi_smooth<T>(bh_ctx*, int, int, int, const T*, const T*, T*);     \
    template void ilu_sweeps<T>(bh_ctx*, int, const std::vector<int>&, const int*, int, const int64_t*, const int*, \
                                const int64_t*, const T*, const T*, const T*, T*, T*, T*, double, bool);
BH_INST(float)
BH_INST(double)
#undef BH_INST

}  // namespace bh

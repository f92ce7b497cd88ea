// Block-sparse (BlockMsrMatrix) SpMV, block-Jacobi, batched block inversion
// and the vector kernels of the Krylov solvers, for gfx950.
//
// BSR SpMV design: one wave64 per block row.  Blocks are stored column-major
// and the blocks of one row are contiguous, so the row's values form a stream
// of "columns" of bs entries.  The 64 lanes are split into G = 64/bs groups of
// bs lanes (lane = g*bs + r); at step t the groups read columns t*G .. t*G+G-1,
// i.e. G*bs consecutive values: every value load is fully coalesced and each
// value is read exactly once (the HBM-bound part).  x entries are
// group-uniform gathers served by L1/L2.  Group partial sums are combined in
// LDS.  This is the successor of the reference's cell-blocked CSR kernel
// blockMultiply2 (ilPSP.LinSolvers.monkey/cuda-src/CudaMatrixKernelDP.cu:57).
#include <type_traits>

#include "bh_internal.hpp"

namespace bh {

template <typename T, int BS>
__global__ __launch_bounds__(256) void bsr_spmv_kernel(int nrows, const int64_t* __restrict__ rowptr,
                                                       const int* __restrict__ col, const T* __restrict__ vals,
                                                       const T* __restrict__ x, T* __restrict__ y, T alpha, T beta,
                                                       const T* __restrict__ b_sub) {
    constexpr int G = 64 / BS;
    __shared__ T red[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    const bool active = r < nrows;
    const int g = lane / BS, ri = lane - g * BS;
    T acc = 0;
    if (active && g < G) {
        const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
        const int ncol = (int)(p1 - p0) * BS;
        const T* __restrict__ base = vals + (size_t)p0 * BS * BS;
        int q = g;
#pragma unroll 4
        for (; q < ncol; q += G) {
            const int b = q / BS, cc = q - b * BS;
            const T xv = x[(size_t)col[p0 + b] * BS + cc];
            acc += base[(size_t)q * BS + ri] * xv;
        }
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (active && g == 0) {
        T s = acc;
#pragma unroll
        for (int gg = 1; gg < G; ++gg) s += red[wave][gg * BS + ri];
        const size_t o = (size_t)r * BS + ri;
        T out = alpha * s;
        if (b_sub) out = b_sub[o] - s;          // residual form: y = b - A x
        else if (beta != T(0)) out += beta * y[o];
        y[o] = out;
    }
}

// The same stream with VB-byte value loads (16, or 8 where a column is not whole
// 16-byte units, e.g. bs = 10 in FP32), for block sizes whose columns are whole VB-byte
// units: a lane reads VW = VB / sizeof(T) consecutive rows of one column, L = bs / VW
// lanes cover a column, and the G = 64 / L groups read G consecutive columns per step
// (bs = 20 FP64: 10 lanes per column, 6 columns = 960 contiguous bytes per load
// instruction).  Each lane keeps VW partial sums; the groups' partials are combined in LDS.
template <typename T, int BS, int VB = 16>
__global__ __launch_bounds__(256) void bsr_spmv_vec_kernel(int nrows, const int64_t* __restrict__ rowptr,
                                                           const int* __restrict__ col, const T* __restrict__ vals,
                                                           const T* __restrict__ x, T* __restrict__ y, T alpha, T beta,
                                                           const T* __restrict__ b_sub) {
    constexpr int VW = VB / sizeof(T);
    static_assert(BS % VW == 0, "columns must be whole VB-byte units");
    constexpr int L = BS / VW, G = 64 / L;
    using V16 = typename std::conditional<sizeof(T) == 8, double2, float4>::type;
    using V8 = typename std::conditional<sizeof(T) == 8, double, float2>::type;
    using V = typename std::conditional<VB == 16, V16, V8>::type;
    __shared__ T red[4][VW][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    const bool active = r < nrows;
    const int g = lane / L, li = lane - g * L;
    T acc[VW];
#pragma unroll
    for (int v = 0; v < VW; ++v) acc[v] = 0;
    if (active && g < G) {
        const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
        const int ncol = (int)(p1 - p0) * BS;
        const T* __restrict__ base = vals + (size_t)p0 * BS * BS + li * VW;
#pragma unroll 4
        for (int q = g; q < ncol; q += G) {
            const int b = q / BS, cc = q - b * BS;
            const T xv = x[(size_t)col[p0 + b] * BS + cc];
            const V a = *reinterpret_cast<const V*>(base + (size_t)q * BS);
            const T* av = reinterpret_cast<const T*>(&a);
#pragma unroll
            for (int v = 0; v < VW; ++v) acc[v] += av[v] * xv;
        }
    }
#pragma unroll
    for (int v = 0; v < VW; ++v) red[wave][v][lane] = acc[v];
    __syncthreads();
    if (active && g == 0) {
#pragma unroll
        for (int v = 0; v < VW; ++v) {
            T s = acc[v];
#pragma unroll
            for (int gg = 1; gg < G; ++gg) s += red[wave][v][gg * L + li];
            const size_t o = (size_t)r * BS + li * VW + v;
            T out = alpha * s;
            if (b_sub) out = b_sub[o] - s;      // residual form: y = b - A x
            else if (beta != T(0)) out += beta * y[o];
            y[o] = out;
        }
    }
}

// Small blocks (bs <= 6: the k = 1 levels and the aggregation levels): one group of
// bs lanes per block row, RPW = 64 / bs block rows per wave.  Lane j of a group reads
// column j of each block of its row (bs consecutive values), multiplies it by x_j of
// the block's column and keeps bs partial sums; the group's bs lanes then exchange
// their partials so that lane i ends with row i, and consecutive lanes write
// consecutive rows.  A wave thus covers RPW rows at once instead of one, and the
// dependent loads (row pointer, block columns, x) are paid once per RPW rows.
template <typename T, int BS>
__global__ __launch_bounds__(256) void bsr_spmv_small_kernel(int nrows, const int64_t* __restrict__ rowptr,
                                                             const int* __restrict__ col, const T* __restrict__ vals,
                                                             const T* __restrict__ x, T* __restrict__ y, T alpha, T beta,
                                                             const T* __restrict__ b_sub) {
    constexpr int RPW = 64 / BS;
    const int lane = threadIdx.x & 63;
    const int q = lane / BS, j = lane - q * BS;
    const int r = (blockIdx.x * 4 + (threadIdx.x >> 6)) * RPW + q;
    const bool active = q < RPW && r < nrows;
    T acc[BS];
#pragma unroll
    for (int i = 0; i < BS; ++i) acc[i] = 0;
    if (active) {
        const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
#pragma unroll 4
        for (int64_t p = p0; p < p1; ++p) {
            const T xj = x[(size_t)col[p] * BS + j];
            const T* __restrict__ cj = vals + (size_t)p * BS * BS + (size_t)j * BS;
            T v[BS];
#pragma unroll
            for (int i = 0; i < BS; ++i) v[i] = cj[i];
#pragma unroll
            for (int i = 0; i < BS; ++i) acc[i] += v[i] * xj;
        }
    }
    // row i of the group = sum over its lanes j of acc_j[i]; lane i keeps it
    const int base = lane - j;
    T s = 0;
#pragma unroll
    for (int i = 0; i < BS; ++i) {
        T t = 0;
#pragma unroll
        for (int jj = 0; jj < BS; ++jj) t += __shfl(acc[i], base + jj, 64);
        if (j == i) s = t;
    }
    if (active) {
        const size_t o = (size_t)r * BS + j;
        T out = alpha * s;
        if (b_sub) out = b_sub[o] - s;          // residual form: y = b - A x
        else if (beta != T(0)) out += beta * y[o];
        y[o] = out;
    }
}

template <typename T, int BS>
static void launch_spmv_bs(dim3 grd, hipStream_t st, int nrows, const int64_t* rp, const int* cl, const T* vals, const T* x,
                           T* y, T alpha, T beta, const T* bsub) {
    // small blocks: several rows per wave; 16-byte loads where a column is whole 16-byte
    // units and spans at least 5 lanes (with fewer lanes per column, more than 12 groups
    // would leave the final LDS reduction, serial over the groups, longer than the
    // stream it saves); else the one-row-per-wave stream
    if constexpr (BS <= 6) {
        constexpr int RPW = 64 / BS;
        const dim3 g2((unsigned)((nrows + 4 * RPW - 1) / (4 * RPW)));
        hipLaunchKernelGGL((bsr_spmv_small_kernel<T, BS>), g2, dim3(256), 0, st, nrows, rp, cl, vals, x, y, alpha, beta, bsub);
    } else if constexpr ((BS * sizeof(T)) % 16 == 0 && BS / (16 / sizeof(T)) >= 5)
        hipLaunchKernelGGL((bsr_spmv_vec_kernel<T, BS, 16>), grd, dim3(256), 0, st, nrows, rp, cl, vals, x, y, alpha, beta, bsub);
    else if constexpr (sizeof(T) == 4 && (BS * sizeof(T)) % 8 == 0 && BS / 2 >= 5)
        hipLaunchKernelGGL((bsr_spmv_vec_kernel<T, BS, 8>), grd, dim3(256), 0, st, nrows, rp, cl, vals, x, y, alpha, beta, bsub);
    else
        hipLaunchKernelGGL((bsr_spmv_kernel<T, BS>), grd, dim3(256), 0, st, nrows, rp, cl, vals, x, y, alpha, beta, bsub);
}

template <typename T>
static void launch_spmv(bh_bsr* A, T alpha, const T* x, T beta, T* y, const T* bsub, const T* vals) {
    dim3 grd((A->nrows + 3) / 4);
    hipStream_t st = A->ctx->stream;
    const int64_t* rp = A->rowptr.as<int64_t>();
    const int* cl = A->col.as<int>();
    // the 16-byte value loads need a 16-byte aligned value array (device allocations are)
    BH_CHECK(((uintptr_t)vals & 15) == 0, "bsr_spmv: value array not 16-byte aligned");
#define BH_CASE(B) case B: launch_spmv_bs<T, B>(grd, st, A->nrows, rp, cl, vals, x, y, alpha, beta, bsub); break;
    switch (A->bs) {
        BH_CASE(1) BH_CASE(2) BH_CASE(3) BH_CASE(4) BH_CASE(5) BH_CASE(6) BH_CASE(8) BH_CASE(10)
        BH_CASE(15) BH_CASE(16) BH_CASE(20) BH_CASE(21) BH_CASE(28) BH_CASE(32) BH_CASE(35) BH_CASE(50) BH_CASE(56)
        default: throw Error(-3, "bsr_spmv: unsupported block size " + std::to_string(A->bs));
    }
#undef BH_CASE
    BH_HIP(hipGetLastError());
}

void bsr_spmv(bh_bsr* A, double alpha, const void* x, double beta, void* y, int prec) {
    if (prec == BH_FP64) {
        launch_spmv<double>(A, alpha, (const double*)x, beta, (double*)y, nullptr, A->vals.as<double>());
    } else if (prec == BH_FP32) {
        A->ensure_fp32();
        launch_spmv<float>(A, (float)alpha, (const float*)x, (float)beta, (float*)y, nullptr, A->vals32.as<float>());
    } else throw Error(-3, "bsr_spmv: unsupported precision");
}

// residual r = b - A x
void bsr_residual(bh_bsr* A, const void* b, const void* x, void* r, int prec) {
    if (prec == BH_FP64) {
        launch_spmv<double>(A, 1.0, (const double*)x, 0.0, (double*)r, (const double*)b, A->vals.as<double>());
    } else {
        A->ensure_fp32();
        launch_spmv<float>(A, 1.0f, (const float*)x, 0.0f, (float*)r, (const float*)b, A->vals32.as<float>());
    }
}

// ------------------------------------------------------------ block inverse
// Batched Gauss-Jordan with partial pivoting, one wave per block (bs <= 64),
// lane = column of the augmented system [A | I].  Replaces the per-cell
// InvertInPlace (LAPACK DGETRF/DGETRI, IL/IMatrixExtensions.cs:2221-2335)
// of BlockJacobi.Init (AS/BlockJacobi.cs:100-146).
__global__ __launch_bounds__(64) void invert_blocks_kernel(int nrows, int bs, const int64_t* __restrict__ diag_idx,
                                                           const double* __restrict__ vals, double* __restrict__ inv,
                                                           float* __restrict__ inv32, int* __restrict__ status) {
    extern __shared__ double M[];   // bs x (2 bs), row-major
    const int r = blockIdx.x, lane = threadIdx.x;
    if (r >= nrows) return;
    const int W = 2 * bs;
    const int64_t bi = diag_idx[r];
    for (int i = 0; i < bs; ++i)
        for (int j = lane; j < W; j += 64) {
            double v;
            if (j < bs) v = (bi >= 0) ? vals[(size_t)bi * bs * bs + (size_t)j * bs + i] : 0.0;  // column-major block
            else v = (j - bs == i) ? 1.0 : 0.0;
            M[i * W + j] = v;
        }
    __syncthreads();
    bool ok = true;
    for (int c = 0; c < bs; ++c) {
        // pivot search (every lane scans; bs <= 64 so this is cheap)
        int p = c;
        double best = fabs(M[c * W + c]);
        for (int i = c + 1; i < bs; ++i) {
            double v = fabs(M[i * W + c]);
            if (v > best) { best = v; p = i; }
        }
        if (best == 0.0) { ok = false; break; }
        __syncthreads();
        if (p != c)
            for (int j = lane; j < W; j += 64) { double t = M[p * W + j]; M[p * W + j] = M[c * W + j]; M[c * W + j] = t; }
        __syncthreads();
        const double piv = 1.0 / M[c * W + c];
        __syncthreads();
        for (int j = lane; j < W; j += 64) M[c * W + j] *= piv;
        __syncthreads();
        for (int i = 0; i < bs; ++i) {
            if (i == c) continue;
            const double f = M[i * W + c];
            __syncthreads();
            for (int j = lane; j < W; j += 64) M[i * W + j] -= f * M[c * W + j];
            __syncthreads();
        }
    }
    if (!ok) { if (lane == 0) status[0] = r + 1; return; }
    // store the inverse column-major (as the matrix blocks)
    for (int i = 0; i < bs; ++i)
        for (int j = lane; j < bs; j += 64) {
            const double v = M[i * W + bs + j];
            inv[(size_t)r * bs * bs + (size_t)j * bs + i] = v;
            if (inv32) inv32[(size_t)r * bs * bs + (size_t)j * bs + i] = (float)v;
        }
}

void bsr_invert_diag(bh_bsr* A, DevBuf& dinv64, DevBuf* dinv32) {
    A->ensure_diag();
    const int bs = A->bs;
    BH_CHECK(bs <= 64, "block inversion supports block sizes <= 64");
    dinv64.alloc((size_t)A->nrows * bs * bs * sizeof(double));
    if (dinv32) dinv32->alloc((size_t)A->nrows * bs * bs * sizeof(float));
    DevBuf st(sizeof(int));
    BH_HIP(hipMemsetAsync(st.p, 0, sizeof(int), A->ctx->stream));
    size_t shm = (size_t)bs * 2 * bs * sizeof(double);
    hipLaunchKernelGGL(invert_blocks_kernel, dim3(A->nrows), dim3(64), shm, A->ctx->stream, A->nrows, bs,
                       A->diag_idx.as<int64_t>(), A->vals.as<double>(), dinv64.as<double>(),
                       dinv32 ? dinv32->as<float>() : nullptr, st.as<int>());
    BH_HIP(hipGetLastError());
    int h = 0;
    A->ctx->download(&h, st.p, sizeof(int));
    BH_CHECK(h == 0, "block inversion: singular diagonal block in row " + std::to_string(h - 1));
}

// Inverse of one dense n x n block per workgroup (n <= 128), in place in LDS
// (gj_invert_inplace): blocks larger than invert_blocks_kernel's augmented [A | I] in
// LDS allows.
__global__ __launch_bounds__(256) void invert_inplace_k(int nblocks, int n, const double* __restrict__ blocks,
                                                        double* __restrict__ inv64, float* __restrict__ inv32,
                                                        int* __restrict__ status) {
    extern __shared__ double M[];   // n * n, then n factors
    __shared__ GjScratch gs;
    const int b = blockIdx.x, t = threadIdx.x, nt = blockDim.x;
    if (b >= nblocks) return;
    const double* src = blocks + (size_t)b * n * n;
    for (int e = t; e < n * n; e += nt) M[e] = src[e];
    __syncthreads();
    const bool singular = gj_invert_inplace(M, M + (size_t)n * n, n, gs);
    if (singular && t == 0) status[0] = b + 1;
    for (int e = t; e < n * n; e += nt) {
        if (inv64) inv64[(size_t)b * n * n + e] = M[e];
        if (inv32) inv32[(size_t)b * n * n + e] = (float)M[e];
    }
}

// Inverses of `nblocks` dense column-major n x n blocks: n <= 64 by the diagonal-block
// kernel on [A | I], 64 < n <= 128 in place
void invert_dense_blocks(bh_ctx* ctx, int nblocks, int n, const double* blocks, double* inv64, float* inv32) {
    BH_CHECK(n <= 128, "dense block inversion supports n <= 128");
    if (n > 64) {
        DevBuf st(sizeof(int));
        BH_HIP(hipMemsetAsync(st.p, 0, sizeof(int), ctx->stream));
        const size_t shm = ((size_t)n * n + n) * sizeof(double);
        BH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(invert_inplace_k), hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)shm));
        hipLaunchKernelGGL(invert_inplace_k, dim3(nblocks), dim3(256), shm, ctx->stream, nblocks, n, blocks, inv64, inv32,
                           st.as<int>());
        BH_HIP(hipGetLastError());
        int h = 0;
        ctx->download(&h, st.p, sizeof(int));
        BH_CHECK(h == 0, "dense block inversion: singular block " + std::to_string(h - 1));
        return;
    }
    std::vector<int64_t> iota(nblocks);
    for (int i = 0; i < nblocks; ++i) iota[i] = i;
    DevBuf idx(nblocks * sizeof(int64_t)), st(sizeof(int));
    ctx->upload(idx.p, iota.data(), iota.size() * sizeof(int64_t));
    BH_HIP(hipMemsetAsync(st.p, 0, sizeof(int), ctx->stream));
    hipLaunchKernelGGL(invert_blocks_kernel, dim3(nblocks), dim3(64), (size_t)n * 2 * n * sizeof(double), ctx->stream, nblocks,
                       n, idx.as<int64_t>(), blocks, inv64, inv32, st.as<int>());
    BH_HIP(hipGetLastError());
    int h = 0;
    ctx->download(&h, st.p, sizeof(int));
    BH_CHECK(h == 0, "dense block inversion: singular block " + std::to_string(h - 1));
}

// Gauss-Jordan inverse with partial pivoting of one dense row-major n x n matrix
// (the coarsest multigrid level, n up to a few thousand): per column c one pivot
// search, one row swap + scale of the pivot row (both single-workgroup), and one
// grid-wide rank-1 elimination of every other row of [A | Inv].  The host never
// waits between columns; a zero pivot is flagged and reported at the end.
__global__ __launch_bounds__(1024) void gj_pivot_k(const double* __restrict__ A, int n, int c, int* __restrict__ st) {
    __shared__ double bv[1024];
    __shared__ int bi[1024];
    double best = -1.0;
    int bidx = n;
    for (int r = c + (int)threadIdx.x; r < n; r += blockDim.x) {
        const double v = fabs(A[(size_t)r * n + c]);
        if (v > best) { best = v; bidx = r; }   // rows ascend per thread: the first maximum wins
    }
    bv[threadIdx.x] = best;
    bi[threadIdx.x] = bidx;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const double o = bv[threadIdx.x + off];
            const int oi = bi[threadIdx.x + off];
            if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
                bv[threadIdx.x] = o;
                bi[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        st[1] = bi[0];
        if (!(bv[0] > 0.0) && st[0] == 0) st[0] = c + 1;
    }
}

// swap rows c and p of A and Inv, scale row c by 1 / pivot, keep column c (the
// elimination factors) in fcol with a zero at row c
__global__ __launch_bounds__(1024) void gj_swap_scale_k(double* __restrict__ A, double* __restrict__ Inv, int n, int c,
                                                        const int* __restrict__ st, double* __restrict__ fcol) {
    const int p = st[1];
    const double pv = A[(size_t)p * n + c], cv = A[(size_t)c * n + c];
    const double rp = st[0] ? 0.0 : 1.0 / pv;
    // factors first (rows other than c and p are not written by this kernel; row p receives row c)
    for (int r = threadIdx.x; r < n; r += blockDim.x) fcol[r] = r == c ? 0.0 : (r == p ? cv : A[(size_t)r * n + c]);
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * n; k += blockDim.x) {
        double* M = k < n ? A : Inv;
        const int kk = k < n ? k : k - n;
        const double vp = M[(size_t)p * n + kk], vc = M[(size_t)c * n + kk];
        M[(size_t)c * n + kk] = vp * rp;
        if (p != c) M[(size_t)p * n + kk] = vc;
    }
}

// [A | Inv][r][:] -= fcol[r] * [A | Inv][c][:] for r != c (row c is read-only here)
__global__ __launch_bounds__(256) void gj_eliminate_k(double* __restrict__ A, double* __restrict__ Inv, int n, int c,
                                                      const double* __restrict__ fcol) {
    const int r = blockIdx.y;
    const double f = fcol[r];
    if (f == 0.0) return;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < 2 * n; k += gridDim.x * blockDim.x) {
        double* M = k < n ? A : Inv;
        const int kk = k < n ? k : k - n;
        M[(size_t)r * n + kk] -= f * M[(size_t)c * n + kk];
    }
}

__global__ void gj_identity_k(double* __restrict__ I, int n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int64_t)n * n) I[i] = (i / n == i % n) ? 1.0 : 0.0;
}

void dense_inverse_dev(bh_ctx* ctx, double* A, double* Inv, int n) {
    BH_CHECK(n >= 1 && n <= 16384, "dense inverse: n out of range");
    DevBuf st(2 * sizeof(int)), fcol((size_t)n * sizeof(double));
    const int zero2[2] = {0, 0};
    ctx->upload(st.p, zero2, sizeof(zero2));
    const int64_t nn = (int64_t)n * n;
    hipLaunchKernelGGL(gj_identity_k, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, ctx->stream, Inv, n);
    BH_HIP(hipGetLastError());
    const dim3 eg((unsigned)std::min<int64_t>((2 * (int64_t)n + 255) / 256, 64), (unsigned)n);
    for (int c = 0; c < n; ++c) {
        hipLaunchKernelGGL(gj_pivot_k, dim3(1), dim3(1024), 0, ctx->stream, A, n, c, st.as<int>());
        hipLaunchKernelGGL(gj_swap_scale_k, dim3(1), dim3(1024), 0, ctx->stream, A, Inv, n, c, st.as<int>(), fcol.as<double>());
        hipLaunchKernelGGL(gj_eliminate_k, eg, dim3(256), 0, ctx->stream, A, Inv, n, c, fcol.as<double>());
    }
    BH_HIP(hipGetLastError());
    int h[2] = {0, 0};
    ctx->download(h, st.p, sizeof(h));
    BH_CHECK(h[0] == 0, "multigrid: singular coarsest-level operator (zero pivot in column " + std::to_string(h[0] - 1) + ")");
}

// ------------------------------------------------------------ block Jacobi
// One fused sweep: xout = xin + omega * Dinv (b - A xin).  Algebraically the
// reference's update x = omega*Dinv*(b - A x + D x) + (1-omega) x
// (AS/BlockJacobi.cs:196-255) with one block GEMV instead of two.
// Phase 1 (same wave layout as the SpMV) forms the block-row residual in LDS,
// phase 2 multiplies by the row's Dinv block (column-major, coalesced).
template <typename T, int BS>
__global__ __launch_bounds__(256) void bjacobi_sweep_kernel(int nrows, const int64_t* __restrict__ rowptr,
                                                            const int* __restrict__ col, const T* __restrict__ vals,
                                                            const T* __restrict__ dinv, const T* __restrict__ b,
                                                            const T* __restrict__ xin, T* __restrict__ xout, T omega) {
    constexpr int G = 64 / BS;
    __shared__ T red[4][64];
    __shared__ T res[4][BS];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    const bool active = r < nrows;
    const int g = lane / BS, ri = lane - g * BS;
    T acc = 0;
    if (active && g < G) {
        const int64_t p0 = rowptr[r], p1 = rowptr[r + 1];
        const int ncol = (int)(p1 - p0) * BS;
        const T* __restrict__ base = vals + (size_t)p0 * BS * BS;
#pragma unroll 4
        for (int q = g; q < ncol; q += G) {
            const int bb = q / BS, cc = q - bb * BS;
            acc += base[(size_t)q * BS + ri] * xin[(size_t)col[p0 + bb] * BS + cc];
        }
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (active && g == 0) {
        T s = acc;
#pragma unroll
        for (int gg = 1; gg < G; ++gg) s += red[wave][gg * BS + ri];
        res[wave][ri] = b[(size_t)r * BS + ri] - s;
    }
    __syncthreads();
    // phase 2: (Dinv q)_ri = sum_c Dinv[c*BS + ri] q_c, columns split over groups
    T acc2 = 0;
    if (active && g < G) {
        const T* __restrict__ Di = dinv + (size_t)r * BS * BS;
#pragma unroll 4
        for (int cc = g; cc < BS; cc += G) acc2 += Di[cc * BS + ri] * res[wave][cc];
    }
    __syncthreads();
    red[wave][lane] = acc2;
    __syncthreads();
    if (active && g == 0) {
        T s = acc2;
#pragma unroll
        for (int gg = 1; gg < G; ++gg) s += red[wave][gg * BS + ri];
        const size_t o = (size_t)r * BS + ri;
        xout[o] = xin[o] + omega * s;
    }
}

void bjacobi_sweep(bh_bsr* A, const void* dinv, double omega, const void* b, const void* xin, void* xout, int prec) {
    dim3 blk(256), grd((A->nrows + 3) / 4);
    hipStream_t st = A->ctx->stream;
    const int64_t* rp = A->rowptr.as<int64_t>();
    const int* cl = A->col.as<int>();
    if (prec == BH_FP64) {
#define BH_CASE(B) case B: hipLaunchKernelGGL((bjacobi_sweep_kernel<double, B>), grd, blk, 0, st, A->nrows, rp, cl, A->vals.as<double>(), (const double*)dinv, (const double*)b, (const double*)xin, (double*)xout, omega); break;
        switch (A->bs) {
            BH_CASE(1) BH_CASE(2) BH_CASE(3) BH_CASE(4) BH_CASE(5) BH_CASE(6) BH_CASE(8) BH_CASE(10)
            BH_CASE(15) BH_CASE(16) BH_CASE(20) BH_CASE(21) BH_CASE(28) BH_CASE(32) BH_CASE(35) BH_CASE(50) BH_CASE(56)
            default: throw Error(-3, "bjacobi: unsupported block size");
        }
#undef BH_CASE
    } else {
        A->ensure_fp32();
#define BH_CASE(B) case B: hipLaunchKernelGGL((bjacobi_sweep_kernel<float, B>), grd, blk, 0, st, A->nrows, rp, cl, A->vals32.as<float>(), (const float*)dinv, (const float*)b, (const float*)xin, (float*)xout, (float)omega); break;
        switch (A->bs) {
            BH_CASE(1) BH_CASE(2) BH_CASE(3) BH_CASE(4) BH_CASE(5) BH_CASE(6) BH_CASE(8) BH_CASE(10)
            BH_CASE(15) BH_CASE(16) BH_CASE(20) BH_CASE(21) BH_CASE(28) BH_CASE(32) BH_CASE(35) BH_CASE(50) BH_CASE(56)
            default: throw Error(-3, "bjacobi: unsupported block size");
        }
#undef BH_CASE
    }
    BH_HIP(hipGetLastError());
}

// ------------------------------------------------------------ vectors
__global__ void cvt_d2f(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (float)s[i];
}
__global__ void cvt_f2d(const float* __restrict__ s, double* __restrict__ d, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (double)s[i];
}
template <typename T>
__global__ void fill_k(T* __restrict__ x, T v, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) x[i] = v;
}
template <typename T>
__global__ void axpby_k(T a, const T* __restrict__ x, T b, T* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (b == T(0) ? T(0) : b * y[i]) + a * x[i];
}

static dim3 grid_for(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void convert_f64_f32(bh_ctx* ctx, const double* s, float* d, int64_t n) {
    hipLaunchKernelGGL(cvt_d2f, grid_for(n), dim3(256), 0, ctx->stream, s, d, n);
}
void convert_f32_f64(bh_ctx* ctx, const float* s, double* d, int64_t n) {
    hipLaunchKernelGGL(cvt_f2d, grid_for(n), dim3(256), 0, ctx->stream, s, d, n);
}
// x_i uniform in [0, 1) from a counter-based hash of (seed, i) (splitmix64 finaliser)
__global__ void fill_random_k(double* __restrict__ x, int64_t n, uint64_t seed) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = seed + (uint64_t)i * 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        x[i] = (double)(z >> 11) * (1.0 / 9007199254740992.0);
    }
}
void fill_random(bh_ctx* ctx, double* x, int64_t n, uint64_t seed) {
    hipLaunchKernelGGL(fill_random_k, grid_for(n), dim3(256), 0, ctx->stream, x, n, seed);
    BH_HIP(hipGetLastError());
}

void fill(bh_ctx* ctx, double* x, double v, int64_t n) {
    hipLaunchKernelGGL((fill_k<double>), grid_for(n), dim3(256), 0, ctx->stream, x, v, n);
}
void fill32(bh_ctx* ctx, float* x, float v, int64_t n) {
    hipLaunchKernelGGL((fill_k<float>), grid_for(n), dim3(256), 0, ctx->stream, x, v, n);
}
void copy(bh_ctx* ctx, double* dst, const double* src, int64_t n) {
    BH_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, ctx->stream));
}

// plain device copy in 16-byte units (+ byte tail): a kernel of known traffic
__global__ __launch_bounds__(256) void copy16_k(const uint4* __restrict__ s, uint4* __restrict__ d, int64_t n16,
                                                const unsigned char* __restrict__ st, unsigned char* __restrict__ dt,
                                                int tail) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) d[i] = s[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < tail) dt[threadIdx.x] = st[threadIdx.x];
}

void copy_bytes(bh_ctx* ctx, void* dst, const void* src, size_t bytes) {
    BH_CHECK(((uintptr_t)dst & 15) == 0 && ((uintptr_t)src & 15) == 0, "device copy: pointers must be 16-byte aligned");
    const int64_t n16 = (int64_t)(bytes / 16);
    const int tail = (int)(bytes % 16);
    // one 16-byte unit per thread (no grid-stride loop below 2^31 blocks): measured
    // fastest on MI355X, 6.2 TB/s for 1 GiB each way against 4.5-5.3 TB/s for
    // grid-stride loops of 2048-32768 workgroups
    int64_t blocks = (n16 + 255) / 256;
    if (blocks > 0x7fffffff) blocks = 0x7fffffff;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(copy16_k, dim3((unsigned)blocks), dim3(256), 0, ctx->stream, (const uint4*)src, (uint4*)dst, n16,
                       (const unsigned char*)src + n16 * 16, (unsigned char*)dst + n16 * 16, tail);
    BH_HIP(hipGetLastError());
}
void axpby(bh_ctx* ctx, double a, const double* x, double b, double* y, int64_t n) {
    hipLaunchKernelGGL((axpby_k<double>), grid_for(n), dim3(256), 0, ctx->stream, a, x, b, y, n);
}
void axpby32(bh_ctx* ctx, float a, const float* x, float b, float* y, int64_t n) {
    hipLaunchKernelGGL((axpby_k<float>), grid_for(n), dim3(256), 0, ctx->stream, a, x, b, y, n);
}

// ------------------------------------------------------------ reductions
// Multi-dot h[k] = <w, V_k>, k < M (M <= 16 per launch, a template parameter):
// each thread keeps M partial sums over 16-byte vector loads, a wave64
// butterfly (__shfl_xor) and LDS combine them per workgroup, and a second
// kernel sums the per-workgroup partials in a fixed order (deterministic, no
// atomics).  (A last-workgroup-finishes variant in one launch needs a
// device-scope fence per workgroup, which costs more than the second launch.)
constexpr int MD_BLOCKS = 1024;
constexpr int MD_M = 16;

template <typename T> struct Vec16;
template <> struct Vec16<float> { typedef float __attribute__((ext_vector_type(4))) type; static constexpr int L = 4; };
template <> struct Vec16<double> { typedef double __attribute__((ext_vector_type(2))) type; static constexpr int L = 2; };

// T: vector precision; partial sums are accumulated in double.  VEC: every row
// of V and w 16-byte aligned and n a multiple of the vector width.
template <typename T, int M, bool VEC>
__global__ __launch_bounds__(256) void multidot_k(const T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int64_t n,
                                                  double* __restrict__ part) {
    double acc[M];
#pragma unroll
    for (int k = 0; k < M; ++k) acc[k] = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if constexpr (VEC) {
        using VT = typename Vec16<T>::type;
        constexpr int L = Vec16<T>::L;
        for (int64_t i = t0; i < n / L; i += stride) {
            VT wv{};
            if (w) wv = reinterpret_cast<const VT*>(w)[i];
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const VT v = reinterpret_cast<const VT*>(V + (size_t)k * ldv)[i];
#pragma unroll
                for (int e = 0; e < L; ++e) acc[k] += (double)(w ? wv[e] : v[e]) * (double)v[e];   // w == nullptr: |V_k|^2
            }
        }
    } else {
        for (int64_t i = t0; i < n; i += stride) {
            const double wi = w ? (double)w[i] : 0.0;
#pragma unroll
            for (int k = 0; k < M; ++k) {
                const double v = (double)V[(size_t)k * ldv + i];
                acc[k] += (w ? wi : v) * v;
            }
        }
    }
    __shared__ double sm[4][M];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < M; ++k) {
        double v = acc[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) sm[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < M) {
        const int k = threadIdx.x;
        part[(size_t)blockIdx.x * M + k] = sm[0][k] + sm[1][k] + sm[2][k] + sm[3][k];
    }
}

// out[k] = sum_b part[b][k], b < nb, in a fixed order (one workgroup per k)
__global__ __launch_bounds__(256) void reduce_parts_k(const double* __restrict__ part, int nb, int M, double* __restrict__ out) {
    const int k = blockIdx.x;
    double s = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[(size_t)i * M + k];
    __shared__ double sm[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) sm[wave] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[k] = sm[0] + sm[1] + sm[2] + sm[3];
}

static void ensure_red(bh_ctx* ctx) {
    if (!ctx->red.p) ctx->red.alloc((size_t)MD_BLOCKS * 32 * sizeof(double));   // multi-dot (16) and fused passes (32)
    if (!ctx->red_out.p) ctx->red_out.alloc(256 * sizeof(double));
    if (!ctx->host_red) BH_HIP(hipHostMalloc((void**)&ctx->host_red, 256 * sizeof(double), hipHostMallocDefault));
}

template <typename T>
static bool vec16_ok(const T* w, const T* V, int64_t ldv, int m, int64_t n) {
    constexpr int L = Vec16<T>::L;
    const bool al = ((uintptr_t)V & 15) == 0 && (!w || ((uintptr_t)w & 15) == 0);
    return al && n % L == 0 && (m == 1 || ldv % L == 0);
}

template <typename T, int M>
static void md_launch(bh_ctx* ctx, dim3 g, bool vec, const T* w, const T* V, int64_t ldv, int64_t n, double* dout) {
    if (vec)
        hipLaunchKernelGGL((multidot_k<T, M, true>), g, dim3(256), 0, ctx->stream, w, V, ldv, n, ctx->red.as<double>());
    else
        hipLaunchKernelGGL((multidot_k<T, M, false>), g, dim3(256), 0, ctx->stream, w, V, ldv, n, ctx->red.as<double>());
    hipLaunchKernelGGL(reduce_parts_k, dim3(M), dim3(256), 0, ctx->stream, ctx->red.as<double>(), (int)g.x, M, dout);
}

template <typename T>
static void multi_dot_dev(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, double* dout) {
    ensure_red(ctx);
    const bool vec = vec16_ok<T>(w, V, ldv, m, n);
    const int64_t units = vec ? n / Vec16<T>::L : n;
    int64_t nb = (units + 511) / 512;       // about two units per thread before the grid is full
    if (nb > MD_BLOCKS) nb = MD_BLOCKS;
    if (nb < 1) nb = 1;
    const dim3 g((unsigned)nb);
    switch (m) {
#define BH_MD(MM) case MM: md_launch<T, MM>(ctx, g, vec, w, V, ldv, n, dout); break;
        BH_MD(1) BH_MD(2) BH_MD(3) BH_MD(4) BH_MD(5) BH_MD(6) BH_MD(7) BH_MD(8)
        BH_MD(9) BH_MD(10) BH_MD(11) BH_MD(12) BH_MD(13) BH_MD(14) BH_MD(15) BH_MD(16)
#undef BH_MD
        default: throw Error(-3, "multi_dot: 1..16 vectors per launch");
    }
    BH_HIP(hipGetLastError());
}

template <typename T>
void multi_dot_t(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, double* h) {
    ensure_red(ctx);
    BH_CHECK(m <= 256, "multi_dot: too many vectors");
    for (int k0 = 0; k0 < m; k0 += MD_M) {
        int mm = std::min(MD_M, m - k0);
        multi_dot_dev<T>(ctx, w, V + (size_t)k0 * ldv, ldv, mm, n, ctx->red_out.as<double>() + k0);
    }
    BH_HIP(hipMemcpyAsync(ctx->host_red, ctx->red_out.p, m * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < m; ++k) h[k] = ctx->host_red[k];
}

template <typename T>
double nrm2_t(bh_ctx* ctx, const T* a, int64_t n) {
    ensure_red(ctx);
    multi_dot_dev<T>(ctx, nullptr, a, 0, 1, n, ctx->red_out.as<double>());
    BH_HIP(hipMemcpyAsync(ctx->host_red, ctx->red_out.p, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    return std::sqrt(ctx->host_red[0]);
}

void multi_dot(bh_ctx* ctx, const double* w, const double* V, int64_t ldv, int m, int64_t n, double* h) {
    multi_dot_t<double>(ctx, w, V, ldv, m, n, h);
}
double dot(bh_ctx* ctx, const double* a, const double* b, int64_t n) {
    double h;
    multi_dot_t<double>(ctx, a, b, 0, 1, n, &h);
    return h;
}
double nrm2(bh_ctx* ctx, const double* a, int64_t n) { return nrm2_t<double>(ctx, a, n); }

// w <- w + sum_k c[k] V_k for m <= 32 vectors, 16-byte vector loads when VEC
template <typename T, bool VEC>
__device__ __forceinline__ void axpy_rows(T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int m, int64_t n,
                                          const double* c) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if constexpr (VEC) {
        using VT = typename Vec16<T>::type;
        constexpr int L = Vec16<T>::L;
        for (int64_t i = t0; i < n / L; i += stride) {
            const VT x = reinterpret_cast<const VT*>(w)[i];
            double s[L];
#pragma unroll
            for (int e = 0; e < L; ++e) s[e] = (double)x[e];
            for (int k = 0; k < m; ++k) {
                const VT v = reinterpret_cast<const VT*>(V + (size_t)k * ldv)[i];
#pragma unroll
                for (int e = 0; e < L; ++e) s[e] += c[k] * (double)v[e];
            }
            VT y;
#pragma unroll
            for (int e = 0; e < L; ++e) y[e] = (T)s[e];
            reinterpret_cast<VT*>(w)[i] = y;
        }
    } else {
        for (int64_t i = t0; i < n; i += stride) {
            double s = (double)w[i];
            for (int k = 0; k < m; ++k) s += c[k] * (double)V[(size_t)k * ldv + i];
            w[i] = (T)s;
        }
    }
}

// w <- w + sign * sum_k h[k] V_k   (coefficients passed by value, m <= 32 per launch)
struct Coefs { double c[32]; };
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void multi_axpy_k(T* __restrict__ w, const T* __restrict__ V, int64_t ldv,
                                                    int m, int64_t n, Coefs h) {
    __shared__ double cs[32];
    if (threadIdx.x < 32) cs[threadIdx.x] = h.c[threadIdx.x];
    __syncthreads();
    axpy_rows<T, VEC>(w, V, ldv, m, n, cs);
}

static dim3 grid_vec(int64_t units) {
    int64_t b = (units + 255) / 256;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

template <typename T>
void multi_axpy_t(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* h, double sign) {
    for (int k0 = 0; k0 < m; k0 += 32) {
        Coefs c{};
        int mm = std::min(32, m - k0);
        for (int k = 0; k < mm; ++k) c.c[k] = sign * h[k0 + k];
        const T* Vk = V + (size_t)k0 * ldv;
        if (vec16_ok<T>(w, Vk, ldv, mm, n))
            hipLaunchKernelGGL((multi_axpy_k<T, true>), grid_vec(n / Vec16<T>::L), dim3(256), 0, ctx->stream, w, Vk, ldv, mm, n, c);
        else
            hipLaunchKernelGGL((multi_axpy_k<T, false>), grid_vec(n), dim3(256), 0, ctx->stream, w, Vk, ldv, mm, n, c);
    }
    BH_HIP(hipGetLastError());
}

void multi_axpy(bh_ctx* ctx, double* w, const double* V, int64_t ldv, int m, int64_t n, const double* h, double sign) {
    multi_axpy_t<double>(ctx, w, V, ldv, m, n, h, sign);
}

// ------------------------------------------------ fused Gram-Schmidt passes (FP64)
// w <- scale * (w - sum_k c_k V_k) for M <= 31 vectors (M a template parameter, so
// the M loads of an element issue together), each element's V values held in
// registers; DOT: the same pass also forms the partial sums of <V_k, w_new>
// (k < M) and <w_new, w_new> (column M), reduced by reduce_parts_k.  One read of
// V where a projection followed by a multi-dot needs two.
template <bool DOT, int M>
__global__ __launch_bounds__(256) void axpy_dot_k(double* __restrict__ w, const double* __restrict__ V, int64_t ldv, int64_t n,
                                                  Coefs c, double scale, double* __restrict__ part) {
    double acc[DOT ? M + 1 : 1];
#pragma unroll
    for (int k = 0; k < (DOT ? M + 1 : 1); ++k) acc[k] = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += stride) {
        double v[M];
#pragma unroll
        for (int k = 0; k < M; ++k) v[k] = V[(size_t)k * ldv + i];
        double s = w[i];
#pragma unroll
        for (int k = 0; k < M; ++k) s -= c.c[k] * v[k];
        s *= scale;
        w[i] = s;
        if constexpr (DOT) {
#pragma unroll
            for (int k = 0; k < M; ++k) acc[k] += v[k] * s;
            acc[M] += s * s;
        }
    }
    if constexpr (DOT) {
        __shared__ double sm[4][M + 1];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k <= M; ++k) {
            double x = acc[k];
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if (lane == 0) sm[wave][k] = x;
        }
        __syncthreads();
        if ((int)threadIdx.x <= M) {
            const int k = threadIdx.x;
            part[(size_t)blockIdx.x * (M + 1) + k] = sm[0][k] + sm[1][k] + sm[2][k] + sm[3][k];
        }
    }
}

template <int M>
static void axpy_dot_launch(bh_ctx* ctx, bool dot, double* w, const double* V, int64_t ldv, int64_t n, const Coefs& c,
                            double scale, int nb) {
    if (dot) {
        hipLaunchKernelGGL((axpy_dot_k<true, M>), dim3(nb), dim3(256), 0, ctx->stream, w, V, ldv, n, c, scale,
                           ctx->red.as<double>());
        hipLaunchKernelGGL(reduce_parts_k, dim3(M + 1), dim3(256), 0, ctx->stream, ctx->red.as<double>(), nb, M + 1,
                           ctx->red_out.as<double>());
    } else {
        hipLaunchKernelGGL((axpy_dot_k<false, M>), dim3(nb), dim3(256), 0, ctx->stream, w, V, ldv, n, c, scale,
                           ctx->red.as<double>());
    }
}

void axpy_scale_dots(bh_ctx* ctx, double* w, const double* V, int64_t ldv, int m, int64_t n, const double* c, double scale,
                     double* dots, double* wn2) {
    BH_CHECK(m >= 1 && m <= 31, "axpy_scale_dots: 1..31 vectors");
    ensure_red(ctx);
    Coefs cf{};
    for (int k = 0; k < m; ++k) cf.c[k] = c[k];
    int64_t nb = (n + 1023) / 1024;
    if (nb > MD_BLOCKS) nb = MD_BLOCKS;
    if (nb < 1) nb = 1;
    const bool dot = dots != nullptr;
    switch (m) {
#define BH_AD(MM) case MM: axpy_dot_launch<MM>(ctx, dot, w, V, ldv, n, cf, scale, (int)nb); break;
        BH_AD(1) BH_AD(2) BH_AD(3) BH_AD(4) BH_AD(5) BH_AD(6) BH_AD(7) BH_AD(8) BH_AD(9) BH_AD(10) BH_AD(11) BH_AD(12)
        BH_AD(13) BH_AD(14) BH_AD(15) BH_AD(16) BH_AD(17) BH_AD(18) BH_AD(19) BH_AD(20) BH_AD(21) BH_AD(22) BH_AD(23)
        BH_AD(24) BH_AD(25) BH_AD(26) BH_AD(27) BH_AD(28) BH_AD(29) BH_AD(30) BH_AD(31)
#undef BH_AD
    }
    BH_HIP(hipGetLastError());
    if (!dot) return;
    BH_HIP(hipMemcpyAsync(ctx->host_red, ctx->red_out.p, (m + 1) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < m; ++k) dots[k] = ctx->host_red[k];
    *wn2 = ctx->host_red[m];
}

// ------------------------------------------------ K-cycle Gram-Schmidt passes (device coefficients)
// Classical Gram-Schmidt twice with the second projection's dots, the norm and the
// normalisation fused (M <= 16 vectors, coefficients read from device memory, so the
// passes can be captured into a graph):
//   kc_axpy_dot:    w <- w - sum_k c_k V_k, partials of <V_k, w> (k < M) and |w|^2
//   kc_axpy_scale:  out <- (w - sum_k c_k V_k) / sqrt(nn), nn = n2 - sum_k c_k^2 (Pythagoras),
//                   nn also stored (workgroup 0) for the Hessenberg column
template <typename T, int M>
__global__ __launch_bounds__(256) void kc_axpy_dot_k(T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int64_t n,
                                                     const double* __restrict__ c, double* __restrict__ part) {
    __shared__ double cs[M];
    __shared__ double sm[4][M + 1];
    if ((int)threadIdx.x < M) cs[threadIdx.x] = c[threadIdx.x];
    __syncthreads();
    double acc[M + 1];
#pragma unroll
    for (int k = 0; k <= M; ++k) acc[k] = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v[M];
#pragma unroll
        for (int k = 0; k < M; ++k) v[k] = (double)V[(size_t)k * ldv + i];
        double s = (double)w[i];
#pragma unroll
        for (int k = 0; k < M; ++k) s -= cs[k] * v[k];
        const T st = (T)s;
        w[i] = st;
        const double ws = (double)st;
#pragma unroll
        for (int k = 0; k < M; ++k) acc[k] += v[k] * ws;
        acc[M] += ws * ws;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k <= M; ++k) {
        double x = acc[k];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if (lane == 0) sm[wave][k] = x;
    }
    __syncthreads();
    if ((int)threadIdx.x <= M) {
        const int k = threadIdx.x;
        part[(size_t)blockIdx.x * (M + 1) + k] = sm[0][k] + sm[1][k] + sm[2][k] + sm[3][k];
    }
}

template <typename T, int M>
__global__ __launch_bounds__(256) void kc_axpy_scale_k(const T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int64_t n,
                                                       const double* __restrict__ c, const double* __restrict__ n2,
                                                       double* __restrict__ nnout, T* __restrict__ out) {
    __shared__ double cs[M];
    __shared__ double sc;
    if ((int)threadIdx.x < M) cs[threadIdx.x] = c[threadIdx.x];
    __syncthreads();
    if (threadIdx.x == 0) {
        double p = 0;
        for (int k = 0; k < M; ++k) p += cs[k] * cs[k];
        const double nn = n2[0] - p;
        if (blockIdx.x == 0) nnout[0] = nn;
        const double nv = sqrt(fmax(nn, 0.0));
        sc = (nv > 1e-300 && nv < 1e300) ? 1.0 / nv : 0.0;
    }
    __syncthreads();
    const double f = sc;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double s = (double)w[i];
#pragma unroll
        for (int k = 0; k < M; ++k) s -= cs[k] * (double)V[(size_t)k * ldv + i];
        out[i] = (T)(s * f);
    }
}

template <typename T, int M>
static void kc_launch(bh_ctx* ctx, bool dots, T* w, const T* V, int64_t ldv, int64_t n, const double* c, double* res,
                      const double* n2, double* nnout, T* out) {
    int64_t nb = (n + 1023) / 1024;
    if (nb > MD_BLOCKS) nb = MD_BLOCKS;
    if (nb < 1) nb = 1;
    if (dots) {
        hipLaunchKernelGGL((kc_axpy_dot_k<T, M>), dim3((unsigned)nb), dim3(256), 0, ctx->stream, w, V, ldv, n, c,
                           ctx->red.as<double>());
        hipLaunchKernelGGL(reduce_parts_k, dim3(M + 1), dim3(256), 0, ctx->stream, ctx->red.as<double>(), (int)nb, M + 1, res);
    } else {
        hipLaunchKernelGGL((kc_axpy_scale_k<T, M>), dim3((unsigned)nb), dim3(256), 0, ctx->stream, w, V, ldv, n, c, n2, nnout,
                           out);
    }
}

template <typename T>
static void kc_dispatch(bh_ctx* ctx, bool dots, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* c, double* res,
                        const double* n2, double* nnout, T* out) {
    ensure_red(ctx);
    switch (m) {
#define BH_KC(MM) case MM: kc_launch<T, MM>(ctx, dots, w, V, ldv, n, c, res, n2, nnout, out); break;
        BH_KC(1) BH_KC(2) BH_KC(3) BH_KC(4) BH_KC(5) BH_KC(6) BH_KC(7) BH_KC(8)
        BH_KC(9) BH_KC(10) BH_KC(11) BH_KC(12) BH_KC(13) BH_KC(14) BH_KC(15) BH_KC(16)
#undef BH_KC
        default: throw Error(-3, "K-cycle Gram-Schmidt: 1..16 vectors");
    }
    BH_HIP(hipGetLastError());
}

template <typename T>
void kc_project_dots(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* c, double* res) {
    kc_dispatch<T>(ctx, true, w, V, ldv, m, n, c, res, nullptr, nullptr, nullptr);
}
template <typename T>
void kc_project_scale(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, const double* c, const double* n2,
                      double* nnout, T* out) {
    kc_dispatch<T>(ctx, false, const_cast<T*>(w), V, ldv, m, n, c, nullptr, n2, nnout, out);
}
template void kc_project_dots<float>(bh_ctx*, float*, const float*, int64_t, int, int64_t, const double*, double*);
template void kc_project_dots<double>(bh_ctx*, double*, const double*, int64_t, int, int64_t, const double*, double*);
template void kc_project_scale<float>(bh_ctx*, const float*, const float*, int64_t, int, int64_t, const double*, const double*,
                                      double*, float*);
template void kc_project_scale<double>(bh_ctx*, const double*, const double*, int64_t, int, int64_t, const double*, const double*,
                                       double*, double*);

// ------------------------------------------------ device-resident variants
// The coefficients stay in device memory: no host round trip, so a sequence
// of them can be captured into a HIP graph.

// dout[k] = <w, V_k> (w == nullptr: |V_k|^2), k < m
template <typename T>
void multi_dot_to_dev(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, double* dout) {
    BH_CHECK(m <= 256, "multi_dot: too many vectors");
    for (int k0 = 0; k0 < m; k0 += MD_M)
        multi_dot_dev<T>(ctx, w, V + (size_t)k0 * ldv, ldv, std::min(MD_M, m - k0), n, dout + k0);
}

// w <- w + sign * scale * sum_k c[k] V_k with c (and scale, if given) read from device memory
template <typename T, bool VEC>
__global__ __launch_bounds__(256) void multi_axpy_dc_k(T* __restrict__ w, const T* __restrict__ V, int64_t ldv, int m,
                                                       int64_t n, const double* __restrict__ c, const double* __restrict__ scale,
                                                       double sign) {
    __shared__ double cs[32];
    const double f = sign * (scale ? scale[0] : 1.0);
    if (threadIdx.x < m) cs[threadIdx.x] = f * c[threadIdx.x];
    __syncthreads();
    axpy_rows<T, VEC>(w, V, ldv, m, n, cs);
}

template <typename T>
void multi_axpy_from_dev(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* dcoef,
                         const double* dscale, double sign) {
    for (int k0 = 0; k0 < m; k0 += 32) {
        const int mm = std::min(32, m - k0);
        const T* Vk = V + (size_t)k0 * ldv;
        if (vec16_ok<T>(w, Vk, ldv, mm, n))
            hipLaunchKernelGGL((multi_axpy_dc_k<T, true>), grid_vec(n / Vec16<T>::L), dim3(256), 0, ctx->stream, w, Vk, ldv,
                               mm, n, dcoef + k0, dscale, sign);
        else
            hipLaunchKernelGGL((multi_axpy_dc_k<T, false>), grid_vec(n), dim3(256), 0, ctx->stream, w, Vk, ldv, mm, n,
                               dcoef + k0, dscale, sign);
    }
    BH_HIP(hipGetLastError());
}

template void multi_dot_t<float>(bh_ctx*, const float*, const float*, int64_t, int, int64_t, double*);
template void multi_dot_t<double>(bh_ctx*, const double*, const double*, int64_t, int, int64_t, double*);
template double nrm2_t<float>(bh_ctx*, const float*, int64_t);
template double nrm2_t<double>(bh_ctx*, const double*, int64_t);
template void multi_axpy_t<float>(bh_ctx*, float*, const float*, int64_t, int, int64_t, const double*, double);
template void multi_axpy_t<double>(bh_ctx*, double*, const double*, int64_t, int, int64_t, const double*, double);
template void multi_dot_to_dev<float>(bh_ctx*, const float*, const float*, int64_t, int, int64_t, double*);
template void multi_dot_to_dev<double>(bh_ctx*, const double*, const double*, int64_t, int, int64_t, double*);
template void multi_axpy_from_dev<float>(bh_ctx*, float*, const float*, int64_t, int, int64_t, const double*, const double*, double);
template void multi_axpy_from_dev<double>(bh_ctx*, double*, const double*, int64_t, int, int64_t, const double*, const double*, double);

}  // namespace bh

// ------------------------------------------------------------ bh_bsr members
void bh_bsr::ensure_fp32() {
    if (vals32.p) return;
    int64_t n = nnzb * bs * bs;
    vals32.alloc(n * sizeof(float));
    bh::convert_f64_f32(ctx, vals.as<double>(), vals32.as<float>(), n);
}

void bh_bsr::ensure_diag() {
    if (diag_idx.p) return;
    std::vector<int64_t> di(nrows, -1);
    for (int r = 0; r < nrows; ++r)
        for (int64_t p = h_rowptr[r]; p < h_rowptr[r + 1]; ++p)
            if (h_col[p] == r) di[r] = p;
    diag_idx.alloc(nrows * sizeof(int64_t));
    ctx->upload(diag_idx.p, di.data(), nrows * sizeof(int64_t));
}

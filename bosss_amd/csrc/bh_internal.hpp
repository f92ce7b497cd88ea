// Internal declarations of libbosss_hip.so (host C++ + HIP for gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bosss_hip.h"

namespace bh {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);

#define BH_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::bh::Error(-2, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                      __FILE__ + ":" + std::to_string(__LINE__));             \
    } while (0)

#define BH_CHECK(cond, msg)                                      \
    do {                                                         \
        if (!(cond)) throw ::bh::Error(-1, std::string(msg));    \
    } while (0)

// Device buffer owned by a context.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n) BH_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return (T*)p; }
};

}  // namespace bh

// ------------------------------------------------------------------ context
struct bh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    // small pinned host staging for reductions
    double* host_red = nullptr;
    bh::DevBuf red;        // reduction partials
    bh::DevBuf red_out;    // final reduction results
    void upload(void* dst, const void* src, size_t bytes);
    void download(void* dst, const void* src, size_t bytes);
    void sync();
};

// ------------------------------------------------------------------ grid
// Tensor-product Cartesian grid: cell widths per index and direction, the
// reference's cell numbering, boundary sides.
struct bh_grid {
    bh_ctx* ctx = nullptr;
    int D = 2;
    int n[3] = {1, 1, 1};
    std::vector<double> nodes[3];
    std::vector<double> width[3];
    int J = 0;
    // edges (own enumeration: per cell, per direction: '-' boundary face, then '+' face)
    std::vector<int> e_in, e_out, e_dir, e_side;
    std::vector<double> cj;
    bh::DevBuf d_width;     // [4][3][nmax] doubles: widths, 1/width, sqrt(width), 1/sqrt(width) per direction (padded)
    int nmax = 1;
    int cell_index(int i, int j, int k) const {
        return D == 2 ? i * n[1] + j : i + j * n[0] + k * n[0] * n[1];
    }
    void cell_ijk(int c, int* idx) const {
        if (D == 2) { idx[0] = c / n[1]; idx[1] = c % n[1]; idx[2] = 0; }
        else { idx[0] = c % n[0]; idx[1] = (c / n[0]) % n[1]; idx[2] = c / (n[0] * n[1]); }
    }
};

// ------------------------------------------------------------------ device helpers
namespace bh {
// f(std::integral_constant<int, i>) for i = 0..N-1, unrolled at compile time
template <typename F, int... I>
__device__ __forceinline__ void static_for_(F& f, std::integer_sequence<int, I...>) {
    (f(std::integral_constant<int, I>{}), ...);
}
template <int N, typename F>
__device__ __forceinline__ void static_for(F&& f) {
    static_for_(f, std::make_integer_sequence<int, N>{});
}
// exponent in direction d of mode n of the degree-k basis (order of basis_exponents:
// degree, then e_0, then e_1)
__host__ __device__ constexpr int mode_exponent(int D, int k, int n, int d) {
    int i = 0;
    for (int deg = 0; deg <= k; ++deg) {
        if (D == 2) {
            for (int a0 = 0; a0 <= deg; ++a0) {
                if (i == n) return d == 0 ? a0 : deg - a0;
                ++i;
            }
        } else {
            for (int a0 = 0; a0 <= deg; ++a0)
                for (int b = 0; b <= deg - a0; ++b) {
                    if (i == n) return d == 0 ? a0 : (d == 1 ? b : deg - a0 - b);
                    ++i;
                }
        }
    }
    return 0;
}
__host__ __device__ constexpr int binom(int n, int k) {
    int r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}
// The modal form of an axis-aligned face between cells of a tensor grid: the trace of
// mode n on a face of direction d is p_j(+-1) times the face's tangential mode, with
// j = the exponent of n in d; the modes with the same tangential exponents (a face mode)
// are addressed by their root, the mode with j = 0.
__host__ __device__ constexpr int face_root(int D, int k, int n, int d) {
    const int Np = binom(k + D, D);
    for (int m = 0; m < Np; ++m) {
        if (mode_exponent(D, k, m, d) != 0) continue;
        bool same = true;
        for (int dd = 0; dd < D; ++dd)
            if (dd != d && mode_exponent(D, k, m, dd) != mode_exponent(D, k, n, dd)) same = false;
        if (same) return m;
    }
    return -1;
}
// s_j = sqrt(j + 1/2), correctly rounded
__host__ __device__ constexpr double sqrt_half(int j) {
    return j == 0 ? 0.7071067811865476 : j == 1 ? 1.224744871391589 : j == 2 ? 1.5811388300841898
         : j == 3 ? 1.8708286933869707 : j == 4 ? 2.1213203435596424 : j == 5 ? 2.345207879911715
         : 2.5495097567963922;
}

}  // namespace bh

// ------------------------------------------------------------------ device helpers
#if defined(__HIPCC__)
// In-place Gauss-Jordan inverse with partial pivoting of an n x n matrix M in LDS
// (n <= 128), by one workgroup of up to 256 threads: per column a pivot search (ties
// to the lowest row), a row swap, the pivot row scaled and every other row
// eliminated; the recorded row swaps are undone as column swaps at the end.  F: n
// doubles of LDS.  Returns whether a pivot was zero (the matrix is then garbage).
// The storage order does not matter: the inverse of the transpose is the transpose
// of the inverse.
struct GjScratch {
    int piv[128];
    double bv[256];
    int bi[256];
};
__device__ inline bool gj_invert_inplace(double* M, double* F, int n, GjScratch& s) {
    const int t = threadIdx.x, nt = blockDim.x;
    bool singular = false;
    for (int k = 0; k < n; ++k) {
        double best = -1.0;
        int bix = k;
        for (int r = k + t; r < n; r += nt) {
            const double v = fabs(M[(size_t)r * n + k]);
            if (v > best) { best = v; bix = r; }
        }
        s.bv[t] = best;
        s.bi[t] = bix;
        __syncthreads();
        for (int s2 = nt / 2; s2 > 0; s2 >>= 1) {
            if (t < s2 && (s.bv[t + s2] > s.bv[t] || (s.bv[t + s2] == s.bv[t] && s.bi[t + s2] < s.bi[t]))) {
                s.bv[t] = s.bv[t + s2];
                s.bi[t] = s.bi[t + s2];
            }
            __syncthreads();
        }
        const int p = s.bi[0];
        if (!(s.bv[0] > 0)) singular = true;
        if (t == 0) s.piv[k] = p;
        if (p != k)
            for (int c = t; c < n; c += nt) {
                const double tmp = M[(size_t)k * n + c];
                M[(size_t)k * n + c] = M[(size_t)p * n + c];
                M[(size_t)p * n + c] = tmp;
            }
        __syncthreads();
        const double d = singular ? 0.0 : 1.0 / M[(size_t)k * n + k];
        for (int r = t; r < n; r += nt) F[r] = M[(size_t)r * n + k];
        __syncthreads();
        if (t == 0) M[(size_t)k * n + k] = 1.0;
        __syncthreads();
        for (int c = t; c < n; c += nt) M[(size_t)k * n + c] *= d;
        for (int r = t; r < n; r += nt)
            if (r != k) M[(size_t)r * n + k] = 0.0;
        __syncthreads();
        for (int e = t; e < n * n; e += nt) {
            const int r = e / n, c = e - r * n;
            if (r != k) M[e] -= F[r] * M[(size_t)k * n + c];
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; --k) {
        const int p = s.piv[k];
        if (p != k)
            for (int r = t; r < n; r += nt) {
                const double tmp = M[(size_t)r * n + k];
                M[(size_t)r * n + k] = M[(size_t)r * n + p];
                M[(size_t)r * n + p] = tmp;
            }
        __syncthreads();
    }
    return singular;
}
#endif

// ------------------------------------------------------------------ tables
// Reference-element tables of the orthonormal basis at the Gauss nodes.
namespace bh {
int basis_count(int D, int k);
void basis_exponents(int D, int k, std::vector<int>& ex);
void basis_eval(int D, int k, const double* xi, double* val, double* grad);
void gauss_rule(int n, double* x, double* w);

struct SipTables {
    int D, k, Np, nq1, nqv, nqf;
    std::vector<double> vw;     // nqv weights
    std::vector<double> G;      // [nqv][D][Np]  d psi_n / d xi_d at volume nodes
    std::vector<double> fw;     // nqf face weights
    std::vector<double> T;      // [2D][nqf][Np] psi_n at face nodes of face f = 2d+s
    std::vector<double> Nn;     // [2D][nqf][Np] d psi_n / d xi_d at face nodes
    std::vector<double> vxi;    // nqv*D reference volume nodes
    std::vector<double> fxi;    // nqf*(D-1) reference face-parameter nodes
    // [D][Np][Np] reference stiffness per direction: Kv[d][n][m] = sum_q vw[q] G[q][d][n] G[q][d][m]
    // (the volume integral of an axis-aligned box cell is sum_d (-nu 4/h_d^2) Kv[d], by the same rule)
    std::vector<double> Kv;
    void build(int D, int k, int nq1);
};
}  // namespace bh

// ------------------------------------------------------------------ BSR
namespace bh {
struct Vbsr;
}

struct bh_bsr {
    bh_ctx* ctx = nullptr;
    int nrows = 0, bs = 0;      // bs = 0: variable blocks (var)
    int64_t nnzb = 0;
    std::vector<int64_t> h_rowptr;
    std::vector<int> h_col;
    bh::DevBuf rowptr, col;     // int64 / int32
    bh::DevBuf vals;            // FP64, nnzb blocks, each column-major
    bh::DevBuf vals32;          // FP32 copy (lazy)
    bh::DevBuf diag_idx;        // int64 per row: index of the diagonal block or -1
    // variable-block matrix (bh_bsr_create_v): the blocks in compressed row order
    // (row r at the prefix sum of the sizes) and the caller's layout ("frame": row r
    // at blk_i0[r], possibly with unused slots between rows, as the XDG constant frame)
    bh::Vbsr* var = nullptr;
    int64_t fr_N = 0;           // frame length
    bool fr_gaps = false;       // frame differs from the compressed layout
    bh::DevBuf fr_fwd;          // int64 per compressed DOF: its frame index
    bh::DevBuf fr_inv;          // int64 per frame DOF: its compressed index or -1
    bh::DevBuf fr_a, fr_b;      // compressed work vectors (FP64, or FP32 when used so)
    void ensure_fp32();
    void ensure_diag();
    bh_bsr() = default;
    bh_bsr(const bh_bsr&) = delete;
    bh_bsr& operator=(const bh_bsr&) = delete;
    ~bh_bsr();
};

// ------------------------------------------------------------------ SIP operator
struct bh_op {
    bh_ctx* ctx = nullptr;
    bh_grid* grid = nullptr;
    int D = 2, k = 1, Np = 1;
    int64_t N = 0;
    double penalty_base = 1.3, penalty_deg = 1.0;
    int bc[6] = {0, 0, 0, 0, 0, 0};
    bh::SipTables tab;
    std::vector<double> h_nu;
    bh::DevBuf nu, nu32;
    bh::DevBuf tab64, tab32;    // device copies of the tables: G | T | Nn | vw | fw
    bh::DevBuf tab16;           // FP16 MFMA tables (built on first FP16-MFMA apply)
    int mfma_waves = 0, mfma_wg_per_xcd = 0;   // autotuned launch shape of the MFMA apply (0: not yet)
    float mfma_tuned_ms = 0.f;
    // FP64 apply: vector kernel or FP64-MFMA kernel, picked by timing on first use
    bh::DevBuf tab64m;          // FP64 MFMA tables
    // 0: not tuned; vector kernel 1: face rule node by node, 3: modal faces,
    // 4: tiled modal faces (f64_tile_w waves per workgroup); 2: FP64 MFMA
    int f64_path = 0;
    int f64_tile_w = 0;
    int f64_waves = 8, f64_wg_per_xcd = 64;
    float f64_ms_vector = 0.f, f64_ms_mfma = 0.f;
    // FP32 apply (multigrid levels), tuned separately on its first use: form 1, 3 or 4
    int f32_path = 0;
    int f32_tile_w = 0;
    float f32_ms = 0.f;
    size_t off_G = 0, off_T = 0, off_N = 0, off_vw = 0, off_fw = 0, off_K = 0;
    // boundary values and RHS -> affine vector
    std::vector<double> gbnd;
    std::vector<int> bnd_slot;  // per edge: first slot or -1
    bh::DevBuf rhs;             // N doubles (or empty)
    bh::DevBuf affine;          // N doubles (valid if affine_valid)
    bool affine_valid = false;
    bool has_affine = false;
    void rebuild_affine();
};

namespace bh {
std::unique_ptr<bh_op> make_sip_op(bh_grid* g, int k, double penalty_base, double penalty_deg, const double* nu,
                                   const int* bc);
}

// ------------------------------------------------------------------ solvers
struct bh_solver {
    virtual ~bh_solver() = default;
    bh_ctx* ctx = nullptr;
    int64_t N = 0;
    int iterations = 0;
    int converged = 0;
    double res0 = 0, res = 0;
    // x, b device pointers (FP64)
    virtual void solve_dev(double* x, const double* b) = 0;
    // preconditioner use: x <- M^-1 b with a zero initial guess (the
    // reference clears the output before Precond.Solve, AS/SoftGMRES.cs:456-462)
    virtual void precondition(double* x, const double* b);
};

namespace bh {
// kernels / launchers (defined in the .hip files)
// y = alpha A u + beta y (+ affine); with b0: y = b0 + alpha A u (the residual form, one pass)
// forget the process-wide tuning decision for op's shape and precision (re-arms the tuner)
void sip_forget_tuning(bh_op* op, int prec);
void sip_apply(bh_op* op, double alpha, const void* u, double beta, void* y, bool add_affine, int prec,
               const void* b0 = nullptr);
// FP32 vectors, FP16 MFMA contractions with FP32 accumulation
void sip_apply_mfma(bh_op* op, double alpha, const float* u, double beta, float* y, bool add_affine,
                    const float* b0 = nullptr);
// FP64 vectors, FP64 MFMA contractions (same results as the vector kernel up to summation order)
bool sip_apply_mfma64_supported(int D, int K, int64_t N);   // N: DOF (arrays below 4 GiB)
void sip_apply_mfma64(bh_op* op, double alpha, const double* u, double beta, double* y, const double* affine, int waves,
                      int per_xcd, const double* b0 = nullptr);
void sip_assemble(bh_op* op, bh_bsr* M);
void bsr_spmv(bh_bsr* A, double alpha, const void* x, double beta, void* y, int prec);
void bsr_invert_diag(bh_bsr* A, DevBuf& dinv64, DevBuf* dinv32);
void bjacobi_sweep(bh_bsr* A, const void* dinv, double omega, const void* b, const void* xin, void* xout, int prec);
void convert_f64_f32(bh_ctx* ctx, const double* src, float* dst, int64_t n);
void convert_f32_f64(bh_ctx* ctx, const float* src, double* dst, int64_t n);
void fill(bh_ctx* ctx, double* x, double v, int64_t n);
void fill_random(bh_ctx* ctx, double* x, int64_t n, uint64_t seed);   // uniform [0, 1)
void fill32(bh_ctx* ctx, float* x, float v, int64_t n);
void copy(bh_ctx* ctx, double* dst, const double* src, int64_t n);
void copy_bytes(bh_ctx* ctx, void* dst, const void* src, size_t bytes);
void axpby(bh_ctx* ctx, double a, const double* x, double b, double* y, int64_t n);  // y = a x + b y
void axpby32(bh_ctx* ctx, float a, const float* x, float b, float* y, int64_t n);
// h[k] = <w, V_k>, k < m; V_k = V + k*ldv ; result in host array h
void multi_dot(bh_ctx* ctx, const double* w, const double* V, int64_t ldv, int m, int64_t n, double* h);
double dot(bh_ctx* ctx, const double* a, const double* b, int64_t n);
double nrm2(bh_ctx* ctx, const double* a, int64_t n);
template <typename T> void multi_dot_t(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, double* h);
template <typename T> double nrm2_t(bh_ctx* ctx, const T* a, int64_t n);
template <typename T> void multi_axpy_t(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* h, double sign);
// device-resident coefficients (graph-capturable): dout[k] = <w, V_k> (w null: |V_k|^2);
// w += sign * (dscale ? dscale[0] : 1) * sum_k dcoef[k] V_k
template <typename T> void multi_dot_to_dev(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, double* dout);
// K-cycle Gram-Schmidt passes with coefficients in device memory (m <= 16):
//   kc_project_dots:  w <- w - sum c_k V_k;  res[k] = <V_k, w> (k < m), res[m] = |w|^2
//   kc_project_scale: out <- (w - sum c_k V_k) / sqrt(n2 - sum c_k^2), *nnout = n2 - sum c_k^2
template <typename T> void kc_project_dots(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* c, double* res);
template <typename T> void kc_project_scale(bh_ctx* ctx, const T* w, const T* V, int64_t ldv, int m, int64_t n, const double* c,
                                            const double* n2, double* nnout, T* out);
template <typename T> void multi_axpy_from_dev(bh_ctx* ctx, T* w, const T* V, int64_t ldv, int m, int64_t n, const double* dcoef,
                                               const double* dscale, double sign);
// w -= sum_k h[k] V_k  (h host array, m <= 128)
void multi_axpy(bh_ctx* ctx, double* w, const double* V, int64_t ldv, int m, int64_t n, const double* h, double sign);
// w <- scale (w - sum_{k<m} c_k V_k), m <= 31, in one pass; with dots: dots[k] = <V_k, w_new>, *wn2 = |w_new|^2 (host)
void axpy_scale_dots(bh_ctx* ctx, double* w, const double* V, int64_t ldv, int m, int64_t n, const double* c, double scale,
                     double* dots, double* wn2);
}  // namespace bh

// extern "C" wrappers: translate exceptions into error codes + last-error text
#define BH_API_BEGIN try {
#define BH_API_END                                   \
    return 0;                                        \
    }                                                \
    catch (const bh::Error& e) {                     \
        bh::set_last_error(e.what());                \
        return e.code;                               \
    }                                                \
    catch (const std::exception& e) {                \
        bh::set_last_error(e.what());                \
        return -99;                                  \
    }

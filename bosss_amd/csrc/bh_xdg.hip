// XDG two-phase SIP Poisson on the device (gfx950): element-driven assembly
// with per-cell quadrature rules, the variable-block matrix (10 / 20 DOF
// blocks), its SpMV and block-Jacobi smoothing.
//
// Forms: XLaplace_Bulk (per species, SIPLaplace with species diffusion and
// species length scales; L3-solution/BoSSS.Solution.XNSECommon/XLaplace.cs:50-200)
// and XLaplace_Interface in SIP mode (XLaplace.cs:206-346), assembled as
// XEvaluatorLinear does (X/XSpatialOperatorMk2_XEvaluators.cs:163) into a
// block matrix over (cell, species) blocks.
//
// Assembly: one wave per element.  Uncut cells and uncut faces use the tensor
// Gauss rule (27 / 9 nodes for k = 2), cut cells and cut faces their own rules
// (a few hundred nodes).  At each node the lanes evaluate the orthonormal basis
// (and gradients) of both sides into LDS, then every lane accumulates its (test,
// trial) entries of the element matrix; the element matrix is added to the global
// blocks with plain adds: the elements run in colour groups (XDG_COLOURS) within
// which no two touch the same entry, one launch per group in a fixed order, so the
// result is bitwise reproducible.  The evaluate runs over the same colour groups:
// uncut items in closed modal form (xdg_modal_item), cut items from the element
// matrices the assembly kept (xdg_emat_item), both in one launch per group
// (xdg_colour_kernel).
#include <algorithm>
#include <cmath>

#include "bh_internal.hpp"
#include "bh_xdg.hpp"
#include "bh_xdgdev.hpp"

namespace bh {

struct XArgs {
    int D, k, Np, nq1;
    int n0, n1, n2;
    const double* __restrict__ nodes1;  // [3][nmax+1] grid node coordinates
    int nstride;
    const XElem* __restrict__ elems;
    int nelem;
    const int* __restrict__ perm;       // element indices of the colour group [start, start + count)
    int start, count;
    const double* __restrict__ nx;
    const double* __restrict__ nw;
    const double* __restrict__ nn;
    const int* __restrict__ ex;         // [Np][D] exponents
    const double* __restrict__ gx;      // 1D Gauss nodes / weights (nq1)
    const double* __restrict__ gw;
    const int* __restrict__ spmask;
    const int64_t* __restrict__ rowptr;
    const int* __restrict__ col;
    const int64_t* __restrict__ voff;
    const int* __restrict__ bsz;
    double* __restrict__ vals;
    double* __restrict__ emat;          // element matrices of the cut items (see XdgDevice)
    const int64_t* __restrict__ eoff;
};

template <int D>
__device__ __forceinline__ void cell_box(const XArgs& a, int c, double* lo, double* h) {
    int idx[3];
    if (D == 2) { idx[0] = c / a.n1; idx[1] = c - idx[0] * a.n1; idx[2] = 0; }
    else { int sl = a.n0 * a.n1; idx[2] = c / sl; int r = c - idx[2] * sl; idx[1] = r / a.n0; idx[0] = r - idx[1] * a.n0; }
    lo[2] = 0;
    h[2] = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = a.nodes1[d * a.nstride + idx[d]];
        h[d] = a.nodes1[d * a.nstride + idx[d] + 1] - lo[d];
    }
}

// basis value / gradient of mode m at physical x in cell (lo, h); lane-parallel over modes.
// Degrees k <= 3: the 1D Legendre values and derivatives are held in fixed-size
// register arrays and picked by unrolled selects (a runtime index into a local
// array would place it in per-lane scratch memory).
constexpr int XMAXK = 3;
__device__ __forceinline__ double pick4(const double* v, int e) {
    return e == 0 ? v[0] : e == 1 ? v[1] : e == 2 ? v[2] : v[3];
}
template <int D>
__device__ __forceinline__ void basis_at(const XArgs& a, const double* lo, const double* h, const double* x, int m, double* val,
                                         double* grad) {
    double P[3][XMAXK + 1], dP[3][XMAXK + 1];
    double detJ = 1;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double xi = d < D ? 2.0 * (x[d] - lo[d]) / h[d] - 1.0 : 0.0;
        if (d < D) detJ *= 0.5 * h[d];
        // orthonormal Legendre p_n = sqrt(n + 1/2) P_n and derivatives, n <= 3
        const double P2 = 1.5 * xi * xi - 0.5, P3 = (2.5 * xi * xi - 1.5) * xi;
        P[d][0] = 0.7071067811865476;
        P[d][1] = 1.224744871391589 * xi;
        P[d][2] = 1.5811388300841898 * P2;
        P[d][3] = 1.8708286933869707 * P3;
        dP[d][0] = 0.0;
        dP[d][1] = 1.224744871391589;
        dP[d][2] = 1.5811388300841898 * 3.0 * xi;
        dP[d][3] = 1.8708286933869707 * (7.5 * xi * xi - 1.5);
    }
    const double sc = 1.0 / sqrt(detJ);
    const int* e = a.ex + m * D;
    double pv[3], dv[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        pv[d] = d < D ? pick4(P[d], e[d]) : 1.0;
        dv[d] = d < D ? pick4(dP[d], e[d]) : 0.0;
    }
    *val = sc * pv[0] * pv[1] * pv[2];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double g = sc * 2.0 / h[d];
#pragma unroll
        for (int dd = 0; dd < 3; ++dd) g *= (dd == d) ? dv[dd] : pv[dd];
        grad[d] = g;
    }
}

__device__ __forceinline__ int64_t find_block(const XArgs& a, int r, int c) {
    for (int64_t p = a.rowptr[r]; p < a.rowptr[r + 1]; ++p)
        if (a.col[p] == c) return p;
    return -1;
}

// position (0 / 1) of species sp inside cell c's block
__device__ __forceinline__ int spos(const XArgs& a, int c, int sp) { return (sp == 1 && (a.spmask[c] & 1)) ? 1 : 0; }

constexpr int XMAXNP = 20;

template <int D>
__global__ __launch_bounds__(64) void xdg_assemble_kernel(XArgs a) {
    __shared__ double sv[2][XMAXNP], sd[2][XMAXNP], sg[XMAXNP][3];
    __shared__ double node[4];
    const int lane = threadIdx.x;
    const int Np = a.Np;
    for (int ei = blockIdx.x; ei < a.count; ei += gridDim.x) {
        const XElem E = a.elems[a.perm[a.start + ei]];
        double loP[3], hP[3], loQ[3], hQ[3];
        cell_box<D>(a, E.cellP, loP, hP);
        cell_box<D>(a, E.cellQ, loQ, hQ);
        const int nblk = E.type == 0 ? 1 : (E.type == 2 ? 1 : 4);
        const int nent = nblk * Np * Np;
        // entries owned by this lane: t = lane + 64 j
        double acc[7] = {0, 0, 0, 0, 0, 0, 0};
        // normal (faces: +-e_dir; interface: per node)
        int nnodes;
        if (E.rn > 0) nnodes = E.rn;
        else if (E.type == 0) { nnodes = 1; for (int d = 0; d < D; ++d) nnodes *= a.nq1; }
        else { nnodes = 1; for (int d = 0; d < D - 1; ++d) nnodes *= a.nq1; }
        for (int qn = 0; qn < nnodes; ++qn) {
            // node coordinates and weight
            if (lane == 0) {
                double x[3] = {0, 0, 0}, w;
                if (E.rn > 0) {
                    for (int d = 0; d < 3; ++d) x[d] = a.nx[(size_t)(E.r0 + qn) * 3 + d];
                    w = a.nw[E.r0 + qn];
                } else if (E.type == 0) {
                    int r = qn;
                    w = 1;
                    for (int d = 0; d < D; ++d) {
                        const int id = r % a.nq1;
                        r /= a.nq1;
                        x[d] = loP[d] + 0.5 * (a.gx[id] + 1.0) * hP[d];
                        w *= a.gw[id] * 0.5 * hP[d];
                    }
                } else {
                    int r = qn;
                    w = 1;
                    for (int d = 0; d < D; ++d) {
                        if (d == E.dir) {
                            x[d] = (E.type == 2 && E.side == 0) ? loP[d] : loP[d] + hP[d];
                            continue;
                        }
                        const int id = r % a.nq1;
                        r /= a.nq1;
                        x[d] = loP[d] + 0.5 * (a.gx[id] + 1.0) * hP[d];
                        w *= a.gw[id] * 0.5 * hP[d];
                    }
                }
                node[0] = x[0]; node[1] = x[1]; node[2] = x[2]; node[3] = w;
            }
            __syncthreads();
            const double x[3] = {node[0], node[1], node[2]};
            const double w = node[3];
            double nrm[3];
            const double fsgn = E.type == 1 ? 1.0 : (E.side ? 1.0 : -1.0);
#pragma unroll
            for (int d = 0; d < 3; ++d)
                nrm[d] = E.type == 3 ? a.nn[(size_t)(E.r0 + qn) * 3 + d] : (E.type != 0 && d == E.dir ? fsgn : 0.0);
            // basis of both sides at the node
            if (lane < 2 * Np) {
                const int side = lane / Np, m = lane - side * Np;
                double v, g[3];
                basis_at<D>(a, side ? loQ : loP, side ? hQ : hP, x, m, &v, g);
                double dn = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) dn += g[d] * nrm[d];
                sv[side][m] = v;
                sd[side][m] = dn;
                if (side == 0) for (int d = 0; d < 3; ++d) sg[m][d] = d < D ? g[d] : 0.0;
            }
            __syncthreads();
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int t = lane + 64 * j;
                if (t >= nent) break;
                const int blk = t / (Np * Np), rem = t - blk * Np * Np;
                const int m = rem / Np, n = rem - m * Np;   // test m, trial n
                double val;
                if (E.type == 0) {
                    double s = 0;
                    for (int d = 0; d < D; ++d) s += sg[m][d] * sg[n][d];
                    val = -E.nuP * s;
                } else if (E.type == 2) {
                    // SIPLaplace.BoundaryEdgeForm, Dirichlet, linear part
                    val = E.nuP * sd[0][n] * sv[0][m] + E.nuP * sd[0][m] * sv[0][n] - E.nuP * E.mu * sv[0][n] * sv[0][m];
                } else {
                    const double vP_m = sv[0][m], dP_m = sd[0][m], vQ_m = sv[1][m], dQ_m = sd[1][m];
                    const double vP_n = sv[0][n], dP_n = sd[0][n], vQ_n = sv[1][n], dQ_n = sd[1][n];
                    const double nP = E.nuP, nQ = E.nuQ, mu = E.mu;
                    if (blk == 0) val = 0.5 * nP * dP_n * vP_m + 0.5 * nP * dP_m * vP_n - mu * vP_n * vP_m;        // PP
                    else if (blk == 1) val = 0.5 * nQ * dQ_n * vP_m - 0.5 * nP * dP_m * vQ_n + mu * vQ_n * vP_m;  // test P, trial Q
                    else if (blk == 2) val = -0.5 * nP * dP_n * vQ_m + 0.5 * nQ * dQ_m * vP_n + mu * vP_n * vQ_m; // test Q, trial P
                    else val = -0.5 * nQ * dQ_n * vQ_m - 0.5 * nQ * dQ_m * vQ_n - mu * vQ_n * vQ_m;              // QQ
                }
                acc[j] += w * val;
            }
            __syncthreads();
        }
        // cut items: keep the element matrix for the evaluate
        if (E.rn > 0 && a.emat) {
            double* em = a.emat + a.eoff[a.perm[a.start + ei]];
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                const int t = lane + 64 * j;
                if (t >= nent) break;
                em[t] = acc[j];
            }
        }
        // scatter into the global blocks
#pragma unroll
        for (int j = 0; j < 7; ++j) {
            const int t = lane + 64 * j;
            if (t >= nent) break;
            const int blk = t / (Np * Np), rem = t - blk * Np * Np;
            const int m = rem / Np, n = rem - m * Np;
            int rc, cc, rsp, csp;
            if (E.type == 0 || E.type == 2) { rc = cc = E.cellP; rsp = csp = E.spP; }
            else {
                const bool testQ = blk >= 2, trialQ = (blk == 1 || blk == 3);
                rc = testQ ? E.cellQ : E.cellP;
                rsp = testQ ? E.spQ : E.spP;
                cc = trialQ ? E.cellQ : E.cellP;
                csp = trialQ ? E.spQ : E.spP;
            }
            const int64_t p = find_block(a, rc, cc);
            if (p < 0) continue;
            const int R = a.bsz[rc];
            const int row = spos(a, rc, rsp) * Np + m, colj = spos(a, cc, csp) * Np + n;
            a.vals[a.voff[p] + (int64_t)colj * R + row] += acc[j];   // one element of the group per entry
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------ matrix-free evaluate
// Node qn of element E: physical point and weight (tensor Gauss rule for uncut
// items, the item's own node list for cut items).
template <int D>
__device__ __forceinline__ void element_node(const XArgs& a, const XElem& E, int qn, const double* loP, const double* hP,
                                             double* x, double* w) {
    x[0] = x[1] = x[2] = 0;
    if (E.rn > 0) {
        for (int d = 0; d < 3; ++d) x[d] = a.nx[(size_t)(E.r0 + qn) * 3 + d];
        *w = a.nw[E.r0 + qn];
        return;
    }
    int r = qn;
    double ww = 1;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if (E.type != 0 && d == E.dir) {
            x[d] = (E.type == 2 && E.side == 0) ? loP[d] : loP[d] + hP[d];
            continue;
        }
        const int id = r % a.nq1;
        r /= a.nq1;
        x[d] = loP[d] + 0.5 * (a.gx[id] + 1.0) * hP[d];
        ww *= a.gw[id] * 0.5 * hP[d];
    }
    *w = ww;
}

struct XEvalArgs {
    XArgs g;
    const int64_t* __restrict__ off;   // full DOF offset per cell
    const double* __restrict__ u;
    double* __restrict__ y;
    double alpha, beta;
};

// y_i += alpha v; the volume group, which holds exactly one item per (cell, species)
// row, applies beta first: y_i = beta y_i + alpha v (beta = 0: y is not read)
template <bool FIRST>
__device__ __forceinline__ void xdg_put(const XEvalArgs& e, int64_t i, double v) {
    if constexpr (FIRST) e.y[i] = (e.beta == 0.0 ? 0.0 : e.beta * e.y[i]) + e.alpha * v;
    else e.y[i] += e.alpha * v;
}

// Cut items of the evaluate: y += alpha M u with the element matrix M that the assembly
// kept (the same values it adds into the global blocks), one lane per (item, test row):
// a row of the test side's two blocks (trial P, trial Q) against both sides' coefficients.
// Colour groups as in the assembly: no two items of a launch touch the same rows.
template <bool TWO, bool FIRST>
__device__ __forceinline__ void xdg_emat_item(const XEvalArgs& e, int start, int count, int64_t t) {
    const XArgs& a = e.g;
    const int Np = a.Np, R = TWO ? 2 * Np : Np;
    if (t >= (int64_t)count * R) return;
    const int ei = (int)(t / R), r = (int)(t - (int64_t)ei * R);
    const int eg = a.perm[start + ei];
    if (a.eoff[eg] < 0) return;   // every cut item has a kept matrix; nothing else reads one
    const XElem E = a.elems[eg];
    const int side = TWO ? r / Np : 0, m = r - side * Np;
    const int64_t oP = e.off[E.cellP] + (int64_t)spos(a, E.cellP, E.spP) * Np;
    const int64_t oQ = TWO ? e.off[E.cellQ] + (int64_t)spos(a, E.cellQ, E.spQ) * Np : oP;
    const double* M = a.emat + a.eoff[eg] + (int64_t)(2 * side) * Np * Np + (int64_t)m * Np;
    double s = 0;
    for (int n = 0; n < Np; ++n) s += M[n] * e.u[oP + n];
    if (TWO)
        for (int n = 0; n < Np; ++n) s += M[Np * Np + n] * e.u[oQ + n];
    xdg_put<FIRST>(e, (side ? oQ : oP) + m, s);
}

// XDG mass matrix (X/MassMatrixFactory.cs:35; the block of cell j and species s is
// M_mn = int_{K_j cap species s} phi_m phi_n): one workgroup per species volume item;
// uncut items (the whole cell, orthonormal basis) give the identity, cut items sum
// w phi_m phi_n over their rule.  out: row r of the full (cell, species) layout holds
// its block row, out[r * Np + n].
template <int D>
__global__ __launch_bounds__(64) void xdg_mass_k(XEvalArgs e, const int* __restrict__ vol, int nvol, double* __restrict__ out) {
    const XArgs& a = e.g;
    const int Np = a.Np, NN = Np * Np;
    for (int v = blockIdx.x; v < nvol; v += gridDim.x) {
        const XElem E = a.elems[vol[v]];
        double lo[3], h[3];
        cell_box<D>(a, E.cellP, lo, h);
        const int64_t o = e.off[E.cellP] + (int64_t)spos(a, E.cellP, E.spP) * Np;
        for (int t = threadIdx.x; t < NN; t += blockDim.x) {
            const int m = t / Np, n = t - m * Np;
            double s = 0;
            if (E.rn == 0) {
                s = m == n ? 1.0 : 0.0;
            } else {
                for (int q = 0; q < E.rn; ++q) {
                    double x[3], w, vm, vn, g[3];
                    element_node<D>(a, E, q, lo, h, x, &w);
                    basis_at<D>(a, lo, h, x, m, &vm, g);
                    basis_at<D>(a, lo, h, x, n, &vn, g);
                    s += w * vm * vn;
                }
            }
            out[(o + m) * Np + n] = s;
        }
    }
}

void xdg_mass(XdgDevice& X, double* out_dev) {
    XdgProblem& P = *X.P;
    std::vector<int> vol;
    for (size_t i = 0; i < P.elems.size(); ++i)
        if (P.elems[i].type == 0) vol.push_back((int)i);
    if (vol.empty()) return;
    DevBuf dv(vol.size() * sizeof(int));
    X.ctx->upload(dv.p, vol.data(), vol.size() * sizeof(int));
    XEvalArgs e{};
    XArgs& a = e.g;
    a.D = P.D; a.k = P.k; a.Np = P.Np; a.nq1 = P.k + 1;
    a.n0 = P.g->n[0]; a.n1 = P.g->n[1]; a.n2 = P.g->n[2];
    a.nodes1 = X.nodes1.as<double>();
    a.nstride = X.nstride;
    a.elems = X.elems.as<XElem>();
    a.nx = X.nx.as<double>(); a.nw = X.nw.as<double>(); a.nn = X.nn.as<double>();
    a.ex = X.ex.as<int>();
    a.gx = X.gx.as<double>(); a.gw = X.gw.as<double>();
    a.spmask = X.spmask.as<int>();
    e.off = X.offf.as<int64_t>();
    BH_HIP(hipMemsetAsync(out_dev, 0, (size_t)P.N_full * P.Np * sizeof(double), X.ctx->stream));
    const int n = (int)vol.size();
    const int blocks = n < 65536 ? n : 65536;
    if (P.D == 2) hipLaunchKernelGGL(xdg_mass_k<2>, dim3(blocks), dim3(64), 0, X.ctx->stream, e, dv.as<int>(), n, out_dev);
    else hipLaunchKernelGGL(xdg_mass_k<3>, dim3(blocks), dim3(64), 0, X.ctx->stream, e, dv.as<int>(), n, out_dev);
    BH_HIP(hipGetLastError());
    X.ctx->sync();
}

// ------------------------------------------------------------------ variable-block SpMV
// One wave per block row; row size R (<= 64) is wave-uniform: lanes = G groups
// x R rows; each group walks the row's blocks column by column (blocks are
// column-major, so the G groups read G*R consecutive values).
template <typename T>
__global__ __launch_bounds__(256) void vbsr_spmv_kernel(int nrows, const int* __restrict__ rsz, const int64_t* __restrict__ roff,
                                                        const int64_t* __restrict__ rowptr, const int* __restrict__ col,
                                                        const int64_t* __restrict__ voff, const T* __restrict__ vals,
                                                        const T* __restrict__ x, T* __restrict__ y, T alpha, T beta,
                                                        const T* __restrict__ bsub) {
    __shared__ T red[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + wave;
    const bool active = r < nrows;
    const int R = active ? rsz[r] : 1;
    const int G = R > 0 ? 64 / R : 1;
    const int g = R > 0 ? lane / R : 64, ri = R > 0 ? lane - g * R : 0;
    T acc = 0;
    if (active && R > 0 && g < G) {
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int c = col[p];
            const int C = rsz[c];
            const T* __restrict__ blk = vals + voff[p];
            const T* __restrict__ xc = x + roff[c];
            for (int cc = g; cc < C; cc += G) acc += blk[(int64_t)cc * R + ri] * xc[cc];
        }
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (active && R > 0 && g == 0) {
        T s = acc;
        for (int gg = 1; gg < G; ++gg) s += red[wave][gg * R + ri];
        const int64_t o = roff[r] + ri;
        T out = alpha * s;
        if (bsub) out = bsub[o] - s;
        else if (beta != T(0)) out += beta * y[o];
        y[o] = out;
    }
}

// batched Gauss-Jordan of the (variable-size) diagonal blocks
__global__ __launch_bounds__(64) void vbsr_invert_kernel(int nrows, const int* __restrict__ rsz, const int64_t* __restrict__ diag,
                                                         const int64_t* __restrict__ voff, const int64_t* __restrict__ doff,
                                                         const double* __restrict__ vals, double* __restrict__ inv,
                                                         float* __restrict__ inv32, int* __restrict__ status) {
    extern __shared__ double M[];
    const int r = blockIdx.x, lane = threadIdx.x;
    const int bs = rsz[r];
    if (bs == 0) return;
    const int W = 2 * bs;
    const int64_t bi = diag[r];
    for (int i = 0; i < bs; ++i)
        for (int j = lane; j < W; j += 64)
            M[i * W + j] = j < bs ? (bi >= 0 ? vals[voff[bi] + (int64_t)j * bs + i] : 0.0) : ((j - bs == i) ? 1.0 : 0.0);
    __syncthreads();
    for (int c = 0; c < bs; ++c) {
        int p = c;
        double best = fabs(M[c * W + c]);
        for (int i = c + 1; i < bs; ++i) { const double v = fabs(M[i * W + c]); if (v > best) { best = v; p = i; } }
        if (best == 0.0) { if (lane == 0) status[0] = r + 1; return; }
        __syncthreads();
        if (p != c) for (int j = lane; j < W; j += 64) { const double t = M[p * W + j]; M[p * W + j] = M[c * W + j]; M[c * W + j] = t; }
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
    for (int i = 0; i < bs; ++i)
        for (int j = lane; j < bs; j += 64) {
            const double v = M[i * W + bs + j];
            inv[doff[r] + (int64_t)j * bs + i] = v;
            if (inv32) inv32[doff[r] + (int64_t)j * bs + i] = (float)v;
        }
}

// x (+)= omega * Dinv r per variable block
template <typename T, typename M>
__global__ __launch_bounds__(256) void vblock_gemv_kernel(int nrows, const int* __restrict__ rsz, const int64_t* __restrict__ roff,
                                                          const int64_t* __restrict__ doff, const M* __restrict__ Dinv,
                                                          const T* __restrict__ r, T* __restrict__ x, M omega, int add) {
    __shared__ M red[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + wave;
    const bool active = row < nrows;
    const int R = active ? rsz[row] : 1;
    const int G = R > 0 ? 64 / R : 1;
    const int g = R > 0 ? lane / R : 64, ri = R > 0 ? lane - g * R : 0;
    M acc = 0;
    if (active && R > 0 && g < G) {
        const M* Di = Dinv + doff[row];
        const T* rr = r + roff[row];
        for (int c = g; c < R; c += G) acc += Di[(int64_t)c * R + ri] * (M)rr[c];
    }
    red[wave][lane] = acc;
    __syncthreads();
    if (active && R > 0 && g == 0) {
        M s = acc;
        for (int gg = 1; gg < G; ++gg) s += red[wave][gg * R + ri];
        const int64_t o = roff[row] + ri;
        x[o] = (T)(add ? (M)x[o] + omega * s : omega * s);
    }
}

// species-merged coarse space: rc(c, m) = sum over the species blocks of cell c of r(c, s, m), m < Npc
template <typename T>
__global__ void xdg_restrict_k(int nrows, int Np, int Npc, const int* __restrict__ rsz, const int64_t* __restrict__ roff,
                               const T* __restrict__ r, T* __restrict__ rc) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrows * Npc) return;
    const int c = (int)(t / Npc), m = (int)(t - (int64_t)c * Npc);
    const int ns = rsz[c] / Np;
    T s = 0;
    for (int q = 0; q < ns; ++q) s += r[roff[c] + q * Np + m];
    rc[t] = s;
}

template <typename T>
__global__ void xdg_prolong_k(int nrows, int Np, int Npc, const int* __restrict__ rsz, const int64_t* __restrict__ roff,
                              const T* __restrict__ xc, T* __restrict__ x) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrows * Npc) return;
    const int c = (int)(t / Npc), m = (int)(t - (int64_t)c * Npc);
    const int ns = rsz[c] / Np;
    const T v = xc[t];
    for (int q = 0; q < ns; ++q) x[roff[c] + q * Np + m] += v;
}

template <typename T>
__global__ void frame_gather_k(int64_t nc, const int64_t* __restrict__ fwd, const T* __restrict__ src, T* __restrict__ dst) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nc; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[fwd[i]];
}
template <typename T>
__global__ void frame_scatter_k(int64_t nf, const int64_t* __restrict__ inv, const T* __restrict__ src, T* __restrict__ dst) {
    for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < nf; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = inv[k];
        dst[k] = i >= 0 ? src[i] : T(0);
    }
}
template <typename T>
void frame_gather(bh_ctx* ctx, int64_t nc, const int64_t* fwd, const T* src_f, T* dst_c) {
    const int64_t b = std::min<int64_t>((nc + 255) / 256, 65536);
    hipLaunchKernelGGL((frame_gather_k<T>), dim3((unsigned)std::max<int64_t>(b, 1)), dim3(256), 0, ctx->stream, nc, fwd, src_f, dst_c);
    BH_HIP(hipGetLastError());
}
template <typename T>
void frame_scatter(bh_ctx* ctx, int64_t nf, const int64_t* inv, const T* src_c, T* dst_f) {
    const int64_t b = std::min<int64_t>((nf + 255) / 256, 65536);
    hipLaunchKernelGGL((frame_scatter_k<T>), dim3((unsigned)std::max<int64_t>(b, 1)), dim3(256), 0, ctx->stream, nf, inv, src_c, dst_f);
    BH_HIP(hipGetLastError());
}
template void frame_gather<double>(bh_ctx*, int64_t, const int64_t*, const double*, double*);
template void frame_gather<float>(bh_ctx*, int64_t, const int64_t*, const float*, float*);
template void frame_scatter<double>(bh_ctx*, int64_t, const int64_t*, const double*, double*);
template void frame_scatter<float>(bh_ctx*, int64_t, const int64_t*, const float*, float*);

// ------------------------------------------------------------------ agglomeration Q^T A Q
// A contribution is one species sub-block S of a full-matrix block (Np x Np, element
// (m, n) at fv[src + n * rf + m], the full block column-major with rf rows), mapped to
// the reduced unit block of its row and column targets, with the extension blocks of
// the row / column species when they are agglomeration sources (E^T S, S E).

// flags[i] = 1 when contribution i's sub-block holds a nonzero
__global__ void qaq_flags_k(int64_t n, int Np, const int64_t* __restrict__ src, const int* __restrict__ rf,
                            const double* __restrict__ fv, unsigned char* __restrict__ flags) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* S = fv + src[i];
    const int R = rf[i];
    bool nz = false;
    for (int c = 0; c < Np && !nz; ++c)
        for (int r = 0; r < Np; ++r)
            if (S[(size_t)c * R + r] != 0.0) { nz = true; break; }
    flags[i] = nz ? 1 : 0;
}

// One workgroup (64 lanes) per reduced unit block b: the sum over its contributions
// k in [cptr[b], cptr[b+1]) of E_pr^T S E_pc (pr, pc < 0: no extension), formed in LDS
// in the order E^T S first, then the product with E; written column-major.
__global__ __launch_bounds__(64) void qaq_gather_k(int64_t nblk, int Np, const int64_t* __restrict__ cptr,
                                                   const int64_t* __restrict__ src, const int* __restrict__ rf,
                                                   const int* __restrict__ pr, const int* __restrict__ pc,
                                                   const double* __restrict__ fv, const double* __restrict__ ext,
                                                   double* __restrict__ out) {
    __shared__ double S[XDG_NPMAX * XDG_NPMAX], W[XDG_NPMAX * XDG_NPMAX], E[XDG_NPMAX * XDG_NPMAX];
    const int64_t b = blockIdx.x;
    if (b >= nblk) return;
    const int lane = threadIdx.x, NN = Np * Np;
    double acc[(XDG_NPMAX * XDG_NPMAX + 63) / 64];
#pragma unroll
    for (int e = 0; e < (XDG_NPMAX * XDG_NPMAX + 63) / 64; ++e) acc[e] = 0.0;
    for (int64_t k = cptr[b]; k < cptr[b + 1]; ++k) {
        const double* Sg = fv + src[k];
        const int R = rf[k], er = pr[k], ec = pc[k];
        for (int e = lane; e < NN; e += 64) {   // S[m][n], row-major in LDS
            const int m = e / Np, n = e - m * Np;
            S[e] = Sg[(size_t)n * R + m];
        }
        __syncthreads();
        if (er >= 0) {   // W = E^T S, E row-major [src mode][tgt mode]
            for (int e = lane; e < NN; e += 64) E[e] = ext[(size_t)er * NN + e];
            __syncthreads();
            for (int e = lane; e < NN; e += 64) {
                const int m = e / Np, n = e - m * Np;
                double s = 0;
                for (int t = 0; t < Np; ++t) s += E[t * Np + m] * S[t * Np + n];
                W[e] = s;
            }
            __syncthreads();
            for (int e = lane; e < NN; e += 64) S[e] = W[e];
            __syncthreads();
        }
        if (ec >= 0) {   // W = S E
            for (int e = lane; e < NN; e += 64) E[e] = ext[(size_t)ec * NN + e];
            __syncthreads();
            for (int e = lane; e < NN; e += 64) {
                const int m = e / Np, n = e - m * Np;
                double s = 0;
                for (int t = 0; t < Np; ++t) s += S[m * Np + t] * E[t * Np + n];
                W[e] = s;
            }
            __syncthreads();
            for (int e = lane; e < NN; e += 64) S[e] = W[e];
            __syncthreads();
        }
        int j = 0;
        for (int e = lane; e < NN; e += 64, ++j) acc[j] += S[e];
        __syncthreads();
    }
    int j = 0;
    for (int e = lane; e < NN; e += 64, ++j) {
        const int m = e / Np, n = e - m * Np;
        out[(size_t)b * NN + (size_t)n * Np + m] = acc[j];
    }
}

// cell diagonal blocks (species coupled) from the unit blocks: cell r of size R = nu * Np
// takes unit block dblk[4 r + a * 2 + c] (or zero, -1) at rows a Np.., columns c Np..
__global__ __launch_bounds__(64) void qaq_dcell_k(int J, int Np, const int* __restrict__ rsz, const int64_t* __restrict__ doff,
                                                  const int64_t* __restrict__ dblk, const double* __restrict__ uv,
                                                  double* __restrict__ dv) {
    const int r = blockIdx.x;
    if (r >= J) return;
    const int R = rsz[r];
    for (int e = threadIdx.x; e < R * R; e += 64) {
        const int j = e / R, i = e - j * R;   // column-major: entry (i, j)
        const int a = i / Np, c = j / Np;
        const int64_t bk = dblk[4 * r + a * 2 + c];
        dv[doff[r] + e] = bk >= 0 ? uv[bk * Np * Np + (size_t)(j - c * Np) * Np + (i - a * Np)] : 0.0;
    }
}

void qaq_flags(bh_ctx* ctx, int64_t n, int Np, const int64_t* src, const int* rf, const double* fv, unsigned char* flags) {
    if (n == 0) return;
    hipLaunchKernelGGL(qaq_flags_k, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, n, Np, src, rf, fv, flags);
    BH_HIP(hipGetLastError());
}
void qaq_gather(bh_ctx* ctx, int64_t nblk, int Np, const int64_t* cptr, const int64_t* src, const int* rf, const int* pr,
                const int* pc, const double* fv, const double* ext, double* out) {
    BH_CHECK(Np <= XDG_NPMAX, "agglomeration: Np too large for the device Q^T A Q");
    if (nblk == 0) return;
    hipLaunchKernelGGL(qaq_gather_k, dim3((unsigned)nblk), dim3(64), 0, ctx->stream, nblk, Np, cptr, src, rf, pr, pc, fv, ext, out);
    BH_HIP(hipGetLastError());
}
void qaq_dcell(bh_ctx* ctx, int J, int Np, const int* rsz, const int64_t* doff, const int64_t* dblk, const double* uv, double* dv) {
    if (J == 0) return;
    hipLaunchKernelGGL(qaq_dcell_k, dim3((unsigned)J), dim3(64), 0, ctx->stream, J, Np, rsz, doff, dblk, uv, dv);
    BH_HIP(hipGetLastError());
}

template <typename T>
void xdg_restrict(bh_ctx* ctx, const Vbsr& A, int Np, int Npc, const T* r, T* rc) {
    const int64_t n = (int64_t)A.nrows * Npc;
    hipLaunchKernelGGL((xdg_restrict_k<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A.nrows, Np, Npc,
                       A.rsz_d.as<int>(), A.roff_d.as<int64_t>(), r, rc);
    BH_HIP(hipGetLastError());
}

template <typename T>
void xdg_prolong_add(bh_ctx* ctx, const Vbsr& A, int Np, int Npc, const T* xc, T* x) {
    const int64_t n = (int64_t)A.nrows * Npc;
    hipLaunchKernelGGL((xdg_prolong_k<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, A.nrows, Np, Npc,
                       A.rsz_d.as<int>(), A.roff_d.as<int64_t>(), xc, x);
    BH_HIP(hipGetLastError());
}
template void xdg_restrict<float>(bh_ctx*, const Vbsr&, int, int, const float*, float*);
template void xdg_restrict<double>(bh_ctx*, const Vbsr&, int, int, const double*, double*);
template void xdg_prolong_add<float>(bh_ctx*, const Vbsr&, int, int, const float*, float*);
template void xdg_prolong_add<double>(bh_ctx*, const Vbsr&, int, int, const double*, double*);

// ------------------------------------------------------------------ host
// Coarse matrix of the XDG preconditioner on the device.  Species units: block q of
// Ac is the sum, in list order, of the leading Npc x Npc parts of the unit blocks
// p listed for q (column-major Np x Np blocks in, Npc x Npc out).
__global__ __launch_bounds__(256) void xdg_coarse_gather_k(int64_t nq, int Np, int Npc, const int64_t* __restrict__ qptr,
                                                           const int64_t* __restrict__ plist, const double* __restrict__ vals,
                                                           double* __restrict__ out) {
    const int E = Npc * Npc;
    const int64_t tot = nq * E;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t q = t / E;
        const int e = (int)(t - q * E);
        const int n = e / Npc, m = e - n * Npc;
        double s = 0.0;
        for (int64_t k = qptr[q]; k < qptr[q + 1]; ++k) s += vals[(size_t)plist[k] * Np * Np + (size_t)n * Np + m];
        out[t] = s;
    }
}

// A caller's cell blocks (R x C, column-major, R and C multiples of Np): block p of Ac
// is Ac_p(m, n) = sum over the species blocks a < R / Np, b < C / Np of A_p(a Np + m, b Np + n)
__global__ __launch_bounds__(256) void xdg_coarse_vblocks_k(int64_t nnzb, int Np, int Npc, const int* __restrict__ prow,
                                                            const int* __restrict__ col, const int* __restrict__ rsz,
                                                            const int64_t* __restrict__ voff, const double* __restrict__ vals,
                                                            double* __restrict__ out) {
    const int E = Npc * Npc;
    const int64_t tot = nnzb * E;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < tot; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = t / E;
        const int e = (int)(t - p * E);
        const int n = e / Npc, m = e - n * Npc;
        const int R = rsz[prow[p]], C = rsz[col[p]];
        const double* __restrict__ B = vals + voff[p];
        double s = 0.0;
        for (int a = 0; a < R / Np; ++a)
            for (int b = 0; b < C / Np; ++b) s += B[(size_t)(b * Np + n) * R + a * Np + m];
        out[t] = s;
    }
}

static dim3 xdg_grid1d(int64_t n) {
    int64_t b = (n + 255) / 256;
    if (b > 16384) b = 16384;
    if (b < 1) b = 1;
    return dim3((unsigned)b);
}

void xdg_coarse_gather(bh_ctx* ctx, int64_t nq, int Np, int Npc, const int64_t* qptr, const int64_t* plist, const double* vals,
                       double* out) {
    hipLaunchKernelGGL(xdg_coarse_gather_k, xdg_grid1d(nq * Npc * Npc), dim3(256), 0, ctx->stream, nq, Np, Npc, qptr, plist, vals, out);
    BH_HIP(hipGetLastError());
}

void xdg_coarse_vblocks(bh_ctx* ctx, const Vbsr& A, int Np, int Npc, const int* prow, double* out) {
    const int64_t nnzb = (int64_t)A.col.size();
    hipLaunchKernelGGL(xdg_coarse_vblocks_k, xdg_grid1d(nnzb * Npc * Npc), dim3(256), 0, ctx->stream, nnzb, Np, Npc, prow,
                       A.col_d.as<int>(), A.rsz_d.as<int>(), A.voff_d.as<int64_t>(), A.vals.as<double>(), out);
    BH_HIP(hipGetLastError());
}

void Vbsr::upload_structure() {
    rowptr_d.alloc((nrows + 1) * sizeof(int64_t));
    col_d.alloc(std::max<size_t>(col.size(), 1) * sizeof(int));
    voff_d.alloc(std::max<size_t>(voff.size(), 1) * sizeof(int64_t));
    rsz_d.alloc(std::max(nrows, 1) * sizeof(int));
    roff_d.alloc((nrows + 1) * sizeof(int64_t));
    ctx->upload(rowptr_d.p, rowptr.data(), (nrows + 1) * sizeof(int64_t));
    ctx->upload(col_d.p, col.data(), col.size() * sizeof(int));
    ctx->upload(voff_d.p, voff.data(), voff.size() * sizeof(int64_t));
    ctx->upload(rsz_d.p, rsz.data(), nrows * sizeof(int));
    ctx->upload(roff_d.p, roff.data(), (nrows + 1) * sizeof(int64_t));
}

void Vbsr::build(bh_ctx* c, int n, const std::vector<int>& sizes, const std::vector<std::vector<int>>& cols) {
    ctx = c;
    nrows = n;
    rsz = sizes;
    roff.assign(n + 1, 0);
    for (int r = 0; r < n; ++r) roff[r + 1] = roff[r] + rsz[r];
    N = roff[n];
    rowptr.assign(n + 1, 0);
    col.clear();
    voff.clear();
    int64_t off = 0;
    for (int r = 0; r < n; ++r) {
        for (int cc : cols[r]) {
            if (rsz[cc] == 0 || rsz[r] == 0) continue;
            col.push_back(cc);
            voff.push_back(off);
            off += (int64_t)rsz[r] * rsz[cc];
        }
        rowptr[r + 1] = (int64_t)col.size();
    }
    voff.push_back(off);
    nvals = off;
    vals.alloc(std::max<int64_t>(nvals, 1) * sizeof(double));
    BH_HIP(hipMemsetAsync(vals.p, 0, vals.bytes, ctx->stream));
    upload_structure();
}

void Vbsr::ensure_fp32() {
    if (vals32.p) return;
    vals32.alloc(std::max<int64_t>(nvals, 1) * sizeof(float));
    convert_f64_f32(ctx, vals.as<double>(), vals32.as<float>(), nvals);
}

void Vbsr::spmv(double alpha, const void* x, double beta, void* y, int prec, const void* bsub) {
    dim3 blk(256), grd((nrows + 3) / 4);
    if (prec == BH_FP32) {
        ensure_fp32();
        hipLaunchKernelGGL((vbsr_spmv_kernel<float>), grd, blk, 0, ctx->stream, nrows, rsz_d.as<int>(), roff_d.as<int64_t>(),
                           rowptr_d.as<int64_t>(), col_d.as<int>(), voff_d.as<int64_t>(), vals32.as<float>(), (const float*)x,
                           (float*)y, (float)alpha, (float)beta, (const float*)bsub);
    } else {
        hipLaunchKernelGGL((vbsr_spmv_kernel<double>), grd, blk, 0, ctx->stream, nrows, rsz_d.as<int>(), roff_d.as<int64_t>(),
                           rowptr_d.as<int64_t>(), col_d.as<int>(), voff_d.as<int64_t>(), vals.as<double>(), (const double*)x,
                           (double*)y, alpha, beta, (const double*)bsub);
    }
    BH_HIP(hipGetLastError());
}

void Vbsr::invert_diag(bool want32) {
    std::vector<int64_t> di(nrows, -1), doff(nrows + 1, 0);
    for (int r = 0; r < nrows; ++r) {
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) if (col[p] == r) di[r] = p;
        doff[r + 1] = doff[r] + (int64_t)rsz[r] * rsz[r];
    }
    diag_d.alloc(std::max(nrows, 1) * sizeof(int64_t));
    doff_d.alloc((nrows + 1) * sizeof(int64_t));
    ctx->upload(diag_d.p, di.data(), nrows * sizeof(int64_t));
    ctx->upload(doff_d.p, doff.data(), (nrows + 1) * sizeof(int64_t));
    dinv64.alloc(std::max<int64_t>(doff[nrows], 1) * sizeof(double));
    if (want32) dinv32.alloc(std::max<int64_t>(doff[nrows], 1) * sizeof(float));
    DevBuf st(sizeof(int));
    BH_HIP(hipMemsetAsync(st.p, 0, sizeof(int), ctx->stream));
    int maxbs = 1;
    for (int v : rsz) maxbs = std::max(maxbs, v);
    BH_CHECK(maxbs <= 64, "variable-block inverse: block size > 64");
    hipLaunchKernelGGL(vbsr_invert_kernel, dim3(nrows), dim3(64), (size_t)maxbs * 2 * maxbs * sizeof(double), ctx->stream, nrows,
                       rsz_d.as<int>(), diag_d.as<int64_t>(), voff_d.as<int64_t>(), doff_d.as<int64_t>(), vals.as<double>(),
                       dinv64.as<double>(), want32 ? dinv32.as<float>() : nullptr, st.as<int>());
    BH_HIP(hipGetLastError());
    int h = 0;
    ctx->download(&h, st.p, sizeof(int));
    BH_CHECK(h == 0, "variable-block inverse: singular diagonal block in row " + std::to_string(h - 1));
}

void Vbsr::block_gemv(const void* r, void* x, double omega, bool add, int prec) {
    dim3 blk(256), grd((nrows + 3) / 4);
    if (prec == BH_FP32 && fp64_inverse && dinv64.p)
        // FP32 vectors, FP64 inverse blocks and accumulation: cut-cell blocks that couple
        // species at high contrast are too ill-conditioned for FP32-rounded inverses
        hipLaunchKernelGGL((vblock_gemv_kernel<float, double>), grd, blk, 0, ctx->stream, nrows, rsz_d.as<int>(),
                           roff_d.as<int64_t>(), doff_d.as<int64_t>(), dinv64.as<double>(), (const float*)r, (float*)x, omega,
                           add ? 1 : 0);
    else if (prec == BH_FP32)
        hipLaunchKernelGGL((vblock_gemv_kernel<float, float>), grd, blk, 0, ctx->stream, nrows, rsz_d.as<int>(), roff_d.as<int64_t>(),
                           doff_d.as<int64_t>(), dinv32.as<float>(), (const float*)r, (float*)x, (float)omega, add ? 1 : 0);
    else
        hipLaunchKernelGGL((vblock_gemv_kernel<double, double>), grd, blk, 0, ctx->stream, nrows, rsz_d.as<int>(),
                           roff_d.as<int64_t>(), doff_d.as<int64_t>(), dinv64.as<double>(), (const double*)r, (double*)x, omega,
                           add ? 1 : 0);
    BH_HIP(hipGetLastError());
}

void xdg_assemble(XdgDevice& X, Vbsr& M) {
    XdgProblem& P = *X.P;
    XArgs a{};
    a.D = P.D; a.k = P.k; a.Np = P.Np; a.nq1 = P.k + 1;
    a.n0 = P.g->n[0]; a.n1 = P.g->n[1]; a.n2 = P.g->n[2];
    a.nodes1 = X.nodes1.as<double>();
    a.nstride = X.nstride;
    a.elems = X.elems.as<XElem>();
    a.nelem = (int)P.elems.size();
    a.nx = X.nx.as<double>(); a.nw = X.nw.as<double>(); a.nn = X.nn.as<double>();
    a.ex = X.ex.as<int>();
    a.gx = X.gx.as<double>(); a.gw = X.gw.as<double>();
    a.spmask = X.spmask.as<int>();
    a.rowptr = M.rowptr_d.as<int64_t>(); a.col = M.col_d.as<int>(); a.voff = M.voff_d.as<int64_t>(); a.bsz = M.rsz_d.as<int>();
    a.vals = M.vals.as<double>();
    BH_CHECK(P.Np <= XMAXNP && 4 * P.Np * P.Np <= 7 * 64, "XDG assembly: degree too high");
    {
        std::vector<int64_t> off(P.elems.size(), -1);
        int64_t tot = 0;
        for (size_t i = 0; i < P.elems.size(); ++i) {
            const XElem& E = P.elems[i];
            if (E.rn == 0) continue;
            off[i] = tot;
            tot += (int64_t)(E.type == 1 || E.type == 3 ? 4 : 1) * P.Np * P.Np;
        }
        X.emat.alloc(std::max<int64_t>(tot, 1) * sizeof(double));
        X.eoff.alloc(std::max<size_t>(off.size(), 1) * sizeof(int64_t));
        X.ctx->upload(X.eoff.p, off.data(), off.size() * sizeof(int64_t));
    }
    a.emat = X.emat.as<double>();
    a.eoff = X.eoff.as<int64_t>();
    a.perm = X.perm.as<int>();
    for (int c = 0; c < XDG_COLOURS; ++c) {
        a.start = X.colour_off[c];
        a.count = X.colour_off[c + 1] - a.start;
        if (a.count == 0) continue;
        const int blocks = a.count < 65536 ? a.count : 65536;
        if (P.D == 2) hipLaunchKernelGGL(xdg_assemble_kernel<2>, dim3(blocks), dim3(64), 0, X.ctx->stream, a);
        else hipLaunchKernelGGL(xdg_assemble_kernel<3>, dim3(blocks), dim3(64), 0, X.ctx->stream, a);
        BH_HIP(hipGetLastError());
    }
}

// Uncut items (tensor rules: the whole cell of one species, or a whole face of one species)
// in modal form, one lane per item: on a Cartesian cell the orthonormal basis makes the
// volume term -nu sum_d (2/h_d)^2 S_d u (S_d the 1D Legendre stiffness in direction d,
// S_ij = sqrt((i + 1/2)(j + 1/2)) min(i,j)(min(i,j) + 1) for i + j even), and a face of
// direction d couples only modes with the same tangential exponents, through the traces
// A = sum_j p_j(+-1) u_j / sqrt(h_d / 2) and B = (2 / h_d) sum_j p_j'(+-1) u_j / sqrt(h_d / 2)
// of each face mode (the face's tangential basis is orthonormal).  The forms are the
// assembly kernel's forms (XLaplace_Bulk, SIPLaplace boundary), integrated exactly.
template <int D, int K>
__host__ __device__ constexpr bool same_tangential(int m, int n, int d) {
    for (int dd = 0; dd < D; ++dd)
        if (dd != d && mode_exponent(D, K, m, dd) != mode_exponent(D, K, n, dd)) return false;
    return true;
}
__host__ __device__ constexpr double leg_stiff(int i, int j) {
    const int m = i < j ? i : j;
    return ((i + j) & 1) ? 0.0 : sqrt_half(i) * sqrt_half(j) * (double)(m * (m + 1));
}
// p_j(+1), p_j(-1), p_j'(+1), p_j'(-1) of the orthonormal Legendre polynomial p_j
__host__ __device__ constexpr double leg_val(int j, int side) { return side ? sqrt_half(j) : ((j & 1) ? -sqrt_half(j) : sqrt_half(j)); }
__host__ __device__ constexpr double leg_der(int j, int side) {
    return (side ? 1.0 : ((j & 1) ? 1.0 : -1.0)) * sqrt_half(j) * (double)(j * (j + 1) / 2);
}

template <int D, int K, int TYPE, int DIR>
__device__ __forceinline__ void xdg_modal_item(const XEvalArgs& e, int start, int count, int t) {
    constexpr int Np = binom(K + D, D);
    const XArgs& a = e.g;
    if (t >= count) return;
    const XElem E = a.elems[a.perm[start + t]];
    double loP[3], hP[3];
    cell_box<D>(a, E.cellP, loP, hP);
    const int64_t oP = e.off[E.cellP] + (int64_t)spos(a, E.cellP, E.spP) * Np;
    double uP[Np], yP[Np];
#pragma unroll
    for (int n = 0; n < Np; ++n) { uP[n] = e.u[oP + n]; yP[n] = 0.0; }
    if constexpr (TYPE == 0) {
        double s2[D];
#pragma unroll
        for (int d = 0; d < D; ++d) s2[d] = -E.nuP * 4.0 / (hP[d] * hP[d]);
        static_for<Np>([&](auto MM) {
            constexpr int m = MM;
            static_for<Np>([&](auto NN) {
                constexpr int n = NN;
                static_for<D>([&](auto DD) {
                    constexpr int d = DD;
                    constexpr int i = mode_exponent(D, K, m, d), j = mode_exponent(D, K, n, d);
                    if constexpr (same_tangential<D, K>(m, n, d) && i > 0 && j > 0 && leg_stiff(i, j) != 0.0)
                        yP[m] += s2[d] * (leg_stiff(i, j) * uP[n]);
                });
            });
        });
    } else {
        constexpr int d = DIR;
        const double sP = 1.0 / sqrt(0.5 * hP[d]), gP = 2.0 / hP[d];
        if constexpr (TYPE == 2) {
            // boundary face: side 1 at x_d = hi (outward +e_d), side 0 at x_d = lo (outward -e_d)
            const int side = E.side;
            const double fsgn = side ? 1.0 : -1.0;
            static_for<Np>([&](auto FF) {
                constexpr int f = FF;
                if constexpr (mode_exponent(D, K, f, d) == 0) {
                    double A = 0, B = 0;
                    static_for<Np>([&](auto NN) {
                        constexpr int n = NN;
                        if constexpr (face_root(D, K, n, d) == f) {
                            constexpr int j = mode_exponent(D, K, n, d);
                            A += (side ? leg_val(j, 1) : leg_val(j, 0)) * uP[n];
                            B += (side ? leg_der(j, 1) : leg_der(j, 0)) * uP[n];
                        }
                    });
                    A *= sP;
                    const double dn = fsgn * gP * sP * B;
                    const double CV = E.nuP * dn - E.nuP * E.mu * A, CD = E.nuP * A * fsgn * gP * sP;
                    static_for<Np>([&](auto MM) {
                        constexpr int m = MM;
                        if constexpr (face_root(D, K, m, d) == f) {
                            constexpr int j = mode_exponent(D, K, m, d);
                            yP[m] += CV * sP * (side ? leg_val(j, 1) : leg_val(j, 0)) + CD * (side ? leg_der(j, 1) : leg_der(j, 0));
                        }
                    });
                }
            });
        } else {
            // interior face from P (x_d = hi of P) to Q (x_d = lo of Q), normal +e_d
            double loQ[3], hQ[3];
            cell_box<D>(a, E.cellQ, loQ, hQ);
            const int64_t oQ = e.off[E.cellQ] + (int64_t)spos(a, E.cellQ, E.spQ) * Np;
            const double sQ = 1.0 / sqrt(0.5 * hQ[d]), gQ = 2.0 / hQ[d];
            double uQ[Np], yQ[Np];
#pragma unroll
            for (int n = 0; n < Np; ++n) { uQ[n] = e.u[oQ + n]; yQ[n] = 0.0; }
            static_for<Np>([&](auto FF) {
                constexpr int f = FF;
                if constexpr (mode_exponent(D, K, f, d) == 0) {
                    double AP = 0, BP = 0, AQ = 0, BQ = 0;
                    static_for<Np>([&](auto NN) {
                        constexpr int n = NN;
                        if constexpr (face_root(D, K, n, d) == f) {
                            constexpr int j = mode_exponent(D, K, n, d);
                            AP += leg_val(j, 1) * uP[n];
                            BP += leg_der(j, 1) * uP[n];
                            AQ += leg_val(j, 0) * uQ[n];
                            BQ += leg_der(j, 0) * uQ[n];
                        }
                    });
                    AP *= sP;
                    BP *= gP * sP;
                    AQ *= sQ;
                    BQ *= gQ * sQ;
                    const double jump = AP - AQ, flux = 0.5 * (E.nuP * BP + E.nuQ * BQ);
                    const double CVP = (flux - E.mu * jump) * sP, CDP = 0.5 * E.nuP * jump * gP * sP;
                    const double CVQ = (E.mu * jump - flux) * sQ, CDQ = 0.5 * E.nuQ * jump * gQ * sQ;
                    static_for<Np>([&](auto MM) {
                        constexpr int m = MM;
                        if constexpr (face_root(D, K, m, d) == f) {
                            constexpr int j = mode_exponent(D, K, m, d);
                            yP[m] += CVP * leg_val(j, 1) + CDP * leg_der(j, 1);
                            yQ[m] += CVQ * leg_val(j, 0) + CDQ * leg_der(j, 0);
                        }
                    });
                }
            });
#pragma unroll
            for (int m = 0; m < Np; ++m) e.y[oQ + m] += e.alpha * yQ[m];
        }
    }
#pragma unroll
    for (int m = 0; m < Np; ++m) xdg_put<TYPE == 0>(e, oP + m, yP[m]);
}

// One colour group per launch: its cut items (first nbc workgroups, a lane per item row,
// from the kept element matrices) and its uncut items (the other workgroups, a lane per
// item, modal); the group's items touch disjoint rows, so both parts run together.
template <int D, int K, int TYPE, int DIR>
__global__ __launch_bounds__(64) void xdg_colour_kernel(XEvalArgs e, int cstart, int ccount, int ustart, int ucount, int nbc) {
    if ((int)blockIdx.x < nbc) {
        xdg_emat_item<TYPE == 1 || TYPE == 3, TYPE == 0>(e, cstart, ccount, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
    } else if constexpr (TYPE != 3) {
        xdg_modal_item<D, K, TYPE, DIR>(e, ustart, ucount, ((int)blockIdx.x - nbc) * (int)blockDim.x + (int)threadIdx.x);
    }
}


// one launch per colour group: cut items from their element matrices, uncut items modal
struct ColourRange {
    int cstart, ccount, ustart, ucount;
};
template <int D, int K, int TYPE, int DIR>
static void xdg_eval_type(hipStream_t st, const ColourRange& cr, const XEvalArgs& e) {
    constexpr bool TWO = TYPE == 1 || TYPE == 3;
    const int64_t nc = (int64_t)cr.ccount * (TWO ? 2 : 1) * e.g.Np;
    const int64_t nbc = (nc + 63) / 64, nbu = TYPE == 3 ? 0 : ((int64_t)cr.ucount + 63) / 64;
    BH_CHECK(nbc + nbu < ((int64_t)1 << 31), "XDG evaluate: colour group too large");
    if (nbc + nbu == 0) return;
    hipLaunchKernelGGL((xdg_colour_kernel<D, K, TYPE, DIR>), dim3((unsigned)(nbc + nbu)), dim3(64), 0, st, e, cr.cstart,
                       cr.ccount, cr.ustart, cr.ucount, (int)nbc);
    BH_HIP(hipGetLastError());
}

template <int D, int K, int TYPE>
static void xdg_eval_face(hipStream_t st, int dir, const ColourRange& cr, const XEvalArgs& e) {
    if (dir == 0) xdg_eval_type<D, K, TYPE, 0>(st, cr, e);
    else if (dir == 1) xdg_eval_type<D, K, TYPE, 1>(st, cr, e);
    else if constexpr (D == 3) xdg_eval_type<D, K, TYPE, 2>(st, cr, e);
}

template <int D, int K>
static void xdg_eval_dk(hipStream_t st, int type, int dir, const ColourRange& cr, const XEvalArgs& e) {
    switch (type) {
        case 0: xdg_eval_type<D, K, 0, -1>(st, cr, e); break;
        case 1: xdg_eval_face<D, K, 1>(st, dir, cr, e); break;
        case 2: xdg_eval_face<D, K, 2>(st, dir, cr, e); break;
        default: xdg_eval_type<D, K, 3, -1>(st, cr, e); break;
    }
}

static void xdg_evaluate_launch(hipStream_t st, int D, int k, int type, int dir, const ColourRange& cr, const XEvalArgs& e) {
    if (D == 2) {
        if (k == 1) xdg_eval_dk<2, 1>(st, type, dir, cr, e);
        else if (k == 2) xdg_eval_dk<2, 2>(st, type, dir, cr, e);
        else xdg_eval_dk<2, 3>(st, type, dir, cr, e);
    } else {
        if (k == 1) xdg_eval_dk<3, 1>(st, type, dir, cr, e);
        else if (k == 2) xdg_eval_dk<3, 2>(st, type, dir, cr, e);
        else xdg_eval_dk<3, 3>(st, type, dir, cr, e);
    }
}

void xdg_evaluate(XdgDevice& X, double alpha, const double* u, double beta, double* y) {
    XdgProblem& P = *X.P;
    const int64_t N = P.N_full;
    XEvalArgs e{};
    XArgs& a = e.g;
    a.D = P.D; a.k = P.k; a.Np = P.Np; a.nq1 = P.k + 1;
    a.n0 = P.g->n[0]; a.n1 = P.g->n[1]; a.n2 = P.g->n[2];
    a.nodes1 = X.nodes1.as<double>();
    a.nstride = X.nstride;
    a.elems = X.elems.as<XElem>();
    a.nelem = (int)P.elems.size();
    a.nx = X.nx.as<double>(); a.nw = X.nw.as<double>(); a.nn = X.nn.as<double>();
    a.ex = X.ex.as<int>();
    a.gx = X.gx.as<double>(); a.gw = X.gw.as<double>();
    a.spmask = X.spmask.as<int>();
    a.emat = X.emat.as<double>();
    a.eoff = X.eoff.as<int64_t>();
    e.off = X.offf.as<int64_t>();
    e.u = u;
    e.y = y;
    e.alpha = alpha;
    e.beta = beta;
    BH_CHECK(X.colour_off[1] > X.colour_off[0] || N == 0, "XDG evaluate: no volume items");
    BH_CHECK(P.k >= 1 && P.k <= 3 && (P.D == 2 || P.D == 3), "XDG evaluate: degrees 1 to 3 in 2D or 3D");
    a.perm = X.perm.as<int>();
    for (int c = 0; c < XDG_COLOURS; ++c) {
        // the colour's cut items (own rules) and its uncut items (tensor rules)
        const ColourRange cr{X.colour_off[c], X.cut_end[c] - X.colour_off[c], X.cut_end[c], X.colour_off[c + 1] - X.cut_end[c]};
        const int type = c == 0 ? 0 : c == 1 ? 3 : c < 8 ? 2 : 1;
        const int dir = c < 2 ? -1 : c < 8 ? (c - 2) / 2 : (c - 8) / 2;   // colours of the faces: per direction
        BH_CHECK(type != 3 || cr.ucount == 0, "XDG evaluate: interface items without a rule");
        xdg_evaluate_launch(X.ctx->stream, P.D, P.k, type, dir, cr, e);
    }
}

}  // namespace bh

// XDG device objects: variable-block matrix, device copies of the element
// lists and rules, the agglomerated system and its two-level preconditioner.
#pragma once
#include <memory>
#include <vector>

#include "bh_internal.hpp"
#include "bh_mg.hpp"
#include "bh_xdg.hpp"

namespace bh {

// Block-sparse matrix with per-row block sizes (cell blocks of 10 or 20 DOF
// for k = 2 in 3D).  Blocks are stored column-major, one after another in
// row order; voff[p] is the first value of block p.
struct Vbsr {
    bh_ctx* ctx = nullptr;
    int nrows = 0;
    int64_t N = 0, nvals = 0;
    std::vector<int> rsz, col;
    std::vector<int64_t> roff, rowptr, voff;
    DevBuf vals, vals32;
    DevBuf rowptr_d, col_d, voff_d, rsz_d, roff_d;
    DevBuf diag_d, doff_d, dinv64, dinv32;
    void build(bh_ctx* c, int n, const std::vector<int>& sizes, const std::vector<std::vector<int>>& cols);
    void upload_structure();
    void ensure_fp32();
    // y <- alpha A x + beta y; with bsub: y <- bsub - A x
    void spmv(double alpha, const void* x, double beta, void* y, int prec, const void* bsub = nullptr);
    void invert_diag(bool want32);
    // x (+)= omega Dinv r
    void block_gemv(const void* r, void* x, double omega, bool add, int prec);
    bool fp64_inverse = false;   // FP32 vectors with the FP64 inverse (see block_gemv)
    int64_t find(int r, int c) const {
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) if (col[p] == c) return p;
        return -1;
    }
};

// Elements are processed in groups ("colours") within which no two elements add
// to the same (cell, species) rows or matrix entries: volume items, interface
// items, boundary faces per (direction, side), interior faces per (direction,
// parity of the lower cell's index).  The groups run one after another in a fixed
// order, so every entry receives its contributions in the same order on every
// run: plain adds, no atomics, bitwise reproducible.
constexpr int XDG_COLOURS = 14;

struct XdgDevice {
    bh_ctx* ctx = nullptr;
    XdgProblem* P = nullptr;
    int nstride = 0;
    DevBuf nodes1, elems, nx, nw, nn, ex, gx, gw, spmask, offf;
    DevBuf perm;                            // element indices grouped by colour (cut items first in each)
    std::vector<int> colour_off;            // [XDG_COLOURS + 1] offsets into perm
    std::vector<int> cut_end;               // [XDG_COLOURS] end of each colour's cut items
    // element matrices of the cut items (their own rules), kept by the assembly for the
    // evaluate: eoff[element] = offset into emat (-1 for uncut items); a face or interface
    // item holds 4 Np x Np blocks (test P / Q x trial P / Q), a volume item one
    DevBuf emat, eoff;
    void upload();
};

void xdg_assemble(XdgDevice& X, Vbsr& M);
// agglomeration Q^T A Q on the device (see bh_xdg.hip): nonzero flags of the contributions,
// the unit-block sums, and the cell diagonal blocks gathered from the unit blocks
constexpr int XDG_NPMAX = 10;
void qaq_flags(bh_ctx*, int64_t n, int Np, const int64_t* src, const int* rf, const double* fv, unsigned char* flags);
void qaq_gather(bh_ctx*, int64_t nblk, int Np, const int64_t* cptr, const int64_t* src, const int* rf, const int* pr,
                const int* pc, const double* fv, const double* ext, double* out);
void qaq_dcell(bh_ctx*, int J, int Np, const int* rsz, const int64_t* doff, const int64_t* dblk, const double* uv, double* dv);
// frame <-> compressed layouts of a variable-block matrix: dst_c[i] = src_f[fwd[i]];
// dst_f[k] = inv[k] >= 0 ? src_c[inv[k]] : 0 (unused frame slots become zero)
template <typename T> void frame_gather(bh_ctx*, int64_t nc, const int64_t* fwd, const T* src_f, T* dst_c);
template <typename T> void frame_scatter(bh_ctx*, int64_t nf, const int64_t* inv, const T* src_c, T* dst_f);
// XDG mass matrix: per (cell, species) block of the full layout, out[row * Np + n] (device)
void xdg_mass(XdgDevice& X, double* out_dev);
// y <- beta y + alpha A u on the full (cell, species) DOF layout, matrix-free
void xdg_evaluate(XdgDevice& X, double alpha, const double* u, double beta, double* y);
// coarse-space transfer: per cell, the first Npc modes of every species block are
// summed (restriction) / receive the coarse value (prolongation)
template <typename T> void xdg_restrict(bh_ctx*, const Vbsr& A, int Np, int Npc, const T* r, T* rc);
template <typename T> void xdg_prolong_add(bh_ctx*, const Vbsr& A, int Np, int Npc, const T* xc, T* x);

// coarse matrix of the XDG preconditioner (leading Npc x Npc parts summed per cell pair)
void xdg_coarse_gather(bh_ctx*, int64_t nq, int Np, int Npc, const int64_t* qptr, const int64_t* plist, const double* vals,
                       double* out);
void xdg_coarse_vblocks(bh_ctx*, const Vbsr& A, int Np, int Npc, const int* prow, double* out);

// Two-level preconditioner of the agglomerated XDG system: FP32 (or FP64)
// block-Jacobi smoothing on the variable-block matrix and a coarse
// correction in the species-merged degree-1 space, solved by the p/h
// multigrid of the SIP path (aggregation levels + K-cycle + dense coarsest).
struct XdgPrecond : bh_solver {
    bh_bsr* A = nullptr;                   // agglomerated matrix (species units)
    Vbsr* Av = nullptr;                    // or: a caller's variable-block matrix of cell blocks (then Dc == Av)
    Vbsr* Dc = nullptr;                    // its cell diagonal blocks (smoother; cell layout of the transfers)
    int prec = BH_FP32;
    int Np = 0, Npc = 0;
    int pre = 2, post = 2;
    double omega = 0.8;
    bh_bsr* Ac = nullptr;                  // coarse matrix (uniform blocks of Npc)
    std::unique_ptr<PMultigrid> cmg;
    DevBuf b, x, r, xc;                    // work vectors in the cycle precision
    ~XdgPrecond() override;
    void build(bh_grid* g, int degc, int min_coarse, int kcycle_inner);
    template <typename T> void residual(const T* b, const T* x, T* r);   // r = b - A x
    template <typename T> void cycle();
    void precondition(double* x, const double* b) override;
    void solve_dev(double* x, const double* b) override;   // x <- x + M(b - A x)
};

}  // namespace bh

// C-ABI object
struct bh_xdg {
    bh_ctx* ctx = nullptr;
    bh_grid* grid = nullptr;
    std::unique_ptr<bh::XdgProblem> P;
    bh::XdgDevice X;
    bh::Vbsr Afull;                         // full (cell, species) matrix, assembly target
    std::unique_ptr<bh_bsr> Ared;           // agglomerated matrix, uniform BSR over species units
    bh::Vbsr Dcell;                         // cell diagonal blocks of Ared (species coupled)
    std::vector<double> ext;                // per agglomeration pair: Np x Np extension block, row-major [src mode][tgt mode]
    std::vector<double> bnd_x;              // boundary nodes (3 per) of the Dirichlet data
    std::vector<double> bnd_w, bnd_nu, bnd_mu;
    std::vector<int> bnd_cell, bnd_sp, bnd_dir, bnd_side;
    std::vector<double> rhs_full;           // host copy of the full right-hand side
    bh::DevBuf rhs_red;
    std::vector<double> fsrc{0.0, 0.0};
};

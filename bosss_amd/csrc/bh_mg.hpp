// p/h multigrid preconditioner of libbosss_hip.so.
#pragma once
#include <vector>

#include "bh_internal.hpp"

namespace bh {

struct Mlp;

void extract_subblocks(bh_ctx* ctx, int64_t nnzb, int bsf, int bsc, const double* vf, double* vc);
// coarse Galerkin blocks: cv[q] = sum_{k in qptr[q]..qptr[q+1]} P_{prow[k]}^T fv[plist[k]] P_{fcol[plist[k]]},
// with the P block of fine cell i at P + pidx[i] * bs * bs
void galerkin_gather(bh_ctx* ctx, int64_t nq, int bs, const int64_t* qptr, const int* plist, const int* prow, const int* fcol,
                     const double* fv, const int* pidx, const double* P, double* cv);
template <typename T> void p_restrict(bh_ctx*, int64_t J, int bsf, int bsc, const T* rf, T* rc);
template <typename T> void p_prolong_add(bh_ctx*, int64_t J, int bsf, int bsc, const T* xc, T* xf);
// P: the distinct prolongation blocks; pidx[i]: block of fine cell i
template <typename T> void agg_restrict(bh_ctx*, int Jc, int bs, const int* cptr, const int* cidx, const int* pidx, const T* P,
                                       const T* rf, T* rc);
template <typename T> void agg_prolong_add(bh_ctx*, int Jf, int bs, const int* parent, const int* pidx, const T* P, const T* xc,
                                          T* xf);
template <typename T> void block_gemv(bh_ctx*, int nrows, int bs, const T* Dinv, const T* r, T* x, double omega, bool add);
// cell ILU (block ILU on a pattern, level-scheduled): A's blocks into the pattern, one level
// of the factorisation, and the two sweeps x (+)= omega U^-1 L^-1 r
void ilu_scatter(bh_ctx*, int64_t nnzb, int bs, const int64_t* wpos, const double* A, double* W);
void ilu_factor_level(bh_ctx*, const int* rows, int n, int bs, const int64_t* prp, const int* pcol, const int64_t* pdiag,
                      double* W, double* Uinv, int* status);
template <typename T>
void ilu_sweeps(bh_ctx*, int nlev, const std::vector<int>& lvptr, const int* rows, int bs, const int64_t* prp, const int* pcol,
                const int64_t* pdiag, const T* W, const T* Uinv, const T* r, T* y, T* z, T* x, double omega, bool add);
// LevelPmg: high-order parts (modes lo..bs-1) of the diagonal blocks; x_hi += Hinv r_hi per cell
void hi_blocks(bh_ctx*, int J, int bs, int lo, const bh_bsr* A, double* out);
template <typename T> void hi_smooth(bh_ctx*, int J, int bs, int lo, const T* Hinv, const T* r, T* x);
// block equilibration (DiagBlockEquilib): factors of the diagonal blocks D = C S C^T (S signs) and
// L = C^-1 (L D L^T = S), the transform A_ij <- L_i A_ij L_j^T, transfer blocks Pt_i = C_i^T P_i L_I^T,
// and the smoother x (+)= omega S r of an equilibrated level
void chol_inv_blocks(bh_bsr* A, DevBuf& C, DevBuf& L, DevBuf& S);
template <typename T> void sign_smooth(bh_ctx*, int64_t n, double omega, const T* S, const T* r, T* x, bool add);
void block_sandwich(bh_bsr* A, const double* L);
void transfer_blocks(bh_ctx*, int Jf, int bsf, int bsc, const double* C, const int* pidx, const double* P, const int* parent,
                     const double* L, double* Pt, float* Pt32);
template <typename T> void rect_restrict(bh_ctx*, int Jc, int bsf, int bsc, const int* cptr, const int* cidx, const T* Pt,
                                         const T* rf, T* rc);
template <typename T> void rect_prolong_add(bh_ctx*, int Jf, int bsf, int bsc, const int* parent, const T* Pt, const T* xc,
                                            T* xf);
// FP16 inverse blocks scaled per block (FP32 vectors): bs in {10, 15, 20, 21, 28, 35}
void blocks_to_f16(bh_ctx*, int nblocks, int bs, const float* src, _Float16* dst, float* scale);
void block_gemv_h(bh_ctx*, int nrows, int bs, const _Float16* Dinv, const float* scale, const float* r, float* x, double omega,
                  bool add);
// coarsest solve: FP64 inverse and accumulation for either cycle precision
template <typename T, typename M> void dense_gemv(bh_ctx*, int n, const M* A, const T* b, T* x);
template <typename T> void vcopy(bh_ctx*, const T* s, T* d, int64_t n);
void bsr_residual(bh_bsr* A, const void* b, const void* x, void* r, int prec);

// device-resident scalars of the K-cycle's inner flexible GMRES (offsets in doubles)
struct KcycleScalars {
    int kin = 0;
    int B2 = 0, W2 = 1, H = 8, H1 = 0, H2 = 0, Y = 0, total = 0;
    explicit KcycleScalars(int k = 0) : kin(k) {
        H1 = H + (kin + 1) * kin;
        H2 = H1 + kin + 1;
        Y = H2 + kin + 1;
        total = Y + kin;
    }
};
template <typename T> void scale_inv_norm(bh_ctx*, const T* src, T* dst, int64_t n, const double* s2, double tiny);
// aggregate Schwarz blocks: dense column-major (cmax*bs)^2 per aggregate, then x (+)= omega sum_I R_I^T Ainv_I R_I r
void gather_agg_blocks(bh_ctx*, int nagg, int bs, int cmax, const int* cptr, const int* cidx, const bh_bsr* A, double* out);
template <typename T> void schwarz_apply(bh_ctx*, int nagg, int bs, int cmax, const int* cptr, const int* cidx, const T* Ainv,
                                         const T* r, T* x, double omega, bool add);
// overlapping blocks (bptr/bidx: cells per block; iptr/iidx: per cell its flat (block, slot) positions);
// scale: the block sum of a DOF divided by its multiplicity
template <typename T> void schwarz_apply_overlap(bh_ctx*, int nblk, int bs, int bmax, const int* bptr, const int* bidx,
                                                 const int* iptr, const int* iidx, int ncell, const T* Ainv, const T* r,
                                                 T* Y, T* x, double omega, bool add, bool scale);
void invert_dense_blocks(bh_ctx*, int nblocks, int n, const double* blocks, double* inv64, float* inv32);
// Gauss-Jordan inverse (partial pivoting) of a dense row-major n x n device matrix; A is overwritten
void dense_inverse_dev(bh_ctx*, double* A, double* Inv, int n);
void kcycle_column(bh_ctx*, double* S, const KcycleScalars& L, int j);
void kcycle_lsq(bh_ctx*, double* S, const KcycleScalars& L);
template <typename T> void orth2_step(bh_ctx*, const T* e, T* x, int64_t n, double* st, int* cnt);
template <typename T> void sub_vec(bh_ctx*, const T* b, const T* r, T* q, int64_t n);

struct MgLevel {
    int J = 0, bs = 1;
    int64_t N = 0;
    int n[3] = {1, 1, 1};          // logical grid dims
    int deg = 0;                   // DG degree of the level
    bh_bsr* A = nullptr;           // level operator (Galerkin)
    bool ownA = false;
    bh_op* op = nullptr;           // matrix-free operator (finest level; coarse p-levels of a matrix-free hierarchy)
    std::unique_ptr<bh_op> ownop;  // coarse p-level operator owned by the hierarchy
    DevBuf dinv64, dinv32;
    DevBuf dinv16, dscale;         // FP16-MFMA cycle, level 0: FP16 inverse blocks and their FP32 scales
    // transfer to the next coarser level: 0 = degree truncation, 1 = aggregation
    int kind = 0;
    DevBuf cptr, cidx, parent, pidx;       // aggregation: children CSR, parents, P block index per fine cell
    DevBuf P64, P32;                       // the distinct prolongation blocks (few on a regular grid)
    std::vector<double> lo, h;     // cell boxes (J x D), for the aggregation prolongation
    // work vectors (precision of the cycle)
    DevBuf b, x, r;
    // K-cycle storage (inner flexible Krylov on this level): rhs, V, Z, w
    DevBuf kv, kz, kw, ks;         // ks: KcycleScalars in device memory
    // additive Schwarz smoother over this level's aggregates (sw_cmax > 0: in use)
    int sw_cmax = 0, sw_nagg = 0;
    DevBuf sw64, sw32;
    // with overlap (sw_overlap > 0): block cell lists, per-cell positions, local solutions
    int sw_overlap = 0;
    bool sw_scale = false;
    DevBuf sw_bptr, sw_bidx, sw_iptr, sw_iidx, sw_y64, sw_y32;
    // block equilibration (level in the transformed basis x~ = L^-T x): the inverse
    // factors L (and L^T, level 0 only, for the output) and the transfer blocks to the next level
    DevBuf eqL64, eqL32, eqLt64, eqLt32, eqP64, eqP32, eqcptr, eqcidx, eqparent;
    DevBuf eqS64, eqS32;           // signs of the equilibrated diagonal blocks (L D L^T = diag(eqS))
};

struct PMultigrid : bh_solver {
    int prec = BH_FP64;
    int dim = 3;                   // space dimension of the grid
    bool mfma = false;             // FP32 cycle: matrix-free level 0 on the FP16 matrix cores
    bool l0_mfma_apply = true;     // ... for its operator apply too (timed at set-up against the FP32 kernel)
    double omega = 1.0;
    int pre = 1, post = 1;
    bool matrix_free = true;
    bool equilib = false;          // DiagBlockEquilib: every level but the coarsest in its block-equilibrated basis
    // LevelPmg on the finest level: coarse (low-order) correction, then the high-order modes
    // of each cell with the inverse of their diagonal sub-block
    bool level_pmg = false;
    int lp_lo = 0;
    DevBuf lp64, lp32;
    void set_level_pmg(bool on);
    // cell ILU as the finest level's smoother (ILU(p) pattern = block pattern of A^(p+1)):
    // rows grouped by dependency level, the factors on the pattern and the diagonal inverses
    int ilu_p = -1, ilu_nlev = 0;
    std::vector<int> ilu_lvptr;
    DevBuf ilu_rows, ilu_prp, ilu_pcol, ilu_pdiag, ilu_W64, ilu_W32, ilu_U64, ilu_U32, ilu_y, ilu_z;
    void set_cell_ilu(int p);
    bh_bsr* eqA0 = nullptr;        // the transformed finest matrix (the caller's matrix stays untouched)
    int kcycle_level = -1;         // level solved by `kin` inner FGMRES steps (K-cycle), -1: none
    int kin = 0;
    std::vector<MgLevel> lv;
    int ncoarse = 0;
    DevBuf Ainv64;                 // coarsest level: dense FP64 inverse (FP64 cycle)
    DevBuf Ainv32;                 // its FP32 copy (FP32 cycle)
    DevBuf in32, out32;
    bh_bsr* ownedAf = nullptr;
    // learned correction after the cycle: residual minimisation over the cycle's and the
    // network's corrections (the orthonormalisation multigrid's step with a history of one)
    struct Mlp* learned = nullptr;
    bh_op* fine_op = nullptr;
    DevBuf lX, ls, lY, le, lrq;     // features, scales, outputs, correction, [r' | A z1 | A e]
    DevBuf lstat;                   // Gram entries and right-hand sides, sum of gains, int counters
    void learned_stats(int* acc, int* rej);
    // fallback: after `learned_probe` preconditioner calls, a mean removed fraction of
    // |r'|^2 below `learned_min_gain` switches the learned step off (plain multigrid)
    int learned_probe = 8;
    double learned_min_gain = 1e-2;
    int learned_calls = 0;
    bool learned_off = false;
    double learned_mean_gain = 0.0;
    // cost rule at the end of the probe window: flexible GMRES preconditioned with the
    // learned step must reduce the residual at least as fast per millisecond as with
    // the plain cycle (per-iteration times and log-reductions of both)
    bool learned_cost_rule = true;
    bool in_cost_rule = false;
    double learned_t_cycle = 0.0, learned_t_net = 0.0, learned_l_cycle = 0.0, learned_l_net = 0.0;
    void krylov_rate(bool net, int steps, double& lred, double& ms);
    template <typename T> void learned_correction();
    void attach_learned(struct Mlp* m);
    ~PMultigrid() override;
    void build(bh_op* op, bh_bsr* Af, const std::vector<int>& degrees, int min_coarse, bool matrix_free);
    // hierarchy over a Cartesian grid from an assembled matrix of degree degrees[0] (mfop: optional matrix-free level 0)
    void build_grid(bh_ctx* ctx, bh_grid* grid, bh_bsr* Af, const std::vector<int>& degrees, int min_coarse, bh_op* mfop);
    void solve_dev(double* x, const double* b) override;        // x <- x + V(b - A x)
    void precondition(double* x, const double* b) override;     // x <- V(b)
    template <typename T> void vcycle(int l, const T* b, T* x);   // x = (V-cycle from level l) b; scratch: level l's r and coarser
    template <typename T> void vcycle(int l) { vcycle<T>(l, vec<T>(lv[l].b), vec<T>(lv[l].x)); }
    // V-cycles that contain no host synchronisation (the whole cycle without a
    // K-cycle, or the inner cycle of the K-cycle) are captured once into a HIP
    // graph and replayed: one launch instead of one per kernel of every level
    bool use_graphs = true;
    bool capturing = false;        // inside a capture: nested cycles are recorded, not captured again
    template <typename T> void cycle_body(int l);   // cycle from level l (+ learned step on level 0)
    hipGraphExec_t gexec[2][16] = {};
    int gcalls[2][16] = {};
    template <typename T> void vcycle_g(int l);
    void drop_graphs();
    template <typename T> void kcycle(int l);
    template <typename T> void solve_level(int l) { if (l == kcycle_level && kin > 0) kcycle<T>(l); else vcycle<T>(l); }
    template <typename T> void residual(int l, const T* b, const T* x, T* r);
    template <typename T> void apply(int l, const T* x, T* y);   // y = A_l x
    template <typename T> void smooth(int l, const T* r, T* x, bool add);   // x (+)= omega M_l^-1 r
    template <typename T> void restrict_to(int l, const T* r);     // lv[l+1].b = R_l r
    template <typename T> void prolong_add(int l, T* x);           // x += P_l lv[l+1].x
    // aggregate-block Schwarz on every level that has aggregates, the blocks grown by
    // `overlap` layers of face neighbours; scale: overlap scaling
    void set_schwarz(bool on, int overlap = 0, bool scale = false);
    void set_equilib();            // DiagBlockEquilib (assembled hierarchies; irreversible)
    int op_prec(int l, bool f32) const {
        return f32 ? ((l == 0 && mfma && l0_mfma_apply) ? BH_FP16_MFMA : BH_FP32) : BH_FP64;
    }
    template <typename T> T* vec(DevBuf& d) { return (T*)d.p; }
};

}  // namespace bh

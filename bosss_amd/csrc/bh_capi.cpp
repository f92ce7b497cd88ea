// extern "C" entry points of libbosss_hip.so (declared in include/bosss_hip.h).
#include <algorithm>
#include <functional>
#include <cmath>
#include <cstring>

#include "bh_internal.hpp"
#include "bh_solvers.hpp"
#include "bh_mg.hpp"
#include "bh_mlp.hpp"
#include "bh_euler.hpp"
#include "bh_xdgdev.hpp"
#include "bh_ortho.hpp"

namespace bh {
const std::string& last_error();
}
using namespace bh;


// ------------------------------------------------------------------ helpers
namespace {

// Stage a host or device vector on the device for the duration of a call.
struct Staged {
    bh_ctx* ctx;
    DevBuf buf;
    void* dptr;
    Staged(bh_ctx* c, const void* p, size_t bytes, bool device, bool upload) : ctx(c) {
        if (device || p == nullptr) { dptr = (void*)p; return; }
        buf.alloc(bytes);
        dptr = buf.p;
        if (upload) ctx->upload(dptr, p, bytes);
    }
    void download(void* host, size_t bytes, bool device) {
        if (!device && host) ctx->download(host, dptr, bytes);
    }
};

// A solver in the caller's layout (frame) of a variable-block matrix around a solver in
// its compressed layout: gathers b and x, runs the inner solver, scatters x back with
// the unused frame slots set to zero.
struct FrameSolver : bh_solver {
    bh_bsr* M = nullptr;
    std::unique_ptr<bh_solver> inner;
    DevBuf xc, bc;
    FrameSolver(bh_bsr* m, bh_solver* in) : M(m), inner(in) {
        ctx = m->ctx;
        N = m->fr_N;
        xc.alloc(std::max<int64_t>(in->N, 1) * sizeof(double));
        bc.alloc(std::max<int64_t>(in->N, 1) * sizeof(double));
    }
    void sync_stats() {
        iterations = inner->iterations;
        converged = inner->converged;
        res0 = inner->res0;
        res = inner->res;
    }
    void solve_dev(double* x, const double* b) override {
        frame_gather<double>(ctx, inner->N, M->fr_fwd.as<int64_t>(), b, bc.as<double>());
        frame_gather<double>(ctx, inner->N, M->fr_fwd.as<int64_t>(), x, xc.as<double>());
        inner->solve_dev(xc.as<double>(), bc.as<double>());
        frame_scatter<double>(ctx, N, M->fr_inv.as<int64_t>(), xc.as<double>(), x);
        sync_stats();
    }
    void precondition(double* x, const double* b) override {
        frame_gather<double>(ctx, inner->N, M->fr_fwd.as<int64_t>(), b, bc.as<double>());
        inner->precondition(xc.as<double>(), bc.as<double>());
        frame_scatter<double>(ctx, N, M->fr_inv.as<int64_t>(), xc.as<double>(), x);
        sync_stats();
    }
};

// wraps a compressed-layout solver of M for the caller's layout (when the two differ)
bh_solver* in_frame(bh_bsr* M, bh_solver* s) { return (M && M->var && M->fr_gaps) ? new FrameSolver(M, s) : s; }
// the compressed-layout solver behind a caller-layout one
bh_solver* compressed(bh_solver* s) {
    FrameSolver* f = dynamic_cast<FrameSolver*>(s);
    return f ? f->inner.get() : s;
}

void build_grid_edges(bh_grid* g) {
    const int D = g->D;
    for (int c = 0; c < g->J; ++c) {
        int idx[3];
        g->cell_ijk(c, idx);
        for (int d = 0; d < D; ++d) {
            if (idx[d] == 0) { g->e_in.push_back(c); g->e_out.push_back(-1); g->e_dir.push_back(d); g->e_side.push_back(2 * d); }
            if (idx[d] == g->n[d] - 1) {
                g->e_in.push_back(c); g->e_out.push_back(-1); g->e_dir.push_back(d); g->e_side.push_back(2 * d + 1);
            } else {
                int nb[3] = {idx[0], idx[1], idx[2]};
                nb[d] += 1;
                g->e_in.push_back(c); g->e_out.push_back(g->cell_index(nb[0], nb[1], nb[2]));
                g->e_dir.push_back(d); g->e_side.push_back(-1);
            }
        }
    }
    // cj (F/GridData_CellData.cs:666-700): boundary faces count twice
    g->cj.assign(g->J, 0.0);
    for (int c = 0; c < g->J; ++c) {
        int idx[3];
        g->cell_ijk(c, idx);
        double h[3], vol = 1, surf = 0, bnd = 0;
        for (int d = 0; d < D; ++d) { h[d] = g->width[d][idx[d]]; vol *= h[d]; }
        for (int d = 0; d < D; ++d) {
            double ar = vol / h[d];
            surf += 2 * ar;
            if (idx[d] == 0) bnd += ar;
            if (idx[d] == g->n[d] - 1) bnd += ar;
        }
        g->cj[c] = 0.5 * (surf + bnd) / vol;
    }
}

}  // namespace

void bh_op::rebuild_affine() {
    const int D = this->D;
    const int Np = this->Np;
    std::vector<double> aff;
    bool any_bnd = false;
    for (double v : gbnd) if (v != 0.0) { any_bnd = true; break; }
    affine.alloc(N * sizeof(double));
    fill(ctx, affine.as<double>(), 0.0, N);
    if (any_bnd) {
        aff.assign(N, 0.0);
        bh_grid* g = grid;
        const int nqf = tab.nqf;
        for (size_t e = 0; e < g->e_in.size(); ++e) {
            if (bnd_slot[e] < 0) continue;
            const int c = g->e_in[e], d = g->e_dir[e], side = g->e_side[e], s = side & 1;
            int idx[3];
            g->cell_ijk(c, idx);
            double h[3], detJ = 1;
            for (int dd = 0; dd < D; ++dd) { h[dd] = g->width[dd][idx[dd]]; detJ *= 0.5 * h[dd]; }
            const double dA = detJ / (0.5 * h[d]);
            const double sdinv = 1.0 / std::sqrt(detJ);
            const double sgn = s ? 1.0 : -1.0;
            const double nuP = h_nu[c];
            const double pn = 2.0 * penalty_base * penalty_deg * g->cj[c];
            for (int q = 0; q < nqf; ++q) {
                const double gv = gbnd[bnd_slot[e] + q];
                const double wq = tab.fw[q] * dA * sdinv;
                for (int m = 0; m < Np; ++m) {
                    const double phi = tab.T[((size_t)side * nqf + q) * Np + m];
                    const double dn = sgn * (2.0 / h[d]) * tab.Nn[((size_t)side * nqf + q) * Np + m];
                    double v;
                    if (bc[side] == 0) v = nuP * gv * pn * phi - nuP * gv * dn;   // ipLaplace.cs:224-236 at u = 0
                    else v = nuP * gv * phi;                                     // ipLaplace.cs:238-243
                    aff[(size_t)c * Np + m] += wq * v;
                }
            }
        }
        ctx->upload(affine.p, aff.data(), N * sizeof(double));
    }
    if (rhs.p) axpby(ctx, -1.0, rhs.as<double>(), 1.0, affine.as<double>(), N);  // -rhs*v (SipPoissonMain.cs:431-434)
    affine_valid = true;
}

namespace bh {
// SIP operator of degree k on grid g (penalty_deg <= 0: the degree's own
// penalty factor, ipLaplace.cs:100-110; a coarse p-level of a multigrid passes
// the finest level's factor so that its operator is the Galerkin submatrix)
std::unique_ptr<bh_op> make_sip_op(bh_grid* g, int k, double penalty_base, double penalty_deg, const double* nu,
                                   const int* bc) {
    BH_CHECK(k >= 1 && k <= (g->D == 2 ? 4 : 3), "bh_sip_create: supported degrees are 1..4 (2D), 1..3 (3D)");
    auto op = std::make_unique<bh_op>();
    op->ctx = g->ctx;
    op->grid = g;
    op->D = g->D;
    op->k = k;
    op->tab.build(g->D, k, k + 1);   // Gauss rule of order 2k (precision 2k+1), F/RefElement.cs:855-884
    op->Np = op->tab.Np;
    op->N = (int64_t)g->J * op->Np;
    op->penalty_base = penalty_base;
    const double D = g->D, p = k;
    op->penalty_deg = penalty_deg > 0 ? penalty_deg : std::max((p + 1) * (p + D) / D, (p + 1.0) * (p + 1.0));
    for (int f = 0; f < 2 * g->D; ++f) op->bc[f] = bc ? bc[f] : 0;
    op->h_nu.assign(nu, nu + g->J);
    for (double v : op->h_nu) BH_CHECK(std::isfinite(v), "Inf/NaN in diffusion coefficient");
    bh_ctx* ctx = g->ctx;
    op->nu.alloc(g->J * sizeof(double));
    op->nu32.alloc(g->J * sizeof(float));
    ctx->upload(op->nu.p, nu, g->J * sizeof(double));
    convert_f64_f32(ctx, op->nu.as<double>(), op->nu32.as<float>(), g->J);
    // tables
    std::vector<double> t;
    op->off_G = t.size(); t.insert(t.end(), op->tab.G.begin(), op->tab.G.end());
    op->off_T = t.size(); t.insert(t.end(), op->tab.T.begin(), op->tab.T.end());
    op->off_N = t.size(); t.insert(t.end(), op->tab.Nn.begin(), op->tab.Nn.end());
    op->off_vw = t.size(); t.insert(t.end(), op->tab.vw.begin(), op->tab.vw.end());
    op->off_fw = t.size(); t.insert(t.end(), op->tab.fw.begin(), op->tab.fw.end());
    op->off_K = t.size(); t.insert(t.end(), op->tab.Kv.begin(), op->tab.Kv.end());
    op->tab64.alloc(t.size() * sizeof(double));
    ctx->upload(op->tab64.p, t.data(), t.size() * sizeof(double));
    std::vector<float> t32(t.begin(), t.end());
    op->tab32.alloc(t32.size() * sizeof(float));
    ctx->upload(op->tab32.p, t32.data(), t32.size() * sizeof(float));
    // boundary slots
    op->bnd_slot.assign(g->e_in.size(), -1);
    int slot = 0;
    for (size_t e = 0; e < g->e_in.size(); ++e)
        if (g->e_out[e] < 0) { op->bnd_slot[e] = slot; slot += op->tab.nqf; }
    op->gbnd.assign(slot, 0.0);
    return op;
}
}  // namespace bh

extern "C" {

int bh_last_error(char* buf, int len) {
    const std::string& s = bh::last_error();
    if (buf && len > 0) {
        std::strncpy(buf, s.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return (int)s.size();
}

int bh_init(int device, bh_ctx** out) {
    BH_API_BEGIN
    int n = 0;
    BH_HIP(hipGetDeviceCount(&n));
    BH_CHECK(device >= 0 && device < n, "bh_init: no such HIP device");
    BH_HIP(hipSetDevice(device));
    mlp_noalloc_setup(device);   // device address of the allocation-free workspace on this GPU
    bh_ctx* c = new bh_ctx();
    c->device = device;
    BH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto& e : c->ev) BH_HIP(hipEventCreate(&e));
    *out = c;
    BH_API_END
}

int bh_finalize(bh_ctx* c) {
    BH_API_BEGIN
    if (!c) return 0;
    (void)hipStreamSynchronize(c->stream);
    c->red.release();
    c->red_out.release();
    if (c->host_red) (void)hipHostFree(c->host_red);
    for (auto& e : c->ev) (void)hipEventDestroy(e);
    (void)hipStreamDestroy(c->stream);
    delete c;
    BH_API_END
}

int bh_sync(bh_ctx* c) {
    BH_API_BEGIN
    c->sync();
    BH_API_END
}

int bh_malloc(bh_ctx* c, size_t bytes, void** p) {
    BH_API_BEGIN
    (void)c;
    BH_HIP(hipMalloc(p, bytes));
    BH_API_END
}
int bh_free(bh_ctx* c, void* p) {
    BH_API_BEGIN
    BH_HIP(hipStreamSynchronize(c->stream));
    BH_HIP(hipFree(p));
    BH_API_END
}
int bh_memcpy_h2d(bh_ctx* c, void* dst, const void* src, size_t bytes) {
    BH_API_BEGIN
    c->upload(dst, src, bytes);
    BH_API_END
}
int bh_memcpy_d2h(bh_ctx* c, void* dst, const void* src, size_t bytes) {
    BH_API_BEGIN
    c->download(dst, src, bytes);
    BH_API_END
}
int bh_memcpy_d2d(bh_ctx* c, void* dst, const void* src, size_t bytes) {
    BH_API_BEGIN
    copy_bytes(c, dst, src, bytes);
    BH_API_END
}
int bh_dot(bh_ctx* c, int64_t n, const double* x, const double* y, double* result) {
    BH_API_BEGIN
    BH_CHECK(c && x && y && result && n >= 0, "bh_dot: null argument or negative length");
    multi_dot(c, x, y, n, 1, n, result);   // deterministic two-pass reduction, synchronises
    BH_API_END
}
int bh_event_record(bh_ctx* c, int slot) {
    BH_API_BEGIN
    BH_CHECK(slot >= 0 && slot < 8, "event slot out of range");
    BH_HIP(hipEventRecord(c->ev[slot], c->stream));
    BH_API_END
}
int bh_event_elapsed(bh_ctx* c, int s0, int s1, float* ms) {
    BH_API_BEGIN
    BH_CHECK(s0 >= 0 && s0 < 8 && s1 >= 0 && s1 < 8, "event slot out of range");
    BH_HIP(hipEventSynchronize(c->ev[s1]));
    BH_HIP(hipEventElapsedTime(ms, c->ev[s0], c->ev[s1]));
    BH_API_END
}

// ------------------------------------------------------------------ grid
int bh_grid_cartesian(bh_ctx* ctx, int D, const double* x, int nx1, const double* y, int ny1, const double* z,
                      int nz1, bh_grid** out) {
    BH_API_BEGIN
    BH_CHECK(D == 2 || D == 3, "bh_grid_cartesian: D must be 2 or 3");
    const double* src[3] = {x, y, z};
    int len[3] = {nx1, ny1, nz1};
    auto g = std::make_unique<bh_grid>();
    g->ctx = ctx;
    g->D = D;
    g->nmax = 1;
    for (int d = 0; d < D; ++d) {
        BH_CHECK(src[d] && len[d] >= 2, "bh_grid_cartesian: need at least 2 nodes per direction");
        g->nodes[d].assign(src[d], src[d] + len[d]);
        g->n[d] = len[d] - 1;
        g->width[d].resize(g->n[d]);
        for (int i = 0; i < g->n[d]; ++i) {
            g->width[d][i] = src[d][i + 1] - src[d][i];
            BH_CHECK(g->width[d][i] > 0, "bh_grid_cartesian: nodes must be strictly increasing");  // CheckMonotonicity
        }
        g->nmax = std::max(g->nmax, g->n[d]);
    }
    int64_t J = (int64_t)g->n[0] * g->n[1] * (D == 3 ? g->n[2] : 1);
    BH_CHECK(J < (1LL << 31), "grid too large");
    g->J = (int)J;
    build_grid_edges(g.get());
    // [4][3][nmax]: h, 1/h, sqrt(h), 1/sqrt(h) per direction (the kernels' per-face scalings without
    // divisions or square roots)
    std::vector<double> w(4 * 3 * (size_t)g->nmax, 0.0);
    for (int d = 0; d < D; ++d)
        for (int i = 0; i < g->n[d]; ++i) {
            const double h = g->width[d][i];
            w[(0 * 3 + d) * (size_t)g->nmax + i] = h;
            w[(1 * 3 + d) * (size_t)g->nmax + i] = 1.0 / h;
            w[(2 * 3 + d) * (size_t)g->nmax + i] = std::sqrt(h);
            w[(3 * 3 + d) * (size_t)g->nmax + i] = 1.0 / std::sqrt(h);
        }
    if (ctx) {   // a grid without context serves host-side set-up only (cut-cell rules, agglomeration)
        g->d_width.alloc(w.size() * sizeof(double));
        ctx->upload(g->d_width.p, w.data(), w.size() * sizeof(double));
    }
    *out = g.release();
    BH_API_END
}

int bh_grid_destroy(bh_grid* g) {
    BH_API_BEGIN
    delete g;
    BH_API_END
}

int bh_grid_info(bh_grid* g, int* D, int* J, int* E) {
    BH_API_BEGIN
    if (D) *D = g->D;
    if (J) *J = g->J;
    if (E) *E = (int)g->e_in.size();
    BH_API_END
}

int bh_grid_edges(bh_grid* g, int* in, int* out, int* dir, int* side) {
    BH_API_BEGIN
    std::copy(g->e_in.begin(), g->e_in.end(), in);
    std::copy(g->e_out.begin(), g->e_out.end(), out);
    std::copy(g->e_dir.begin(), g->e_dir.end(), dir);
    std::copy(g->e_side.begin(), g->e_side.end(), side);
    BH_API_END
}

int bh_grid_cj(bh_grid* g, double* cj) {
    BH_API_BEGIN
    std::copy(g->cj.begin(), g->cj.end(), cj);
    BH_API_END
}

// ------------------------------------------------------------------ SIP
int bh_sip_create(bh_grid* g, int k, double penalty_base, const double* nu, const int* bc, bh_op** out) {
    BH_API_BEGIN
    *out = make_sip_op(g, k, penalty_base, 0.0, nu, bc).release();
    BH_API_END
}

int bh_op_destroy(bh_op* op) {
    BH_API_BEGIN
    if (op) (void)hipStreamSynchronize(op->ctx->stream);
    delete op;
    BH_API_END
}

int bh_op_info(bh_op* op, int* Np, int64_t* N) {
    BH_API_BEGIN
    if (Np) *Np = op->Np;
    if (N) *N = op->N;
    BH_API_END
}

int bh_sip_num_bnd_nodes(bh_op* op, int* count) {
    BH_API_BEGIN
    *count = (int)op->gbnd.size();
    BH_API_END
}

int bh_sip_bnd_nodes(bh_op* op, double* X) {
    BH_API_BEGIN
    bh_grid* g = op->grid;
    const int D = op->D, nqf = op->tab.nqf;
    for (size_t e = 0; e < g->e_in.size(); ++e) {
        const int slot = op->bnd_slot[e];
        if (slot < 0) continue;
        const int c = g->e_in[e], d = g->e_dir[e], s = g->e_side[e] & 1;
        int idx[3];
        g->cell_ijk(c, idx);
        for (int q = 0; q < nqf; ++q) {
            int m = 0;
            for (int dd = 0; dd < D; ++dd) {
                double xi = (dd == d) ? (s ? 1.0 : -1.0) : op->tab.fxi[q * (D - 1) + m++];
                X[(size_t)(slot + q) * D + dd] = g->nodes[dd][idx[dd]] + 0.5 * (xi + 1.0) * g->width[dd][idx[dd]];
            }
        }
    }
    BH_API_END
}

int bh_sip_set_bnd_values(bh_op* op, const double* gv) {
    BH_API_BEGIN
    std::copy(gv, gv + op->gbnd.size(), op->gbnd.begin());
    bool any = false;
    for (double v : op->gbnd) any |= (v != 0.0);
    op->has_affine = any || op->rhs.p;
    op->affine_valid = false;
    BH_API_END
}

int bh_sip_set_rhs(bh_op* op, const double* f, int flags) {
    BH_API_BEGIN
    if (!f) {
        op->rhs.release();
    } else {
        op->rhs.alloc(op->N * sizeof(double));
        if (flags & BH_DEVICE_PTR) copy(op->ctx, op->rhs.as<double>(), f, op->N);
        else op->ctx->upload(op->rhs.p, f, op->N * sizeof(double));
    }
    bool any = false;
    for (double v : op->gbnd) any |= (v != 0.0);
    op->has_affine = any || op->rhs.p;
    op->affine_valid = false;
    BH_API_END
}

int bh_op_evaluate(bh_op* op, double alpha, const double* u, double beta, double* y, int affine, int prec, int flags) {
    BH_API_BEGIN
    const bool dev = flags & BH_DEVICE_PTR;
    bh_ctx* ctx = op->ctx;
    const int64_t N = op->N;
    if (prec == BH_FP64) {
        Staged su(ctx, u, N * sizeof(double), dev, true);
        Staged sy(ctx, y, N * sizeof(double), dev, beta != 0.0);
        sip_apply(op, alpha, su.dptr, beta, sy.dptr, affine != 0, BH_FP64);
        sy.download(y, N * sizeof(double), dev);
    } else if (prec == BH_FP32 || prec == BH_FP16_MFMA) {
        const bool mfma = prec == BH_FP16_MFMA;
        BH_CHECK(mfma || !affine || !op->has_affine, "FP32 evaluate: linear part only");
        if (dev) {
            sip_apply(op, alpha, u, beta, y, mfma && affine, prec);   // float* device vectors
        } else {
            DevBuf d64(N * sizeof(double)), u32(N * sizeof(float)), y32(N * sizeof(float));
            ctx->upload(d64.p, u, N * sizeof(double));
            convert_f64_f32(ctx, d64.as<double>(), u32.as<float>(), N);
            if (beta != 0.0) {
                ctx->upload(d64.p, y, N * sizeof(double));
                convert_f64_f32(ctx, d64.as<double>(), y32.as<float>(), N);
            }
            sip_apply(op, alpha, u32.p, beta, y32.p, mfma && affine, prec);
            convert_f32_f64(ctx, y32.as<float>(), d64.as<double>(), N);
            ctx->download(y, d64.p, N * sizeof(double));
        }
    } else {
        throw Error(-3, "bh_op_evaluate: unsupported precision");
    }
    if (!dev) ctx->sync();
    BH_API_END
}

int bh_op_tuning(bh_op* op, int* cfg, float* ms) {
    BH_API_BEGIN
    if (cfg) {
        cfg[0] = op->mfma_waves;
        cfg[1] = op->mfma_wg_per_xcd;
        cfg[2] = op->f64_path;
        cfg[3] = op->f64_path == 4 ? op->f64_tile_w : op->f64_waves;
        cfg[4] = op->f64_wg_per_xcd;
    }
    if (ms) { ms[0] = op->mfma_tuned_ms; ms[1] = op->f64_ms_vector; ms[2] = op->f64_ms_mfma; }
    BH_API_END
}

int bh_op_tuning_fp32(bh_op* op, int* path, int* waves, float* ms) {
    BH_API_BEGIN
    BH_CHECK(op, "bh_op_tuning_fp32: null operator");
    if (path) *path = op->f32_path;
    if (waves) *waves = op->f32_tile_w;
    if (ms) *ms = op->f32_ms;
    BH_API_END
}

int bh_op_set_fp64_path(bh_op* op, int path, int waves, int wg_per_xcd) {
    BH_API_BEGIN
    BH_CHECK(op, "bh_op_set_fp64_path: null operator");
    BH_CHECK(path >= 0 && path <= 4,
             "bh_op_set_fp64_path: path must be 0 (autotune), 1 (vector, face rule node by node), 2 (MFMA), 3 "
             "(vector, modal face terms) or 4 (vector, tiled modal face terms)");
    if (path == 2) {
        BH_CHECK(bh::sip_apply_mfma64_supported(op->D, op->k, op->N),
                 "bh_op_set_fp64_path: the FP64-MFMA apply supports D in {2, 3}, k in {2, 3} and vectors below 4 GiB");
        BH_CHECK(waves == 4 || waves == 8, "bh_op_set_fp64_path: waves must be 4 or 8");
        BH_CHECK(wg_per_xcd >= 1 && wg_per_xcd <= 4096, "bh_op_set_fp64_path: wg_per_xcd out of range");
        op->f64_waves = waves;
        op->f64_wg_per_xcd = wg_per_xcd;
    }
    if (path == 4) {
        BH_CHECK(waves == 2 || waves == 4 || waves == 8, "bh_op_set_fp64_path: the tiled kernel takes 2, 4 or 8 waves");
        op->f64_tile_w = waves;
    }
    if (path == 0) bh::sip_forget_tuning(op, BH_FP64);
    op->f64_path = path;
    BH_API_END
}

int bh_op_set_fp32_path(bh_op* op, int path, int waves) {
    BH_API_BEGIN
    BH_CHECK(op, "bh_op_set_fp32_path: null operator");
    BH_CHECK(path == 0 || path == 1 || path == 3 || path == 4,
             "bh_op_set_fp32_path: path must be 0 (autotune), 1 (face rule node by node), 3 (modal face terms) or 4 "
             "(tiled modal face terms)");
    if (path == 4) BH_CHECK(waves == 2 || waves == 4 || waves == 8, "bh_op_set_fp32_path: the tiled kernel takes 2, 4 or 8 waves");
    if (path == 0) bh::sip_forget_tuning(op, BH_FP32);
    op->f32_path = path;
    op->f32_tile_w = path == 4 ? waves : 0;
    BH_API_END
}

int bh_op_assemble(bh_op* op, bh_bsr** out, double* affine, int flags) {
    BH_API_BEGIN
    bh_grid* g = op->grid;
    bh_ctx* ctx = op->ctx;
    const int Np = op->Np, J = g->J;
    auto M = std::make_unique<bh_bsr>();
    M->ctx = ctx;
    M->nrows = J;
    M->bs = Np;
    // sparsity: self + face neighbours, ascending columns (SortedDictionary rows of P/BlockMsrMatrix.cs:191-340)
    M->h_rowptr.assign(J + 1, 0);
    std::vector<int> cols;
    cols.reserve((size_t)J * (2 * g->D + 1));
    for (int c = 0; c < J; ++c) {
        int idx[3];
        g->cell_ijk(c, idx);
        int tmp[7], nt = 0;
        tmp[nt++] = c;
        for (int d = 0; d < g->D; ++d) {
            if (idx[d] > 0) { int nb[3] = {idx[0], idx[1], idx[2]}; nb[d]--; tmp[nt++] = g->cell_index(nb[0], nb[1], nb[2]); }
            if (idx[d] < g->n[d] - 1) { int nb[3] = {idx[0], idx[1], idx[2]}; nb[d]++; tmp[nt++] = g->cell_index(nb[0], nb[1], nb[2]); }
        }
        std::sort(tmp, tmp + nt);
        cols.insert(cols.end(), tmp, tmp + nt);
        M->h_rowptr[c + 1] = M->h_rowptr[c] + nt;
    }
    M->h_col = std::move(cols);
    M->nnzb = M->h_rowptr[J];
    M->rowptr.alloc((J + 1) * sizeof(int64_t));
    M->col.alloc(M->nnzb * sizeof(int));
    ctx->upload(M->rowptr.p, M->h_rowptr.data(), (J + 1) * sizeof(int64_t));
    ctx->upload(M->col.p, M->h_col.data(), M->nnzb * sizeof(int));
    M->vals.alloc((size_t)M->nnzb * Np * Np * sizeof(double));
    BH_HIP(hipMemsetAsync(M->vals.p, 0, M->vals.bytes, ctx->stream));
    sip_assemble(op, M.get());
    if (affine) {
        if (op->has_affine && !op->affine_valid) op->rebuild_affine();
        if (op->has_affine) {
            if (flags & BH_DEVICE_PTR) copy(ctx, affine, op->affine.as<double>(), op->N);
            else ctx->download(affine, op->affine.p, op->N * sizeof(double));
        } else {
            if (flags & BH_DEVICE_PTR) fill(ctx, affine, 0.0, op->N);
            else std::fill(affine, affine + op->N, 0.0);
        }
    }
    ctx->sync();
    *out = M.release();
    BH_API_END
}

// ------------------------------------------------------------------ BSR
int bh_bsr_create(bh_ctx* ctx, int nrows, int bs, const int64_t* rowptr, const int* col, const double* vals,
                  bh_bsr** out) {
    BH_API_BEGIN
    BH_CHECK(nrows >= 0 && bs >= 1 && bs <= 64, "bh_bsr_create: invalid sizes");
    auto M = std::make_unique<bh_bsr>();
    M->ctx = ctx;
    M->nrows = nrows;
    M->bs = bs;
    M->h_rowptr.assign(rowptr, rowptr + nrows + 1);
    M->nnzb = M->h_rowptr[nrows];
    BH_CHECK(M->h_rowptr[0] == 0, "bh_bsr_create: rowptr[0] must be 0");
    for (int r = 0; r < nrows; ++r) BH_CHECK(M->h_rowptr[r + 1] >= M->h_rowptr[r], "bh_bsr_create: rowptr not monotone");
    M->h_col.assign(col, col + M->nnzb);
    for (int64_t p = 0; p < M->nnzb; ++p) BH_CHECK(M->h_col[p] >= 0 && M->h_col[p] < nrows, "bh_bsr_create: column out of range");
    M->rowptr.alloc((nrows + 1) * sizeof(int64_t));
    M->col.alloc(std::max<int64_t>(M->nnzb, 1) * sizeof(int));
    ctx->upload(M->rowptr.p, M->h_rowptr.data(), (nrows + 1) * sizeof(int64_t));
    ctx->upload(M->col.p, M->h_col.data(), M->nnzb * sizeof(int));
    // blocks: row-major in, column-major on the device
    std::vector<double> v((size_t)M->nnzb * bs * bs);
    for (int64_t p = 0; p < M->nnzb; ++p)
        for (int i = 0; i < bs; ++i)
            for (int j = 0; j < bs; ++j) v[(size_t)p * bs * bs + (size_t)j * bs + i] = vals[(size_t)p * bs * bs + (size_t)i * bs + j];
    M->vals.alloc(std::max<size_t>(v.size(), 1) * sizeof(double));
    ctx->upload(M->vals.p, v.data(), v.size() * sizeof(double));
    *out = M.release();
    BH_API_END
}

}  // extern "C"

bh_bsr::~bh_bsr() { delete var; }

extern "C" {

int bh_bsr_create_v(bh_ctx* ctx, int nrows, const int64_t* rowptr, const int* col, const int64_t* blk_i0,
                    const int* blk_len, const double* vals, bh_bsr** out) {
    BH_API_BEGIN
    BH_CHECK(ctx && rowptr && col && blk_i0 && blk_len && out && nrows >= 0, "bh_bsr_create_v: null argument");
    BH_CHECK(rowptr[0] == 0, "bh_bsr_create_v: rowptr[0] must be 0");
    std::vector<int> sizes(blk_len, blk_len + nrows);
    int64_t frN = 0;
    for (int r = 0; r < nrows; ++r) {
        BH_CHECK(sizes[r] >= 0 && sizes[r] <= 64, "bh_bsr_create_v: block sizes must be in 0..64");
        BH_CHECK(blk_i0[r] >= 0, "bh_bsr_create_v: negative block offset");
        BH_CHECK(rowptr[r + 1] >= rowptr[r], "bh_bsr_create_v: rowptr not monotone");
        frN = std::max<int64_t>(frN, blk_i0[r] + sizes[r]);
    }
    // the rows may not overlap in the caller's layout
    std::vector<int> order(nrows);
    for (int r = 0; r < nrows; ++r) order[r] = r;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return blk_i0[a] < blk_i0[b]; });
    for (int i = 1; i < nrows; ++i)
        BH_CHECK(blk_i0[order[i - 1]] + sizes[order[i - 1]] <= blk_i0[order[i]] || sizes[order[i - 1]] == 0,
                 "bh_bsr_create_v: block rows overlap in the vector layout");
    const int64_t nnzb = rowptr[nrows];
    std::vector<std::vector<int>> cols(nrows);
    for (int r = 0; r < nrows; ++r)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            BH_CHECK(col[p] >= 0 && col[p] < nrows, "bh_bsr_create_v: column out of range");
            cols[r].push_back(col[p]);
        }
    auto M = std::make_unique<bh_bsr>();
    M->ctx = ctx;
    M->nrows = nrows;
    M->bs = 0;
    M->nnzb = nnzb;
    M->h_rowptr.assign(rowptr, rowptr + nrows + 1);
    M->h_col.assign(col, col + nnzb);
    M->var = new Vbsr();
    Vbsr& V = *M->var;
    V.build(ctx, nrows, sizes, cols);
    // values: the caller's blocks (row-major, every block listed) -> the kept blocks, column-major
    std::vector<double> v(std::max<int64_t>(V.nvals, 1), 0.0);
    int64_t src = 0, q = 0;
    for (int r = 0; r < nrows; ++r)
        for (int64_t p = rowptr[r]; p < rowptr[r + 1]; ++p) {
            const int R = sizes[r], C = sizes[col[p]];
            if (R > 0 && C > 0) {
                double* dst = &v[V.voff[q++]];
                for (int i = 0; i < R; ++i)
                    for (int j = 0; j < C; ++j) dst[(size_t)j * R + i] = vals[src + (size_t)i * C + j];
            }
            src += (int64_t)R * C;
        }
    ctx->upload(V.vals.p, v.data(), V.nvals * sizeof(double));
    // layouts: compressed row r at V.roff[r], caller's at blk_i0[r]
    M->fr_N = frN;
    M->fr_gaps = frN != V.N;
    for (int r = 0; r < nrows && !M->fr_gaps; ++r) M->fr_gaps = sizes[r] > 0 && blk_i0[r] != V.roff[r];
    std::vector<int64_t> fwd(std::max<int64_t>(V.N, 1)), inv(std::max<int64_t>(frN, 1), -1);
    for (int r = 0; r < nrows; ++r)
        for (int i = 0; i < sizes[r]; ++i) {
            fwd[V.roff[r] + i] = blk_i0[r] + i;
            inv[blk_i0[r] + i] = V.roff[r] + i;
        }
    M->fr_fwd.alloc(fwd.size() * sizeof(int64_t));
    M->fr_inv.alloc(inv.size() * sizeof(int64_t));
    ctx->upload(M->fr_fwd.p, fwd.data(), fwd.size() * sizeof(int64_t));
    ctx->upload(M->fr_inv.p, inv.data(), inv.size() * sizeof(int64_t));
    ctx->sync();
    *out = M.release();
    BH_API_END
}

int bh_bsr_destroy(bh_bsr* M) {
    BH_API_BEGIN
    if (M) (void)hipStreamSynchronize(M->ctx->stream);
    delete M;
    BH_API_END
}

int bh_bsr_info(bh_bsr* M, int* nrows, int* bs, int64_t* nnzb) {
    BH_API_BEGIN
    if (nrows) *nrows = M->nrows;
    if (bs) *bs = M->bs;   // 0: variable blocks
    if (nnzb) *nnzb = M->nnzb;
    BH_API_END
}

int bh_bsr_export(bh_bsr* M, int64_t* rowptr, int* col, double* vals) {
    BH_API_BEGIN
    BH_CHECK(!M->var, "bh_bsr_export: uniform-block matrices only");
    const int bs = M->bs;
    if (rowptr) std::copy(M->h_rowptr.begin(), M->h_rowptr.end(), rowptr);
    if (col) std::copy(M->h_col.begin(), M->h_col.end(), col);
    if (vals) {
        std::vector<double> v((size_t)M->nnzb * bs * bs);
        M->ctx->download(v.data(), M->vals.p, v.size() * sizeof(double));
        for (int64_t p = 0; p < M->nnzb; ++p)
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j) vals[(size_t)p * bs * bs + (size_t)i * bs + j] = v[(size_t)p * bs * bs + (size_t)j * bs + i];
    }
    BH_API_END
}

int bh_bsr_spmv(bh_bsr* M, double alpha, const void* x, double beta, void* y, int prec, int flags) {
    BH_API_BEGIN
    const bool dev = flags & BH_DEVICE_PTR;
    bh_ctx* ctx = M->ctx;
    if (M->var) {
        // variable blocks: vectors in the caller's layout; unused slots are zero rows
        Vbsr& V = *M->var;
        BH_CHECK(prec == BH_FP64 || (prec == BH_FP32 && dev), "bh_bsr_spmv: FP64, or FP32 on device float vectors");
        const size_t es = prec == BH_FP32 ? sizeof(float) : sizeof(double);
        Staged sx(ctx, x, M->fr_N * es, dev, true);
        Staged sy(ctx, y, M->fr_N * es, dev, beta != 0.0);
        if (!M->fr_gaps) {
            V.spmv(alpha, sx.dptr, beta, sy.dptr, prec);
        } else {
            if (M->fr_a.bytes < (size_t)V.N * es) { M->fr_a.alloc(std::max<int64_t>(V.N, 1) * es); M->fr_b.alloc(std::max<int64_t>(V.N, 1) * es); }
            const int64_t* fw = M->fr_fwd.as<int64_t>();
            const int64_t* iv = M->fr_inv.as<int64_t>();
            if (prec == BH_FP32) {
                frame_gather<float>(ctx, V.N, fw, (const float*)sx.dptr, M->fr_a.as<float>());
                if (beta != 0.0) frame_gather<float>(ctx, V.N, fw, (const float*)sy.dptr, M->fr_b.as<float>());
                V.spmv(alpha, M->fr_a.p, beta, M->fr_b.p, BH_FP32);
                frame_scatter<float>(ctx, M->fr_N, iv, M->fr_b.as<float>(), (float*)sy.dptr);
            } else {
                frame_gather<double>(ctx, V.N, fw, (const double*)sx.dptr, M->fr_a.as<double>());
                if (beta != 0.0) frame_gather<double>(ctx, V.N, fw, (const double*)sy.dptr, M->fr_b.as<double>());
                V.spmv(alpha, M->fr_a.p, beta, M->fr_b.p, BH_FP64);
                frame_scatter<double>(ctx, M->fr_N, iv, M->fr_b.as<double>(), (double*)sy.dptr);
            }
        }
        sy.download(y, M->fr_N * es, dev);
        if (!dev) ctx->sync();
        return 0;
    }
    const int64_t N = (int64_t)M->nrows * M->bs;
    if (prec == BH_FP64) {
        Staged sx(ctx, x, N * sizeof(double), dev, true);
        Staged sy(ctx, y, N * sizeof(double), dev, beta != 0.0);
        bsr_spmv(M, alpha, sx.dptr, beta, sy.dptr, BH_FP64);
        sy.download(y, N * sizeof(double), dev);
    } else if (prec == BH_FP32) {
        BH_CHECK(dev, "FP32 SpMV takes device float vectors");
        bsr_spmv(M, alpha, x, beta, y, BH_FP32);
    } else throw Error(-3, "bh_bsr_spmv: unsupported precision");
    if (!dev) ctx->sync();
    BH_API_END
}

// ------------------------------------------------------------------ solvers
int bh_bjacobi_create(bh_bsr* M, double omega, int iterations, int prec, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(prec == BH_FP64 || prec == BH_FP32, "bh_bjacobi_create: precision must be FP64 or FP32");
    BH_CHECK(!M->var, "bh_bjacobi_create: uniform-block matrices only (variable blocks: bh_xdg_precond_create_v)");
    *out = new BlockJacobiSolver(M, omega, iterations, prec);
    BH_API_END
}

static int make_gmres(bh_bsr* M, bh_op* op, bh_solver* P, int restart, double tol, int maxit, bool flex, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(M || op, "GMRES needs a matrix or an operator");
    BH_CHECK(restart >= 1 && restart <= 256, "GMRES restart length must be in 1..256");
    LinOp A;
    if (M && M->var) {
        A.V = M->var;
        A.N = M->var->N;
    } else {
        A.M = M;
        A.op = M ? nullptr : op;
        A.N = M ? (int64_t)M->nrows * M->bs : op->N;
    }
    bh_solver* Pc = P ? compressed(P) : nullptr;
    if (Pc) BH_CHECK(Pc->N == A.N, "preconditioner size mismatch");
    *out = in_frame(M, new GmresSolver(A, Pc, restart, tol, maxit, flex));
    BH_API_END
}

int bh_gmres_create(bh_bsr* M, bh_op* op, bh_solver* P, int restart, double tol, int maxit, bh_solver** out) {
    return make_gmres(M, op, P, restart, tol, maxit, false, out);
}
int bh_fgmres_create(bh_bsr* M, bh_op* op, bh_solver* P, int restart, double tol, int maxit, bh_solver** out) {
    return make_gmres(M, op, P, restart, tol, maxit, true, out);
}

int bh_pmg_create(bh_op* op, bh_bsr* fine_mat, const int* degrees, int ndeg, int min_coarse, double omega,
                  int pre, int post, int kcycle_inner, int prec, int matrix_free, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(prec == BH_FP64 || prec == BH_FP32 || prec == BH_FP16_MFMA,
             "bh_pmg_create: precision must be FP64, FP32 or FP16_MFMA");
    BH_CHECK(pre >= 1 && post >= 0, "bh_pmg_create: need pre >= 1, post >= 0");
    BH_CHECK(prec != BH_FP16_MFMA || matrix_free, "bh_pmg_create: FP16_MFMA needs the matrix-free fine level");
    auto mg = std::make_unique<PMultigrid>();
    mg->prec = prec == BH_FP16_MFMA ? BH_FP32 : prec;
    mg->mfma = prec == BH_FP16_MFMA;
    mg->omega = omega;
    mg->pre = pre;
    mg->post = post;
    mg->kin = kcycle_inner;
    mg->kcycle_level = kcycle_inner > 0 ? ndeg - 1 : -1;
    BH_CHECK(!fine_mat || !fine_mat->var, "bh_pmg_create: uniform-block fine matrix only");
    bh_bsr* Af = fine_mat;
    if (!Af) {
        BH_CHECK(bh_op_assemble(op, &Af, nullptr, 0) == 0, "bh_pmg_create: assembly failed: " + bh::last_error());
        mg->ownedAf = Af;
    }
    std::vector<int> deg(degrees, degrees + ndeg);
    mg->build(op, Af, deg, min_coarse, matrix_free != 0);
    *out = mg.release();
    BH_API_END
}

// ------------------------------------------------------------------ learned correction
int bh_mlp_create(bh_ctx* ctx, int n_in, int n_hidden, int n_out, uint64_t seed, bh_mlp** out) {
    BH_API_BEGIN
    BH_CHECK(n_in >= 1 && n_in <= 32 && n_out >= 1 && n_out <= 32, "bh_mlp_create: widths must be in 1..32");
    auto w = std::make_unique<bh_mlp>();
    w->m = new Mlp(ctx, n_in, n_hidden, n_out, seed);
    *out = w.release();
    BH_API_END
}
int bh_mlp_destroy(bh_mlp* w) {
    BH_API_BEGIN
    if (w) {
        (void)hipStreamSynchronize(w->m->ctx->stream);
        delete w->m;
        delete w;
    }
    BH_API_END
}
int bh_mlp_info(bh_mlp* w, int* n, int* in_p, int* out_p) {
    BH_API_BEGIN
    if (n) *n = w->m->nw;
    if (in_p) *in_p = w->m->in_p;
    if (out_p) *out_p = w->m->out_p;
    BH_API_END
}
int bh_mlp_get_params(bh_mlp* w, float* p) {
    BH_API_BEGIN
    w->m->ctx->download(p, w->m->params.p, w->m->nw * sizeof(float));
    BH_API_END
}
int bh_mlp_set_params(bh_mlp* w, const float* p) {
    BH_API_BEGIN
    w->m->set_params(p);
    w->m->ctx->sync();
    BH_API_END
}
int bh_mlp_forward(bh_mlp* w, int n, const float* X, float* Y, int flags) {
    BH_API_BEGIN
    Mlp* m = w->m;
    const bool dev = flags & BH_DEVICE_PTR;
    Staged sx(m->ctx, X, (size_t)n * m->in_p * sizeof(float), dev, true);
    Staged sy(m->ctx, Y, (size_t)n * m->out_p * sizeof(float), dev, false);
    m->infer(n, (const float*)sx.dptr, (float*)sy.dptr);
    sy.download(Y, (size_t)n * m->out_p * sizeof(float), dev);
    if (!dev) m->ctx->sync();
    BH_API_END
}
int bh_mlp_gradient(bh_mlp* w, int n, const float* X, const float* T, float* grad, double* loss, int flags) {
    BH_API_BEGIN
    Mlp* m = w->m;
    BH_CHECK(n >= 1, "bh_mlp_gradient: empty batch");
    const bool dev = flags & BH_DEVICE_PTR;
    Staged sx(m->ctx, X, (size_t)n * m->in_p * sizeof(float), dev, true);
    Staged st(m->ctx, T, (size_t)n * m->out_p * sizeof(float), dev, true);
    const double l = m->gradient(n, nullptr, (const float*)sx.dptr, (const float*)st.dptr);
    if (loss) *loss = l;
    if (grad) m->ctx->download(grad, m->grad.p, m->nw * sizeof(float));
    BH_API_END
}
int bh_mlp_gradient_noalloc(bh_mlp* w, int n, const float* X, const float* T, float* grad, double* loss) {
    BH_API_BEGIN
    BH_CHECK(w && w->m, "bh_mlp_gradient_noalloc: null network");
    BH_CHECK(n >= 1, "bh_mlp_gradient_noalloc: empty batch");
    BH_CHECK(X && T, "bh_mlp_gradient_noalloc: X and T must be device pointers");
    w->m->gradient_noalloc(n, X, T, grad, loss);
    BH_API_END
}
int bh_mlp_adam_step(bh_mlp* w, float lr) {
    BH_API_BEGIN
    w->m->adam(lr);
    BH_API_END
}
int bh_mlp_features(bh_op* op, const double* r, float* X, float* s, int in_p, int flags) {
    BH_API_BEGIN
    BH_CHECK(in_p >= op->Np + 4 * op->D, "bh_mlp_features: in_p too small (Np + 4 D features)");
    const bool dev = flags & BH_DEVICE_PTR;
    const int J = op->grid->J;
    Staged sr(op->ctx, r, op->N * sizeof(double), dev, true);
    Staged sx(op->ctx, X, (size_t)J * in_p * sizeof(float), dev, false);
    Staged ss(op->ctx, s, (size_t)J * sizeof(float), dev, false);
    mlp_features(op, sr.dptr, BH_FP64, (float*)sx.dptr, (float*)ss.dptr, in_p);
    sx.download(X, (size_t)J * in_p * sizeof(float), dev);
    ss.download(s, (size_t)J * sizeof(float), dev);
    if (!dev) op->ctx->sync();
    BH_API_END
}
int bh_mlp_targets(bh_op* op, const float* s, const double* e, float* T, int out_p, int flags) {
    BH_API_BEGIN
    BH_CHECK(out_p >= op->Np, "bh_mlp_targets: out_p too small");
    const bool dev = flags & BH_DEVICE_PTR;
    const int J = op->grid->J;
    Staged ss(op->ctx, s, (size_t)J * sizeof(float), dev, true);
    Staged se(op->ctx, e, op->N * sizeof(double), dev, true);
    Staged st(op->ctx, T, (size_t)J * out_p * sizeof(float), dev, false);
    mlp_targets(op, (const float*)ss.dptr, (const double*)se.dptr, out_p, (float*)st.dptr);
    st.download(T, (size_t)J * out_p * sizeof(float), dev);
    if (!dev) op->ctx->sync();
    BH_API_END
}
int bh_mlp_backward(bh_mlp* w, int n, const float* X, const float* dY, float* grad, int flags) {
    BH_API_BEGIN
    Mlp* m = w->m;
    BH_CHECK(n >= 1, "bh_mlp_backward: empty batch");
    const bool dev = flags & BH_DEVICE_PTR;
    Staged sx(m->ctx, X, (size_t)n * m->in_p * sizeof(float), dev, true);
    Staged sd(m->ctx, dY, (size_t)n * m->out_p * sizeof(float), dev, true);
    m->backward(n, (const float*)sx.dptr, (const float*)sd.dptr);
    if (grad) {
        if (dev) BH_HIP(hipMemcpyAsync(grad, m->grad.p, m->nw * sizeof(float), hipMemcpyDeviceToDevice, m->ctx->stream));
        else m->ctx->download(grad, m->grad.p, m->nw * sizeof(float));
    }
    if (!dev) m->ctx->sync();
    BH_API_END
}
int bh_mlp_correction(bh_op* op, const float* s, const float* Y, double* e, int out_p, int flags) {
    BH_API_BEGIN
    BH_CHECK(out_p >= op->Np && (out_p == 16 || out_p == 32), "bh_mlp_correction: out_p must be 16 or 32 and >= Np");
    const bool dev = flags & BH_DEVICE_PTR;
    const int J = op->grid->J;
    Staged ss(op->ctx, s, (size_t)J * sizeof(float), dev, true);
    Staged sy(op->ctx, Y, (size_t)J * out_p * sizeof(float), dev, true);
    Staged se(op->ctx, e, op->N * sizeof(double), dev, false);
    mlp_correction(op, (const float*)ss.dptr, (const float*)sy.dptr, out_p, se.dptr, BH_FP64);
    se.download(e, op->N * sizeof(double), dev);
    if (!dev) op->ctx->sync();
    BH_API_END
}
int bh_mlp_upstream(bh_op* op, const float* s, const double* g, double c, float* dY, int out_p, int flags) {
    BH_API_BEGIN
    BH_CHECK(out_p >= op->Np, "bh_mlp_upstream: out_p too small");
    const bool dev = flags & BH_DEVICE_PTR;
    const int J = op->grid->J;
    Staged ss(op->ctx, s, (size_t)J * sizeof(float), dev, true);
    Staged sg(op->ctx, g, op->N * sizeof(double), dev, true);
    Staged sd(op->ctx, dY, (size_t)J * out_p * sizeof(float), dev, false);
    mlp_upstream(op, (const float*)ss.dptr, (const double*)sg.dptr, c, out_p, (float*)sd.dptr);
    sd.download(dY, (size_t)J * out_p * sizeof(float), dev);
    if (!dev) op->ctx->sync();
    BH_API_END
}
int bh_pmg_set_learned(bh_solver* s, bh_mlp* w) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_learned: solver is not a multigrid preconditioner");
    mg->attach_learned(w ? w->m : nullptr);
    BH_API_END
}
int bh_orthomg_create(bh_solver* mgs, bh_bsr* M, bh_op* op, int pre, int post_sweeps, int coarse_reps, int max_hist,
                      double tol, int maxit, bh_solver** out) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(mgs);
    BH_CHECK(mg && op && out, "bh_orthomg_create: needs a multigrid preconditioner (bh_pmg_create) and the operator");
    BH_CHECK(!mg->equilib, "bh_orthomg_create: the block-equilibrated hierarchy is not supported");
    BH_CHECK(mg->lv.size() >= 2, "bh_orthomg_create: the multigrid needs a coarser level");
    BH_CHECK(!M || !M->var, "bh_orthomg_create: uniform-block matrices only");
    BH_CHECK(pre == 0 || pre == 1, "bh_orthomg_create: pre must be 0 or 1");
    BH_CHECK(post_sweeps >= 0 && coarse_reps >= 0 && pre + post_sweeps + coarse_reps > 0,
             "bh_orthomg_create: need at least one correction per iteration");
    BH_CHECK(max_hist >= 1 && max_hist <= 128, "bh_orthomg_create: max_hist must be in 1..128");
    BH_CHECK(maxit >= 1 && tol >= 0, "bh_orthomg_create: need maxit >= 1 and tol >= 0");
    LinOp A;
    A.M = M;
    A.op = M ? nullptr : op;
    A.N = M ? (int64_t)M->nrows * M->bs : op->N;
    BH_CHECK(A.N == mg->N && op->N == mg->N, "bh_orthomg_create: operator and multigrid sizes differ");
    *out = new OrthoMG(mg, A, op, pre, post_sweeps, coarse_reps, max_hist, tol, maxit);
    BH_API_END
}

int bh_orthomg_set_learned(bh_solver* s, bh_mlp* w) {
    BH_API_BEGIN
    OrthoMG* om = dynamic_cast<OrthoMG*>(s);
    BH_CHECK(om, "bh_orthomg_set_learned: solver is not an orthonormalization multigrid");
    om->attach_learned(w ? w->m : nullptr);
    BH_API_END
}

int bh_orthomg_info(bh_solver* s, int* info) {
    BH_API_BEGIN
    OrthoMG* om = dynamic_cast<OrthoMG*>(s);
    BH_CHECK(om && info, "bh_orthomg_info: solver is not an orthonormalization multigrid");
    info[0] = om->n_added;
    info[1] = om->n_cancel;
    info[2] = om->n_random;
    info[3] = om->n_restart;
    BH_API_END
}

int bh_pmg_learned_stats(bh_solver* s, int* acc, int* rej) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_learned_stats: solver is not a multigrid preconditioner");
    mg->learned_stats(acc, rej);
    BH_API_END
}
int bh_pmg_set_learned_fallback(bh_solver* s, int probe, double min_gain) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_learned_fallback: solver is not a multigrid preconditioner");
    BH_CHECK(probe >= 0 && std::isfinite(min_gain), "bh_pmg_set_learned_fallback: need probe >= 0 and a finite min_gain");
    mg->learned_probe = probe;
    mg->learned_min_gain = min_gain;
    BH_API_END
}
int bh_pmg_set_learned_cost_rule(bh_solver* s, int on) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_learned_cost_rule: solver is not a multigrid preconditioner");
    mg->learned_cost_rule = on != 0;
    BH_API_END
}
int bh_pmg_learned_costs(bh_solver* s, double* out) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_learned_costs: solver is not a multigrid preconditioner");
    BH_CHECK(out, "bh_pmg_learned_costs: out is NULL");
    out[0] = mg->learned_t_cycle;
    out[1] = mg->learned_t_net;
    out[2] = mg->learned_l_cycle;
    out[3] = mg->learned_l_net;
    BH_API_END
}
int bh_pmg_learned_info(bh_solver* s, int* active, double* mean_gain) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_learned_info: solver is not a multigrid preconditioner");
    if (active) *active = (mg->learned && !mg->learned_off) ? 1 : 0;
    if (mean_gain) {
        // mean over the probe window once it is closed, else over the calls so far
        *mean_gain = mg->learned_mean_gain;
        if (mg->learned && mg->lstat.p && (mg->learned_probe == 0 || mg->learned_calls < mg->learned_probe)) {
            double st[8];
            mg->ctx->download(st, mg->lstat.p, sizeof(st));
            int c[2];
            std::memcpy(c, &st[7], sizeof(c));
            const int calls = c[0] + c[1];
            *mean_gain = calls > 0 ? st[6] / calls : 0.0;
        }
    }
    BH_API_END
}

int bh_xdg_precond_coarse_export(bh_solver* s, int* nrows, int* bs, int64_t* nnzb, int64_t* rowptr, int* col, double* vals) {
    BH_API_BEGIN
    XdgPrecond* xp = dynamic_cast<XdgPrecond*>(compressed(s));
    BH_CHECK(xp && xp->Ac, "bh_xdg_precond_coarse_export: solver is not an XDG preconditioner");
    bh_bsr* M = xp->Ac;
    if (nrows) *nrows = M->nrows;
    if (bs) *bs = M->bs;
    if (nnzb) *nnzb = M->nnzb;
    if (rowptr || col || vals) {
        (void)hipStreamSynchronize(M->ctx->stream);
        BH_CHECK(bh_bsr_export(M, rowptr, col, vals) == 0, "bh_xdg_precond_coarse_export: export failed");
    }
    BH_API_END
}

int bh_xdg_precond_set_fp64_blocks(bh_solver* s, int on) {
    BH_API_BEGIN
    XdgPrecond* xp = dynamic_cast<XdgPrecond*>(compressed(s));
    BH_CHECK(xp, "bh_xdg_precond_set_fp64_blocks: solver is not an XDG preconditioner");
    xp->Dc->fp64_inverse = on != 0;
    BH_API_END
}

int bh_pmg_set_cell_ilu(bh_solver* s, int ilu_level) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_cell_ilu: solver is not a multigrid preconditioner");
    (void)hipStreamSynchronize(mg->ctx->stream);
    mg->set_cell_ilu(ilu_level);
    BH_API_END
}

int bh_pmg_level_info(bh_solver* s, int level, int* info, int64_t* nnzb) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_level_info: solver is not a multigrid preconditioner");
    const int nl = (int)mg->lv.size();
    if (level < 0) {   // the number of levels
        if (info) info[0] = nl;
        return 0;
    }
    BH_CHECK(level < nl, "bh_pmg_level_info: level out of range");
    const MgLevel& L = mg->lv[level];
    if (info) {
        info[0] = L.J;
        info[1] = L.bs;
        info[2] = level + 1 < nl ? L.kind : -1;
        info[3] = L.deg;
        for (int d = 0; d < 3; ++d) info[4 + d] = L.n[d];
    }
    if (nnzb) *nnzb = L.A ? L.A->nnzb : 0;
    BH_API_END
}

int bh_pmg_level_transfer(bh_solver* s, int level, int* parent, double* P) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_level_transfer: solver is not a multigrid preconditioner");
    BH_CHECK(level >= 0 && level + 1 < (int)mg->lv.size() && mg->lv[level].kind == 1,
             "bh_pmg_level_transfer: not an aggregation level");
    const MgLevel& L = mg->lv[level];
    const int bs = L.bs;
    std::vector<int> par(L.J), pidx(L.J);
    mg->ctx->download(par.data(), L.parent.p, L.J * sizeof(int));
    mg->ctx->download(pidx.data(), L.pidx.p, L.J * sizeof(int));
    const size_t nP = L.P64.bytes / sizeof(double);
    std::vector<double> Pd(nP);
    mg->ctx->download(Pd.data(), L.P64.p, nP * sizeof(double));
    if (parent) std::copy(par.begin(), par.end(), parent);
    if (P)
        for (int i = 0; i < L.J; ++i)
            std::copy(Pd.begin() + (size_t)pidx[i] * bs * bs, Pd.begin() + (size_t)(pidx[i] + 1) * bs * bs, P + (size_t)i * bs * bs);
    BH_API_END
}

int bh_pmg_level_matrix(bh_solver* s, int level, int64_t* rowptr, int* col, double* vals) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_level_matrix: solver is not a multigrid preconditioner");
    BH_CHECK(level >= 0 && level < (int)mg->lv.size() && mg->lv[level].A, "bh_pmg_level_matrix: no assembled matrix on this level");
    (void)hipStreamSynchronize(mg->ctx->stream);
    BH_CHECK(bh_bsr_export(mg->lv[level].A, rowptr, col, vals) == 0, "bh_pmg_level_matrix: export failed");
    BH_API_END
}

int bh_pmg_smooth(bh_solver* s, int level, const double* r, double* x) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_smooth: solver is not a multigrid preconditioner");
    BH_CHECK(level >= 0 && level + 1 < (int)mg->lv.size(), "bh_pmg_smooth: level out of range");
    MgLevel& L = mg->lv[level];
    bh_ctx* ctx = mg->ctx;
    DevBuf d(L.N * sizeof(double));
    ctx->upload(d.p, r, L.N * sizeof(double));
    if (mg->prec == BH_FP32) {
        convert_f64_f32(ctx, d.as<double>(), (float*)L.b.p, L.N);
        mg->smooth<float>(level, (const float*)L.b.p, (float*)L.x.p, false);
        convert_f32_f64(ctx, (const float*)L.x.p, d.as<double>(), L.N);
    } else {
        copy(ctx, (double*)L.b.p, d.as<double>(), L.N);
        mg->smooth<double>(level, (const double*)L.b.p, (double*)L.x.p, false);
        copy(ctx, d.as<double>(), (const double*)L.x.p, L.N);
    }
    ctx->download(x, d.p, L.N * sizeof(double));
    BH_API_END
}

int bh_pmg_set_level_pmg(bh_solver* s, int enable) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_level_pmg: solver is not a multigrid preconditioner");
    (void)hipStreamSynchronize(mg->ctx->stream);
    mg->set_level_pmg(enable != 0);
    BH_API_END
}

int bh_pmg_set_schwarz(bh_solver* s, int enable) { return bh_pmg_set_schwarz_overlap(s, enable, 0, 0); }

int bh_pmg_set_schwarz_overlap(bh_solver* s, int enable, int overlap, int overlap_scaling) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    XdgPrecond* xp = dynamic_cast<XdgPrecond*>(compressed(s));
    BH_CHECK(mg || xp, "bh_pmg_set_schwarz: solver is not a multigrid preconditioner");
    BH_CHECK(overlap >= 0 && overlap <= 2, "bh_pmg_set_schwarz: overlap must be 0, 1 or 2");
    if (xp) mg = xp->cmg.get();
    (void)hipStreamSynchronize(mg->ctx->stream);
    mg->set_schwarz(enable != 0, overlap, overlap_scaling != 0);
    BH_API_END
}

int bh_pmg_set_diag_equilib(bh_solver* s) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    BH_CHECK(mg, "bh_pmg_set_diag_equilib: solver is not a multigrid preconditioner");
    (void)hipStreamSynchronize(mg->ctx->stream);
    mg->set_equilib();
    BH_API_END
}

int bh_pmg_set_graphs(bh_solver* s, int enable) {
    BH_API_BEGIN
    PMultigrid* mg = dynamic_cast<PMultigrid*>(s);
    XdgPrecond* xp = dynamic_cast<XdgPrecond*>(compressed(s));
    BH_CHECK(mg || xp, "bh_pmg_set_graphs: solver is not a multigrid preconditioner");
    if (xp) mg = xp->cmg.get();
    (void)hipStreamSynchronize(mg->ctx->stream);
    mg->drop_graphs();
    mg->use_graphs = enable != 0;
    BH_API_END
}

// ------------------------------------------------------------------ Euler
int bh_euler_create(bh_grid* g, int k, int nq1, double gamma, double mach, bh_euler** out) {
    BH_API_BEGIN
    BH_CHECK(gamma > 1.0 && mach > 0.0, "bh_euler_create: need gamma > 1, mach > 0");
    auto w = std::make_unique<bh_euler>();
    w->e = new EulerOp(g, k, nq1, gamma, mach);
    *out = w.release();
    BH_API_END
}
int bh_euler_destroy(bh_euler* w) {
    BH_API_BEGIN
    if (w) {
        (void)hipStreamSynchronize(w->e->ctx->stream);
        delete w->e;
        delete w;
    }
    BH_API_END
}
int bh_euler_info(bh_euler* w, int64_t* N, int* Np) {
    BH_API_BEGIN
    if (N) *N = w->e->N;
    if (Np) *Np = w->e->Np;
    BH_API_END
}
int bh_euler_rhs(bh_euler* w, const void* U, void* Rv, int prec, int flags) {
    BH_API_BEGIN
    EulerOp* e = w->e;
    const bool dev = flags & BH_DEVICE_PTR;
    BH_CHECK(prec == BH_FP64 || dev, "bh_euler_rhs: FP32 takes device float vectors");
    Staged su(e->ctx, U, e->N * sizeof(double), dev, true);
    Staged sr(e->ctx, Rv, e->N * sizeof(double), dev, false);
    e->rhs(su.dptr, sr.dptr, prec);
    sr.download(Rv, e->N * sizeof(double), dev);
    if (!dev) e->ctx->sync();
    BH_API_END
}
int bh_euler_implicit_step(bh_euler* w, const double* Un, double* U, double dt, double conv, int maxit, int restart,
                           int* info, double* res, int flags) {
    BH_API_BEGIN
    BH_CHECK(w && Un && U, "bh_euler_implicit_step: null argument");
    BH_CHECK(dt > 0 && conv > 0 && maxit >= 1 && restart >= 1 && restart <= 200,
             "bh_euler_implicit_step: need dt > 0, conv > 0, maxit >= 1, 1 <= restart <= 200");
    EulerOp* e = w->e;
    const bool dev = flags & BH_DEVICE_PTR;
    Staged sn(e->ctx, Un, e->N * sizeof(double), dev, true);
    Staged su(e->ctx, U, e->N * sizeof(double), dev, true);
    newton_euler(e, (const double*)sn.dptr, (double*)su.dptr, dt, conv, maxit, restart, info, res);
    su.download(U, e->N * sizeof(double), dev);
    if (!dev) e->ctx->sync();
    BH_API_END
}
int bh_euler_tvd3(bh_euler* w, void* U, double dt, int steps, int prec, int flags) {
    BH_API_BEGIN
    EulerOp* e = w->e;
    const bool dev = flags & BH_DEVICE_PTR;
    BH_CHECK(prec == BH_FP64 || dev, "bh_euler_tvd3: FP32 takes device float vectors");
    Staged su(e->ctx, U, e->N * sizeof(double), dev, true);
    e->tvd3(su.dptr, dt, steps, prec);
    su.download(U, e->N * sizeof(double), dev);
    if (!dev) e->ctx->sync();
    BH_API_END
}

int bh_solver_solve(bh_solver* s, double* x, const double* b, int flags) {
    BH_API_BEGIN
    const bool dev = flags & BH_DEVICE_PTR;
    const int64_t N = s->N;
    Staged sx(s->ctx, x, N * sizeof(double), dev, true);
    Staged sb(s->ctx, b, N * sizeof(double), dev, true);
    s->solve_dev((double*)sx.dptr, (const double*)sb.dptr);
    sx.download(x, N * sizeof(double), dev);
    s->ctx->sync();
    if (!std::isfinite(s->res)) throw Error(-5, "solver: NaN/Inf residual");  // CheckAndGetNorm, AS/SoftGMRES.cs:462-470
    BH_API_END
}

int bh_solver_stats(bh_solver* s, int* it, int* conv, double* r0, double* r) {
    BH_API_BEGIN
    if (it) *it = s->iterations;
    if (conv) *conv = s->converged;
    if (r0) *r0 = s->res0;
    if (r) *r = s->res;
    BH_API_END
}

int bh_solver_destroy(bh_solver* s) {
    BH_API_BEGIN
    if (s) (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
    BH_API_END
}

}  // extern "C"

// ------------------------------------------------------------------ XDG two-phase Poisson
namespace bh {
void xdg_build(bh_xdg* X);
void xdg_set_rhs(bh_xdg* X, const double* f, const double* g);
void xdg_extend(const bh_xdg* X, const double* ured, double* ufull);
}

static bh_xdg* xdg_make(bh_grid* g, int k, double nu_in, double nu_out, double penalty_base, double agg_threshold,
                        int quad_order, const std::function<void(XdgProblem&)>& set_ls) {
    BH_CHECK(k >= 1 && k <= 3, "bh_xdg_create: degree must be 1..3");
    BH_CHECK(!(g->D == 3 && k == 3), "bh_xdg_create: D = 3 supports k = 1, 2 (the element kernels hold 4 Np^2 <= 448 entries per wave)");
    BH_CHECK(agg_threshold >= 0 && agg_threshold < 0.5, "bh_xdg_create: agglomeration threshold must be in [0, 0.5)");
    auto X = std::make_unique<bh_xdg>();
    X->ctx = g->ctx;
    X->grid = g;
    X->P = std::make_unique<XdgProblem>();
    XdgProblem& p = *X->P;
    p.ctx = g->ctx;
    p.g = g;
    p.k = k;
    p.q = quad_order > 0 ? quad_order : k + 1;
    set_ls(p);
    p.nu[0] = nu_in;
    p.nu[1] = nu_out;
    p.penalty_base = penalty_base;
    p.agg_threshold = agg_threshold;
    p.setup();
    BH_CHECK(p.n_failed == 0, "bh_xdg_create: agglomeration failed for " + std::to_string(p.n_failed) + " cells");
    if (X->ctx) xdg_build(X.get());
    return X.release();
}

int bh_xdg_create(bh_grid* g, int k, const double* center, double radius, double nu_in, double nu_out,
                  double penalty_base, double agg_threshold, int quad_order, bh_xdg** out) {
    BH_API_BEGIN
    BH_CHECK(g && center && out, "bh_xdg_create: null argument");
    BH_CHECK(radius > 0, "bh_xdg_create: radius must be positive");
    *out = xdg_make(g, k, nu_in, nu_out, penalty_base, agg_threshold, quad_order, [&](XdgProblem& p) {
        for (int d = 0; d < 3; ++d) p.S.c[d] = d < g->D ? center[d] : 0.0;
        p.S.r = radius;
        p.ls_kind = 0;
    });
    BH_API_END
}

int bh_xdg_create_poly(bh_grid* g, int k, int deg, const double* coeffs, double nu_in, double nu_out, double penalty_base,
                       double agg_threshold, int quad_order, bh_xdg** out) {
    BH_API_BEGIN
    BH_CHECK(g && coeffs && out, "bh_xdg_create_poly: null argument");
    BH_CHECK(deg >= 1 && deg <= 8, "bh_xdg_create_poly: degree per variable must be 1..8");
    int ncoef = 1;
    for (int d = 0; d < g->D; ++d) ncoef *= deg + 1;
    *out = xdg_make(g, k, nu_in, nu_out, penalty_base, agg_threshold, quad_order, [&](XdgProblem& p) {
        p.ls_kind = 1;
        p.L.D = g->D;
        p.L.n = deg;
        p.L.c.assign(coeffs, coeffs + ncoef);
        for (int d = 0; d < 3; ++d) p.S.c[d] = 0.0;
        p.S.r = 0.0;
    });
    BH_API_END
}

int bh_xdg_cut_rule_fallbacks(long* count) {
    BH_API_BEGIN
    BH_CHECK(count, "bh_xdg_cut_rule_fallbacks: null argument");
    *count = poly_fallbacks();
    BH_API_END
}

int bh_xdg_destroy(bh_xdg* X) {
    BH_API_BEGIN
    if (X && X->ctx) (void)hipStreamSynchronize(X->ctx->stream);
    delete X;
    BH_API_END
}

int bh_xdg_info(bh_xdg* X, int64_t* info) {
    BH_API_BEGIN
    const XdgProblem& p = *X->P;
    info[0] = p.N_full;
    info[1] = p.N_red;
    info[2] = p.ncut;
    info[3] = (int64_t)p.agg_pairs.size() / 3;
    info[4] = p.n_failed;
    info[5] = (int64_t)p.elems.size();
    info[6] = (int64_t)p.nw.size();
    info[7] = X->Afull.nvals;
    info[8] = X->Ared ? X->Ared->nnzb * p.Np * p.Np : 0;
    info[9] = (int64_t)X->bnd_w.size();
    info[10] = p.Np;
    info[11] = p.g->J;
    info[12] = X->Ared ? X->Ared->nrows : 0;
    BH_API_END
}

int bh_xdg_cells(bh_xdg* X, int* spmask, double* vol, double* iface_area, double* len) {
    BH_API_BEGIN
    const XdgProblem& p = *X->P;
    const int J = p.g->J;
    if (spmask) std::copy(p.spmask.begin(), p.spmask.end(), spmask);
    for (int s = 0; s < 2; ++s) {
        if (vol) std::copy(p.vol[s].begin(), p.vol[s].end(), vol + (size_t)s * J);
        if (len) std::copy(p.len[s].begin(), p.len[s].end(), len + (size_t)s * J);
    }
    if (iface_area) std::copy(p.iface_area.begin(), p.iface_area.end(), iface_area);
    BH_API_END
}

int bh_xdg_mass_matrix(bh_xdg* X, double* out) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "bh_xdg_mass_matrix: needs a device context");
    const int64_t n = X->P->N_full * X->P->Np;
    DevBuf d(std::max<int64_t>(n, 1) * sizeof(double));
    xdg_mass(X->X, d.as<double>());
    X->ctx->download(out, d.p, n * sizeof(double));
    BH_API_END
}

int bh_xdg_agglomeration(bh_xdg* X, int* pairs) {
    BH_API_BEGIN
    std::copy(X->P->agg_pairs.begin(), X->P->agg_pairs.end(), pairs);
    BH_API_END
}

int bh_xdg_elements(bh_xdg* X, int* ints, double* dbls, double* nx, double* nw, double* nn) {
    BH_API_BEGIN
    const XdgProblem& p = *X->P;
    for (size_t e = 0; e < p.elems.size(); ++e) {
        const XElem& E = p.elems[e];
        if (ints) {
            int* o = ints + 9 * e;
            o[0] = E.type; o[1] = E.cellP; o[2] = E.cellQ; o[3] = E.spP; o[4] = E.spQ;
            o[5] = E.dir; o[6] = E.side; o[7] = E.r0; o[8] = E.rn;
        }
        if (dbls) { dbls[3 * e] = E.nuP; dbls[3 * e + 1] = E.nuQ; dbls[3 * e + 2] = E.mu; }
    }
    if (nx) std::copy(p.nx.begin(), p.nx.end(), nx);
    if (nw) std::copy(p.nw.begin(), p.nw.end(), nw);
    if (nn) std::copy(p.nn.begin(), p.nn.end(), nn);
    BH_API_END
}

int bh_xdg_matrix(bh_xdg* X, int which, int64_t* dof_off, int64_t* rowptr, int* col, int64_t* voff, double* vals) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    BH_CHECK(which == 0 || which == 1, "bh_xdg_matrix: which must be 0 (full) or 1 (agglomerated)");
    if (which == 0) {
        const Vbsr& M = X->Afull;
        if (dof_off) std::copy(M.roff.begin(), M.roff.end(), dof_off);
        if (rowptr) std::copy(M.rowptr.begin(), M.rowptr.end(), rowptr);
        if (col) std::copy(M.col.begin(), M.col.end(), col);
        if (voff) std::copy(M.voff.begin(), M.voff.end(), voff);
        if (vals) X->ctx->download(vals, M.vals.p, M.nvals * sizeof(double));
    } else {
        // species units of Np rows: uniform blocks
        const bh_bsr& M = *X->Ared;
        const int64_t b2 = (int64_t)M.bs * M.bs;
        for (int u = 0; dof_off && u <= M.nrows; ++u) dof_off[u] = (int64_t)u * M.bs;
        if (rowptr) std::copy(M.h_rowptr.begin(), M.h_rowptr.end(), rowptr);
        if (col) std::copy(M.h_col.begin(), M.h_col.end(), col);
        for (int64_t p = 0; voff && p <= M.nnzb; ++p) voff[p] = p * b2;
        if (vals) X->ctx->download(vals, M.vals.p, M.nnzb * b2 * sizeof(double));
    }
    BH_API_END
}

int bh_xdg_bnd_nodes(bh_xdg* X, double* xyz) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    std::copy(X->bnd_x.begin(), X->bnd_x.end(), xyz);
    BH_API_END
}

int bh_xdg_set_rhs(bh_xdg* X, const double* f, const double* g) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    BH_CHECK(f, "bh_xdg_set_rhs: source values needed");
    xdg_set_rhs(X, f, g);
    BH_API_END
}

int bh_xdg_rhs(bh_xdg* X, int which, double* b, int flags) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    const XdgProblem& p = *X->P;
    if (which == 0) {
        BH_CHECK(!(flags & BH_DEVICE_PTR), "bh_xdg_rhs: the full right-hand side is a host vector");
        BH_CHECK((int64_t)X->rhs_full.size() == p.N_full, "bh_xdg_rhs: no right-hand side set");
        std::copy(X->rhs_full.begin(), X->rhs_full.end(), b);
    } else {
        BH_CHECK((int64_t)X->rhs_full.size() == p.N_full, "bh_xdg_rhs: no right-hand side set");
        if (flags & BH_DEVICE_PTR) BH_HIP(hipMemcpyAsync(b, X->rhs_red.p, p.N_red * sizeof(double), hipMemcpyDeviceToDevice, X->ctx->stream));
        else X->ctx->download(b, X->rhs_red.p, p.N_red * sizeof(double));
        X->ctx->sync();
    }
    BH_API_END
}

int bh_xdg_spmv(bh_xdg* X, int which, double alpha, const void* x, double beta, void* y, int prec, int flags) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    BH_CHECK(which == 0 || which == 1, "bh_xdg_spmv: which must be 0 (full) or 1 (agglomerated)");
    BH_CHECK(prec == BH_FP64 || prec == BH_FP32, "bh_xdg_spmv: precision must be FP64 or FP32");
    const bool dev = flags & BH_DEVICE_PTR;
    const size_t es = prec == BH_FP32 ? sizeof(float) : sizeof(double);
    const int64_t N = which ? X->P->N_red : X->P->N_full;
    Staged sx(X->ctx, x, N * es, dev, true);
    Staged sy(X->ctx, y, N * es, dev, beta != 0.0);
    if (which) bsr_spmv(X->Ared.get(), alpha, sx.dptr, beta, sy.dptr, prec);
    else X->Afull.spmv(alpha, sx.dptr, beta, sy.dptr, prec);
    sy.download(y, N * es, dev);
    X->ctx->sync();
    BH_API_END
}

int bh_xdg_evaluate(bh_xdg* X, double alpha, const double* u, double beta, double* y, int flags) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    const bool dev = flags & BH_DEVICE_PTR;
    const int64_t N = X->P->N_full;
    Staged su(X->ctx, u, N * sizeof(double), dev, true);
    Staged sy(X->ctx, y, N * sizeof(double), dev, beta != 0.0);
    xdg_evaluate(X->X, alpha, (const double*)su.dptr, beta, (double*)sy.dptr);
    sy.download(y, N * sizeof(double), dev);
    X->ctx->sync();
    BH_API_END
}

int bh_xdg_extend(bh_xdg* X, const double* ured, double* ufull) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    xdg_extend(X, ured, ufull);
    BH_API_END
}

int bh_xdg_precond_create(bh_xdg* X, int coarse_degree, int min_coarse, double omega, int pre, int post,
                          int kcycle_inner, int prec, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    BH_CHECK(prec == BH_FP64 || prec == BH_FP32, "bh_xdg_precond_create: precision must be FP64 or FP32");
    BH_CHECK(pre >= 1 && post >= 0, "bh_xdg_precond_create: need pre >= 1, post >= 0");
    BH_CHECK(coarse_degree >= 0 && coarse_degree <= X->P->k, "bh_xdg_precond_create: coarse degree out of range");
    auto P = std::make_unique<XdgPrecond>();
    P->A = X->Ared.get();
    P->Dc = &X->Dcell;
    P->prec = prec;
    P->Np = X->P->Np;
    P->omega = omega;
    P->pre = pre;
    P->post = post;
    P->build(X->grid, coarse_degree, min_coarse, kcycle_inner);
    *out = P.release();
    BH_API_END
}

int bh_xdg_precond_create_v(bh_grid* g, bh_bsr* M, int Np, int coarse_degree, int min_coarse, double omega, int pre,
                            int post, int kcycle_inner, int prec, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(g && M && out, "bh_xdg_precond_create_v: null argument");
    BH_CHECK(M->var, "bh_xdg_precond_create_v: needs a variable-block matrix (bh_bsr_create_v)");
    BH_CHECK(M->nrows == g->J, "bh_xdg_precond_create_v: one block row per grid cell expected");
    BH_CHECK(g->ctx, "bh_xdg_precond_create_v: the grid needs a device context");
    BH_CHECK(prec == BH_FP64 || prec == BH_FP32, "bh_xdg_precond_create_v: precision must be FP64 or FP32");
    BH_CHECK(pre >= 1 && post >= 0, "bh_xdg_precond_create_v: need pre >= 1, post >= 0");
    BH_CHECK(Np >= 1 && coarse_degree >= 0, "bh_xdg_precond_create_v: invalid Np or coarse degree");
    BH_CHECK(basis_count(g->D, coarse_degree) <= Np, "bh_xdg_precond_create_v: coarse degree above the DG degree");
    auto P = std::make_unique<XdgPrecond>();
    P->Av = M->var;
    P->Dc = M->var;
    P->prec = prec;
    P->Np = Np;
    P->omega = omega;
    P->pre = pre;
    P->post = post;
    P->build(g, coarse_degree, min_coarse, kcycle_inner);
    *out = in_frame(M, P.release());
    BH_API_END
}

int bh_xdg_dof_frame(bh_xdg* X, int which, int64_t* frame_i0, int* len, int64_t* compressed_i0) {
    BH_API_BEGIN
    BH_CHECK(which == 0 || which == 1, "bh_xdg_dof_frame: which must be 0 (full) or 1 (agglomerated)");
    const XdgProblem& p = *X->P;
    const int J = p.g->J;
    const std::vector<int64_t>& off = which ? p.off_red : p.off_full;
    const int nmax = 2 * p.Np;   // Np times the number of species (XDGBasis.MaximalLength)
    for (int j = 0; j < J; ++j) {
        if (frame_i0) frame_i0[j] = (int64_t)j * nmax;
        if (len) len[j] = (int)(off[j + 1] - off[j]);
        if (compressed_i0) compressed_i0[j] = off[j];
    }
    BH_API_END
}

int bh_xdg_fgmres_create(bh_xdg* X, bh_solver* P, int restart, double tol, int maxit, bh_solver** out) {
    BH_API_BEGIN
    BH_CHECK(X->ctx, "XDG: this call needs a grid created with a device context");
    BH_CHECK(restart >= 1 && restart <= 256, "GMRES restart length must be in 1..256");
    LinOp A;
    A.M = X->Ared.get();
    A.N = X->P->N_red;
    if (P) BH_CHECK(P->N == A.N, "preconditioner size mismatch");
    *out = new GmresSolver(A, P, restart, tol, maxit, true);
    BH_API_END
}


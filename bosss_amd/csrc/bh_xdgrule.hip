// Cut-cell, interface and cut-face rules on the device, one thread per rule: for a
// sphere level set the construction of bh_sphrule.hpp, for a polynomial one that of
// bh_sayerule.hpp (the host set-up runs the same code).  Two passes: the node count of every rule, then, at the offsets of their
// prefix sum, the nodes themselves, so every rule lands in its own slice in the
// order the items are listed and the result does not depend on the launch.
#include <vector>

#include "bh_internal.hpp"
#include "bh_sayerule.hpp"
#include "bh_sphrule.hpp"
#include "bh_xdg.hpp"

namespace bh {

namespace {

struct CountEmit {
    int64_t n = 0;
    __device__ void operator()(const double*, double, const double*) { ++n; }
};

struct FillEmit {
    double* X;
    double* W;
    double* NRM;   // interface rules only
    int64_t i;
    __device__ void operator()(const double* x, double w, const double* nrm) {
        for (int d = 0; d < 3; ++d) X[3 * i + d] = x[d];
        W[i] = w;
        if (NRM)
            for (int d = 0; d < 3; ++d) NRM[3 * i + d] = nrm ? nrm[d] : 0.0;
        ++i;
    }
};

template <typename E>
__device__ bool run_item(int D, const SphereRuleItem& it, const sph::Ball& S, const double* gx, const double* gw, int q, E& e) {
    switch (it.kind) {
        case 0: return sph::volume_or_interface(D, it.lo, it.hi, S, -1, 0, gx, gw, q, e);
        case 1: return sph::volume_or_interface(D, it.lo, it.hi, S, +1, 0, gx, gw, q, e);
        case 2: return sph::volume_or_interface(D, it.lo, it.hi, S, 0, 1, gx, gw, q, e);
        case 3: return sph::face(D, it.lo, it.hi, it.dir, S, -1, gx, gw, q, e);
        default: return sph::face(D, it.lo, it.hi, it.dir, S, +1, gx, gw, q, e);
    }
}

__global__ __launch_bounds__(64) void rule_count_k(int n, const SphereRuleItem* __restrict__ items, int D, sph::Ball S,
                                                   const double* __restrict__ gxw, int q, int64_t* __restrict__ counts,
                                                   int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CountEmit e;
    const bool ok = run_item(D, items[i], S, gxw, gxw + q, q, e);
    counts[i] = e.n;
    if (!ok) bad[0] = 1;
}

__global__ __launch_bounds__(64) void rule_fill_k(int n, const SphereRuleItem* __restrict__ items, int D, sph::Ball S,
                                                  const double* __restrict__ gxw, int q, const int64_t* __restrict__ off,
                                                  double* __restrict__ X, double* __restrict__ W, double* __restrict__ NRM) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FillEmit e{X, W, items[i].kind == 2 ? NRM : nullptr, off[i]};
    run_item(D, items[i], S, gxw, gxw + q, q, e);
}

// polynomial level set: power-basis coefficients c, D directions, degree n <= POLY_NM
constexpr int POLY_NM = 4;
// threads per workgroup of the polynomial rule kernels: a rule's construction is long
// and branchy, and the launches hold a few thousand rules, so the rules are spread
// over more waves with fewer lanes each instead of filling waves
constexpr int POLY_TPB = 16;
struct PolyArg {
    const double* c;
    int D, n;
};

template <typename E>
__device__ bool run_poly_item(const PolyArg& L, const SphereRuleItem& it, saye::Ctx& ctx, E& e) {
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = it.lo[d]; hi[d] = it.hi[d]; }
    int sp = 0;
    bool iface = false;
    switch (it.kind) {
        case 0: sp = -1; break;
        case 1: sp = +1; break;
        case 2: iface = true; break;
        case 3: sp = -1; hi[it.dir] = lo[it.dir]; break;
        default: sp = +1; hi[it.dir] = lo[it.dir]; break;
    }
    return saye::rule<POLY_NM>(L.c, L.D, L.n, lo, hi, sp, iface, ctx, e);
}

__global__ __launch_bounds__(64) void poly_count_k(int n, const SphereRuleItem* __restrict__ items, PolyArg L,
                                                   const double* __restrict__ gtw, int q, int64_t* __restrict__ counts,
                                                   int* __restrict__ fallbacks, int* __restrict__ bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CountEmit e;
    saye::Ctx ctx{gtw, gtw + q, q, 0, false};
    const bool ok = run_poly_item(L, items[i], ctx, e);
    counts[i] = e.n;
    fallbacks[i] = ctx.fallbacks;
    if (!ok) bad[0] = 1;
}

__global__ __launch_bounds__(64) void poly_fill_k(int n, const SphereRuleItem* __restrict__ items, PolyArg L,
                                                  const double* __restrict__ gtw, int q, const int64_t* __restrict__ off,
                                                  double* __restrict__ X, double* __restrict__ W, double* __restrict__ NRM) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    FillEmit e{X, W, items[i].kind == 2 ? NRM : nullptr, off[i]};
    saye::Ctx ctx{gtw, gtw + q, q, 0, false};
    run_poly_item(L, items[i], ctx, e);
}

__global__ __launch_bounds__(64) void poly_classify_k(int n, const SphereRuleItem* __restrict__ items, PolyArg L,
                                                      int* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = saye::classify<POLY_NM>(L.c, L.D, L.n, items[i].lo, items[i].hi);
}

// rule slices of the items from the node arrays at the prefix offsets
void unpack_rules(bh_ctx* ctx, const std::vector<SphereRuleItem>& items, const std::vector<int64_t>& off, const DevBuf& dX,
                  const DevBuf& dW, const DevBuf& dN, std::vector<QRule>& out) {
    const int n = (int)items.size();
    const int64_t tot = off[n];
    std::vector<double> X(3 * tot), W(tot), NRM(3 * tot);
    ctx->download(X.data(), dX.p, X.size() * sizeof(double));
    ctx->download(W.data(), dW.p, W.size() * sizeof(double));
    ctx->download(NRM.data(), dN.p, NRM.size() * sizeof(double));
    for (int i = 0; i < n; ++i) {
        QRule& R = out[i];
        const int64_t a = off[i], b = off[i + 1];
        R.x.assign(X.begin() + 3 * a, X.begin() + 3 * b);
        R.w.assign(W.begin() + a, W.begin() + b);
        if (items[i].kind == 2) R.nrm.assign(NRM.begin() + 3 * a, NRM.begin() + 3 * b);
    }
}

}  // namespace

bool poly_rules_on_device(const PolyLS& L) { return L.n >= 0 && L.n <= POLY_NM; }

void poly_rules_device(bh_ctx* ctx, const PolyLS& L, int q, const std::vector<SphereRuleItem>& items, std::vector<QRule>& out) {
    BH_CHECK(q >= 1 && q <= saye::QMAX, "XDG cut rules: 1 to 32 Gauss points per direction");
    BH_CHECK(poly_rules_on_device(L), "XDG polynomial cut rules on the device: degree 0 to 4");
    const int n = (int)items.size();
    out.assign(n, QRule{});
    if (n == 0) return;
    std::vector<double> gtw(2 * q);
    gauss01(q, gtw.data(), gtw.data() + q);
    const size_t nc = L.c.size();
    DevBuf ditems(n * sizeof(SphereRuleItem)), dg(2 * q * sizeof(double)), dc(nc * sizeof(double)),
        dcnt(n * sizeof(int64_t)), dfb(n * sizeof(int)), dbad(sizeof(int));
    ctx->upload(ditems.p, items.data(), n * sizeof(SphereRuleItem));
    ctx->upload(dg.p, gtw.data(), 2 * q * sizeof(double));
    ctx->upload(dc.p, L.c.data(), nc * sizeof(double));
    const int zero = 0;
    ctx->upload(dbad.p, &zero, sizeof(int));
    const PolyArg P{dc.as<double>(), L.D, L.n};
    const dim3 grid((unsigned)((n + POLY_TPB - 1) / POLY_TPB));
    hipLaunchKernelGGL(poly_count_k, grid, dim3(POLY_TPB), 0, ctx->stream, n, ditems.as<SphereRuleItem>(), P, dg.as<double>(), q,
                       dcnt.as<int64_t>(), dfb.as<int>(), dbad.as<int>());
    BH_HIP(hipGetLastError());
    std::vector<int64_t> cnt(n), off(n + 1, 0);
    std::vector<int> fb(n);
    int bad = 0;
    ctx->download(cnt.data(), dcnt.p, n * sizeof(int64_t));
    ctx->download(fb.data(), dfb.p, n * sizeof(int));
    ctx->download(&bad, dbad.p, sizeof(int));
    BH_CHECK(!bad, "XDG polynomial cut rules: breakpoint or subdivision capacity exceeded");
    long nfb = 0;
    for (int i = 0; i < n; ++i) {
        off[i + 1] = off[i] + cnt[i];
        nfb += fb[i];
    }
    poly_fallbacks_add(nfb);
    const int64_t tot = off[n];
    if (tot == 0) return;
    DevBuf doff(n * sizeof(int64_t)), dX(3 * tot * sizeof(double)), dW(tot * sizeof(double)), dN(3 * tot * sizeof(double));
    ctx->upload(doff.p, off.data(), n * sizeof(int64_t));
    hipLaunchKernelGGL(poly_fill_k, grid, dim3(POLY_TPB), 0, ctx->stream, n, ditems.as<SphereRuleItem>(), P, dg.as<double>(), q,
                       doff.as<int64_t>(), dX.as<double>(), dW.as<double>(), dN.as<double>());
    BH_HIP(hipGetLastError());
    unpack_rules(ctx, items, off, dX, dW, dN, out);
}

void poly_classify_device(bh_ctx* ctx, const PolyLS& L, const std::vector<SphereRuleItem>& items, std::vector<int>& out) {
    BH_CHECK(poly_rules_on_device(L), "XDG polynomial cut rules on the device: degree 0 to 4");
    const int n = (int)items.size();
    out.assign(n, 0);
    if (n == 0) return;
    const size_t nc = L.c.size();
    DevBuf ditems(n * sizeof(SphereRuleItem)), dc(nc * sizeof(double)), dout(n * sizeof(int));
    ctx->upload(ditems.p, items.data(), n * sizeof(SphereRuleItem));
    ctx->upload(dc.p, L.c.data(), nc * sizeof(double));
    const PolyArg P{dc.as<double>(), L.D, L.n};
    hipLaunchKernelGGL(poly_classify_k, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->stream, n,
                       ditems.as<SphereRuleItem>(), P, dout.as<int>());
    BH_HIP(hipGetLastError());
    ctx->download(out.data(), dout.p, n * sizeof(int));
}

void sphere_rules_device(bh_ctx* ctx, int D, int q, const Sphere& S, const std::vector<SphereRuleItem>& items,
                         std::vector<QRule>& out) {
    BH_CHECK(q >= 1 && q <= sph::QMAX, "XDG cut rules: 1 to 32 Gauss points per direction");
    const int n = (int)items.size();
    out.assign(n, QRule{});
    if (n == 0) return;
    std::vector<double> gxw(2 * q);
    gauss_rule(q, gxw.data(), gxw.data() + q);
    DevBuf ditems(n * sizeof(SphereRuleItem)), dg(2 * q * sizeof(double)), dcnt(n * sizeof(int64_t)), dbad(sizeof(int));
    ctx->upload(ditems.p, items.data(), n * sizeof(SphereRuleItem));
    ctx->upload(dg.p, gxw.data(), 2 * q * sizeof(double));
    const int zero = 0;
    ctx->upload(dbad.p, &zero, sizeof(int));
    const sph::Ball B{{S.c[0], S.c[1], S.c[2]}, S.r};
    const dim3 grid((unsigned)((n + 63) / 64));
    hipLaunchKernelGGL(rule_count_k, grid, dim3(64), 0, ctx->stream, n, ditems.as<SphereRuleItem>(), D, B, dg.as<double>(), q,
                       dcnt.as<int64_t>(), dbad.as<int>());
    BH_HIP(hipGetLastError());
    std::vector<int64_t> cnt(n), off(n + 1, 0);
    int bad = 0;
    ctx->download(cnt.data(), dcnt.p, n * sizeof(int64_t));
    ctx->download(&bad, dbad.p, sizeof(int));
    BH_CHECK(!bad, "XDG cut rules: breakpoint or subdivision capacity exceeded");
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + cnt[i];
    const int64_t tot = off[n];
    if (tot == 0) return;
    DevBuf doff(n * sizeof(int64_t)), dX(3 * tot * sizeof(double)), dW(tot * sizeof(double)), dN(3 * tot * sizeof(double));
    ctx->upload(doff.p, off.data(), n * sizeof(int64_t));
    hipLaunchKernelGGL(rule_fill_k, grid, dim3(64), 0, ctx->stream, n, ditems.as<SphereRuleItem>(), D, B, dg.as<double>(), q,
                       doff.as<int64_t>(), dX.as<double>(), dW.as<double>(), dN.as<double>());
    BH_HIP(hipGetLastError());
    unpack_rules(ctx, items, off, dX, dW, dN, out);
}

}  // namespace bh

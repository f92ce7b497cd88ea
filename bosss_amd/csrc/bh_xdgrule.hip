// Cut-cell, interface and cut-face rules of a sphere level set on the device: one
// thread per rule, the construction of bh_sphrule.hpp (the host set-up runs the same
// code).  Two passes: the node count of every rule, then, at the offsets of their
// prefix sum, the nodes themselves, so every rule lands in its own slice in the
// order the items are listed and the result does not depend on the launch.
#include <vector>

#include "bh_internal.hpp"
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

}  // namespace

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

}  // namespace bh

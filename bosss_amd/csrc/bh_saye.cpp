// Cut-cell quadrature for a polynomial level set on axis-aligned boxes (Saye's
// dimension reduction, X/Quadrature/SayeGaussRule_Algorithm.cs): the host entry
// points over the construction in bh_sayerule.hpp, which the device runs as well.
#include <atomic>
#include <cmath>

#include "bh_internal.hpp"
#include "bh_sayerule.hpp"
#include "bh_xdg.hpp"

namespace bh {

namespace {

std::atomic<long> g_fallbacks{0};

struct RuleEmit {
    QRule& R;
    void operator()(const double* x, double w, const double* nrm) {
        R.x.insert(R.x.end(), x, x + 3);
        R.w.push_back(w);
        if (nrm) R.nrm.insert(R.nrm.end(), nrm, nrm + 3);
    }
};

void run_rule(const PolyLS& L, const double* lo, const double* hi, int species, bool iface, int q, QRule& out) {
    BH_CHECK(L.n >= 0 && L.n <= saye::DEGMAX, "polynomial level set: degree 0 to 8");
    BH_CHECK(q >= 1 && q <= saye::QMAX, "XDG cut rules: 1 to 32 Gauss points per direction");
    double gt[saye::QMAX], gw[saye::QMAX];
    gauss01(q, gt, gw);
    saye::Ctx ctx{gt, gw, q, 0, false};
    out.x.clear();
    out.w.clear();
    out.nrm.clear();
    RuleEmit e{out};
    const bool ok = saye::rule<saye::DEGMAX>(L.c.data(), L.D, L.n, lo, hi, species, iface, ctx, e);
    g_fallbacks += ctx.fallbacks;
    BH_CHECK(ok, "XDG polynomial cut rules: breakpoint or subdivision capacity exceeded");
}

}  // namespace

void gauss01(int q, double* t, double* w) {
    gauss_rule(q, t, w);
    for (int i = 0; i < q; ++i) {
        t[i] = 0.5 * (t[i] + 1.0);
        w[i] *= 0.5;
    }
}

double PolyLS::eval(const double* x) const {
    const int m = n + 1;
    double s = 0;
    const int N = (int)c.size();
    for (int i = 0; i < N; ++i) {
        if (c[i] == 0) continue;
        double t = c[i];
        int r = i;
        for (int d = 0; d < D; ++d) {
            t *= std::pow(x[d], r % m);
            r /= m;
        }
        s += t;
    }
    return s;
}

long poly_fallbacks() { return g_fallbacks.load(); }

void poly_fallbacks_add(long n) { g_fallbacks += n; }

int poly_classify(const PolyLS& L, const double* lo, const double* hi) {
    BH_CHECK(L.n >= 0 && L.n <= saye::DEGMAX, "polynomial level set: degree 0 to 8");
    return saye::classify<saye::DEGMAX>(L.c.data(), L.D, L.n, lo, hi);
}

void poly_volume_rule(const PolyLS& L, const double* lo, const double* hi, int species, int q, QRule& out) {
    run_rule(L, lo, hi, species, false, q, out);
}

void poly_face_rule(const PolyLS& L, const double* lo, const double* hi, int dir, int species, int q, QRule& out) {
    double fl[3], fh[3];
    for (int d = 0; d < 3; ++d) { fl[d] = lo[d]; fh[d] = hi[d]; }
    fh[dir] = fl[dir];
    run_rule(L, fl, fh, species, false, q, out);
}

void poly_interface_rule(const PolyLS& L, const double* lo, const double* hi, int q, QRule& out) {
    run_rule(L, lo, hi, 0, true, q, out);
}

}  // namespace bh

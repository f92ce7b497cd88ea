// Cut-cell quadrature for a polynomial level set on axis-aligned boxes, by Saye's
// dimension reduction (R. Saye, "High-order quadrature methods for implicitly
// defined surfaces and volumes in hyperrectangles", SIAM J. Sci. Comput. 37 (2015)
// A993-A1019), the algorithm of the reference's default cut-cell rules
// (X/Quadrature/SayeGaussRule_Algorithm.cs, SayeGaussRule_Cube.cs):
//  * on a box, functions psi_i with sign conditions s_i (or none, for functions
//    that only split the domain); a function without a root on the box is dropped
//    (or empties the set when its sign is wrong), with the range bounded by the
//    Bernstein coefficients;
//  * a height direction k in which every psi_i is monotone (the Bernstein
//    coefficients of d psi_i / dx_k exclude 0) reduces the problem to the box
//    without k, with the functions psi_i(x', lo_k) and psi_i(x', hi_k) splitting it
//    (where roots enter or leave the k-interval); above every node x' of that rule
//    the k-interval is split at the roots of psi_i(x', .) (one each, found by a
//    Newton iteration safeguarded by bisection) and the pieces where the sign
//    conditions hold get Gauss rules;
//  * without a height direction the box is halved along its longest side (de
//    Casteljau subdivision of the Bernstein form);
//  * the interface rule takes the root of phi above every node of the reduced
//    rule, weight |grad phi| / |d phi / dx_k|, normal grad phi / |grad phi|.
// Polynomials are tensor Bernstein forms of degree n per direction on the box.
#include <algorithm>
#include <atomic>
#include <cmath>

#include "bh_xdg.hpp"

namespace bh {

namespace {

std::atomic<long> g_fallbacks{0};
constexpr int MAXDEPTH = 10;

double binom_d(int n, int k) {
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

struct Gauss01 {
    std::vector<double> t, w;
    explicit Gauss01(int q) : t(q), w(q) {
        gauss_rule(q, t.data(), w.data());
        for (int i = 0; i < q; ++i) { t[i] = 0.5 * (t[i] + 1.0); w[i] *= 0.5; }
    }
};

// Bernstein polynomial of degree n per direction in dim <= 3 local coordinates
struct BPoly {
    int dim = 0, n = 0;
    std::vector<double> b;   // b[i0 + (n+1) i1 + (n+1)^2 i2]
    int stride(int k) const {
        int s = 1;
        for (int j = 0; j < k; ++j) s *= n + 1;
        return s;
    }
    int size() const { return stride(dim); }
};

// value (and derivative) of a univariate Bernstein polynomial at t
double casteljau(const double* c, int n, double t, double* deriv = nullptr) {
    double beta[16];
    for (int i = 0; i <= n; ++i) beta[i] = c[i];
    double d = 0;
    for (int r = 1; r <= n; ++r) {
        if (r == n) d = n * (beta[1] - beta[0]);
        for (int i = 0; i <= n - r; ++i) beta[i] = (1 - t) * beta[i] + t * beta[i + 1];
    }
    if (deriv) *deriv = d;
    return beta[0];
}

// remove direction d of a tensor array (dims, degree n) by evaluating it at t
std::vector<double> collapse(const std::vector<double>& a, int dim, int n, int d, double t) {
    const int m = n + 1;
    int sd = 1, so = 1;
    for (int j = 0; j < d; ++j) sd *= m;
    for (int j = d + 1; j < dim; ++j) so *= m;
    std::vector<double> out((size_t)sd * so);
    double c[16];
    for (int o = 0; o < so; ++o)
        for (int i = 0; i < sd; ++i) {
            for (int j = 0; j <= n; ++j) c[j] = a[(size_t)(o * m + j) * sd + i];
            out[(size_t)o * sd + i] = casteljau(c, n, t);
        }
    return out;
}

double eval(const BPoly& P, const double* t) {
    std::vector<double> a = P.b;
    for (int d = P.dim - 1; d >= 0; --d) a = collapse(a, d + 1, P.n, d, t[d]);
    return a[0];
}

// univariate coefficients along k at the other coordinates t (t[k] ignored)
void fiber(const BPoly& P, int k, const double* t, double* c) {
    std::vector<double> a = P.b;
    int dim = P.dim, kk = k;
    for (int d = P.dim - 1; d >= 0; --d) {
        if (d == k) continue;
        a = collapse(a, dim, P.n, d, t[d]);
        --dim;
        if (d < kk) --kk;
    }
    for (int j = 0; j <= P.n; ++j) c[j] = a[j];
}

void grad(const BPoly& P, const double* t, double* g) {
    double c[16];
    for (int k = 0; k < P.dim; ++k) {
        fiber(P, k, t, c);
        casteljau(c, P.n, t[k], &g[k]);
    }
}

void range(const BPoly& P, double& lo, double& hi) {
    lo = hi = P.b[0];
    for (double v : P.b) { lo = std::min(lo, v); hi = std::max(hi, v); }
}

// bounds of d/dt_k
void deriv_range(const BPoly& P, int k, double& lo, double& hi) {
    const int m = P.n + 1, sk = P.stride(k), N = P.size();
    lo = 1e300;
    hi = -1e300;
    if (P.n == 0) { lo = hi = 0; return; }
    for (int i = 0; i < N; ++i) {
        if ((i / sk) % m == P.n) continue;
        const double v = P.n * (P.b[i + sk] - P.b[i]);
        lo = std::min(lo, v);
        hi = std::max(hi, v);
    }
}

// phi is monotone in direction k on the box, with a margin: a derivative that is
// zero on a face (a surface perpendicular to the face) leaves Bernstein coefficients
// at rounding level whose sign means nothing, and taking such a direction would put
// 1 / (d phi / dx_k) -> infinity into the surface weights
bool monotone(const BPoly& P, int k) {
    double lo, hi;
    deriv_range(P, k, lo, hi);
    const double tol = 1e-9 * std::max(std::fabs(lo), std::fabs(hi));
    return lo > tol || hi < -tol;
}

// the face t_k = side of the box
BPoly restrict_face(const BPoly& P, int k, int side) {
    BPoly R;
    R.dim = P.dim - 1;
    R.n = P.n;
    R.b.resize(R.size());
    const int m = P.n + 1, sk = P.stride(k), N = P.size();
    int o = 0;
    for (int i = 0; i < N; ++i)
        if ((i / sk) % m == (side ? P.n : 0)) R.b[o++] = P.b[i];
    return R;
}

// the half t_k in [0, 1/2] (half = 0) or [1/2, 1] (half = 1), reparametrised to [0, 1]
BPoly split(const BPoly& P, int k, int half) {
    BPoly R = P;
    const int m = P.n + 1, sk = P.stride(k), N = P.size();
    double beta[16], out[16];
    for (int i = 0; i < N; ++i) {
        if ((i / sk) % m != 0) continue;
        for (int j = 0; j <= P.n; ++j) beta[j] = P.b[i + j * sk];
        for (int r = 0; r <= P.n; ++r) {
            if (half == 0) out[r] = beta[0];
            else out[P.n - r] = beta[P.n - r];
            for (int j = 0; j < P.n - r; ++j) beta[j] = 0.5 * (beta[j] + beta[j + 1]);
        }
        for (int j = 0; j <= P.n; ++j) R.b[i + j * sk] = out[j];
    }
    return R;
}

// phi on the box [lo, hi] (lo[d] == hi[d]: that coordinate fixed), in the box's active
// directions; act[] receives the physical direction of each local one
BPoly from_poly(const PolyLS& L, const double* lo, const double* hi, int* act) {
    const int D = L.D, n = L.n, m = n + 1;
    std::vector<double> a = L.c;
    int S[3] = {1, m, m * m};
    const int N = (int)a.size();
    // x_d = lo_d + h_d t_d: power basis in t_d
    for (int d = 0; d < D; ++d) {
        const double h = hi[d] - lo[d], x0 = lo[d];
        std::vector<double> nb(a.size(), 0.0);
        for (int i = 0; i < N; ++i) {
            if ((i / S[d]) % m != 0) continue;
            for (int j = 0; j <= n; ++j) {
                double s = 0;
                for (int ii = j; ii <= n; ++ii) s += a[i + ii * S[d]] * binom_d(ii, j) * std::pow(x0, ii - j);
                nb[i + j * S[d]] = s * std::pow(h, j);
            }
        }
        a.swap(nb);
    }
    // power -> Bernstein in the active directions; fixed ones keep only degree 0
    BPoly P;
    P.n = n;
    P.dim = 0;
    for (int d = 0; d < D; ++d)
        if (hi[d] > lo[d]) act[P.dim++] = d;
    for (int e = 0; e < P.dim; ++e) {
        const int d = act[e];
        std::vector<double> nb(a.size(), 0.0);
        for (int i = 0; i < N; ++i) {
            if ((i / S[d]) % m != 0) continue;
            for (int j = 0; j <= n; ++j) {
                double s = 0;
                for (int ii = 0; ii <= j; ++ii) s += binom_d(j, ii) / binom_d(n, ii) * a[i + ii * S[d]];
                nb[i + j * S[d]] = s;
            }
        }
        a.swap(nb);
    }
    P.b.resize(P.size());
    for (int i = 0; i < N; ++i) {
        bool keep = true;
        int idx = 0, mul = 1;
        for (int d = 0; d < D; ++d) {
            const int id = (i / S[d]) % m;
            if (hi[d] > lo[d]) { idx += id * mul; mul *= m; }
            else if (id != 0) keep = false;
        }
        if (keep) P.b[idx] = a[i];
    }
    return P;
}

// root of a monotone univariate Bernstein polynomial with c0 c_n <= 0 in [0, 1]:
// Newton's iteration, safeguarded by the bisection bracket
double mono_root(const double* c, int n) {
    if (c[0] == 0) return 0.0;
    if (c[n] == 0) return 1.0;
    double lo = 0, hi = 1;
    const bool neg0 = c[0] < 0;
    double t = c[0] / (c[0] - c[n]);
    for (int it = 0; it < 100; ++it) {
        double fp;
        const double f = casteljau(c, n, t, &fp);
        if (f == 0) return t;
        if ((f < 0) == neg0) lo = t;
        else hi = t;
        double tn = (fp != 0) ? t - f / fp : 0.5 * (lo + hi);
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        if (std::fabs(tn - t) <= 1e-16 || hi - lo <= 1e-16) return tn;
        t = tn;
    }
    return t;
}

// all roots in [u, v] of a univariate Bernstein polynomial given on [u, v]
void roots_rec(const double* c, int n, double u, double v, int depth, std::vector<double>& out) {
    bool pos = true, neg = true;
    for (int i = 0; i <= n; ++i) {
        if (c[i] <= 0) pos = false;
        if (c[i] >= 0) neg = false;
    }
    if (pos || neg) return;
    bool inc = true, dec = true;
    for (int i = 0; i < n; ++i) {
        if (c[i + 1] < c[i]) inc = false;
        if (c[i + 1] > c[i]) dec = false;
    }
    if (inc || dec || depth >= 60) {
        if (c[0] * c[n] <= 0) out.push_back(u + (v - u) * mono_root(c, n));
        return;
    }
    double beta[16], L[16], R[16];
    for (int j = 0; j <= n; ++j) beta[j] = c[j];
    for (int r = 0; r <= n; ++r) {
        L[r] = beta[0];
        R[n - r] = beta[n - r];
        for (int j = 0; j < n - r; ++j) beta[j] = 0.5 * (beta[j] + beta[j + 1]);
    }
    const double mid = 0.5 * (u + v);
    roots_rec(L, n, u, mid, depth + 1, out);
    roots_rec(R, n, mid, v, depth + 1, out);
}

struct Fn {
    BPoly p;
    int s;   // required sign (-1 / +1), 0: splits the domain only
};

// split [0, 1] at the breakpoints bp, keep the pieces whose midpoint satisfies ok(t)
template <class OK, class EMIT>
void pieces01(std::vector<double>& bp, OK ok, EMIT emit) {
    bp.push_back(0.0);
    bp.push_back(1.0);
    std::sort(bp.begin(), bp.end());
    for (size_t i = 0; i + 1 < bp.size(); ++i) {
        const double u = std::max(0.0, bp[i]), v = std::min(1.0, bp[i + 1]);
        if (!(v - u > 1e-14)) continue;
        if (ok(0.5 * (u + v))) emit(u, v);
    }
}

// Nodes (dim coordinates each, in the top box's t-space) and weights (t-measure) of
// {t in [a, b]: s_i psi_i(t) > 0 for s_i != 0}, the psi_i given on [a, b] as local
// Bernstein forms.
void vol_rec(int dim, const std::vector<Fn>& fs, const double* a, const double* b, const Gauss01& G, int depth,
             std::vector<double>& X, std::vector<double>& W) {
    std::vector<Fn> keep;
    for (const Fn& f : fs) {
        double lo, hi;
        range(f.p, lo, hi);
        if (lo > 0 || hi < 0) {
            if (f.s != 0 && (lo > 0 ? 1 : -1) != f.s) return;
            continue;
        }
        keep.push_back(f);
    }
    const int q = (int)G.t.size();
    auto tensor = [&](auto accept) {
        int tot = 1;
        for (int d = 0; d < dim; ++d) tot *= q;
        for (int i = 0; i < tot; ++i) {
            double tl[3], x[3], w = 1;
            int r = i;
            for (int d = 0; d < dim; ++d) {
                const int id = r % q;
                r /= q;
                tl[d] = G.t[id];
                x[d] = a[d] + (b[d] - a[d]) * G.t[id];
                w *= (b[d] - a[d]) * G.w[id];
            }
            if (!accept(tl)) continue;
            for (int d = 0; d < dim; ++d) X.push_back(x[d]);
            W.push_back(w);
        }
    };
    if (keep.empty()) {
        tensor([](const double*) { return true; });
        return;
    }
    if (dim == 1) {
        std::vector<double> bp;
        for (const Fn& f : keep) roots_rec(f.p.b.data(), f.p.n, 0.0, 1.0, 0, bp);
        pieces01(
            bp,
            [&](double tm) {
                for (const Fn& f : keep)
                    if (f.s != 0 && (casteljau(f.p.b.data(), f.p.n, tm) > 0 ? 1 : -1) != f.s) return false;
                return true;
            },
            [&](double u, double v) {
                for (int i = 0; i < q; ++i) {
                    X.push_back(a[0] + (b[0] - a[0]) * (u + (v - u) * G.t[i]));
                    W.push_back((b[0] - a[0]) * (v - u) * G.w[i]);
                }
            });
        return;
    }
    // height direction: largest |d psi / dt_k| at the centre first, monotone for every psi
    const Fn& f0 = keep[0];
    double tc[3] = {0.5, 0.5, 0.5}, g0[3];
    grad(f0.p, tc, g0);
    int order[3] = {0, 1, 2};
    std::sort(order, order + dim, [&](int x, int y) { return std::fabs(g0[x]) > std::fabs(g0[y]); });
    int k = -1;
    for (int oi = 0; oi < dim && k < 0; ++oi) {
        bool ok = true;
        for (const Fn& f : keep)
            if (!monotone(f.p, order[oi])) { ok = false; break; }
        if (ok) k = order[oi];
    }
    // at the subdivision limit (a point where every direction degenerates, e.g. a
    // function that touches zero without changing sign): the direction of the largest
    // derivative, with all roots of each fibre instead of the single monotone one
    bool all_roots = false;
    if (k < 0 && depth >= MAXDEPTH) {
        ++g_fallbacks;
        k = order[0];
        all_roots = true;
    }
    if (k < 0) {
        int sd = 0;
        for (int d = 1; d < dim; ++d)
            if (b[d] - a[d] > b[sd] - a[sd]) sd = d;
        for (int half = 0; half < 2; ++half) {
            std::vector<Fn> hs;
            for (const Fn& f : keep) hs.push_back({split(f.p, sd, half), f.s});
            double a2[3], b2[3];
            for (int d = 0; d < dim; ++d) { a2[d] = a[d]; b2[d] = b[d]; }
            const double mid = 0.5 * (a[sd] + b[sd]);
            if (half == 0) b2[sd] = mid;
            else a2[sd] = mid;
            vol_rec(dim, hs, a2, b2, G, depth + 1, X, W);
        }
        return;
    }
    // reduced problem on the box without k
    std::vector<Fn> outer;
    for (const Fn& f : keep) {
        outer.push_back({restrict_face(f.p, k, 0), 0});
        outer.push_back({restrict_face(f.p, k, 1), 0});
    }
    double ao[3], bo[3];
    for (int d = 0, e = 0; d < dim; ++d)
        if (d != k) { ao[e] = a[d]; bo[e] = b[d]; ++e; }
    std::vector<double> Xo, Wo;
    vol_rec(dim - 1, outer, ao, bo, G, depth, Xo, Wo);
    const int no = (int)Wo.size();
    std::vector<std::vector<double>> fib(keep.size(), std::vector<double>(keep[0].p.n + 1));
    for (int i = 0; i < no; ++i) {
        double tl[3], xo[3];
        for (int d = 0, e = 0; d < dim; ++d)
            if (d != k) {
                xo[d] = Xo[(size_t)i * (dim - 1) + e];
                tl[d] = (xo[d] - a[d]) / (b[d] - a[d]);
                ++e;
            }
        tl[k] = 0;
        std::vector<double> bp;
        for (size_t j = 0; j < keep.size(); ++j) {
            fiber(keep[j].p, k, tl, fib[j].data());
            const int n = keep[j].p.n;
            if (all_roots) roots_rec(fib[j].data(), n, 0.0, 1.0, 0, bp);
            else if (fib[j][0] * fib[j][n] <= 0) bp.push_back(mono_root(fib[j].data(), n));
        }
        pieces01(
            bp,
            [&](double tm) {
                for (size_t j = 0; j < keep.size(); ++j)
                    if (keep[j].s != 0 && (casteljau(fib[j].data(), keep[j].p.n, tm) > 0 ? 1 : -1) != keep[j].s) return false;
                return true;
            },
            [&](double u, double v) {
                for (int gq = 0; gq < q; ++gq) {
                    for (int d = 0; d < dim; ++d)
                        X.push_back(d == k ? a[k] + (b[k] - a[k]) * (u + (v - u) * G.t[gq]) : xo[d]);
                    W.push_back(Wo[i] * (b[k] - a[k]) * (v - u) * G.w[gq]);
                }
            });
    }
}

// interface {phi = 0} in the box [a, b] (top-box t-space): nodes, physical weights and
// unit normals grad phi / |grad phi| (physical); h: physical sizes of the top box
void surf_rec(const BPoly& P, const double* a, const double* b, const double* h, const Gauss01& G, int depth,
              std::vector<double>& X, std::vector<double>& W, std::vector<double>& Nn) {
    const int dim = P.dim;
    double lo, hi;
    range(P, lo, hi);
    if (lo > 0 || hi < 0) return;
    double tc[3] = {0.5, 0.5, 0.5}, g0[3];
    grad(P, tc, g0);
    int order[3] = {0, 1, 2};
    // compare physical derivatives
    std::sort(order, order + dim, [&](int x, int y) {
        return std::fabs(g0[x] / (h[x] * (b[x] - a[x]))) > std::fabs(g0[y] / (h[y] * (b[y] - a[y])));
    });
    int k = -1;
    for (int oi = 0; oi < dim && k < 0; ++oi)
        if (monotone(P, order[oi])) k = order[oi];
    bool all_roots = false;
    if (k < 0 && depth >= MAXDEPTH) {   // as in vol_rec: every root of the fibre
        ++g_fallbacks;
        k = order[0];
        all_roots = true;
    }
    if (k < 0) {
        int sd = 0;
        for (int d = 1; d < dim; ++d)
            if ((b[d] - a[d]) * h[d] > (b[sd] - a[sd]) * h[sd]) sd = d;
        for (int half = 0; half < 2; ++half) {
            double a2[3], b2[3];
            for (int d = 0; d < dim; ++d) { a2[d] = a[d]; b2[d] = b[d]; }
            const double mid = 0.5 * (a[sd] + b[sd]);
            if (half == 0) b2[sd] = mid;
            else a2[sd] = mid;
            surf_rec(split(P, sd, half), a2, b2, h, G, depth + 1, X, W, Nn);
        }
        return;
    }
    std::vector<Fn> outer = {{restrict_face(P, k, 0), 0}, {restrict_face(P, k, 1), 0}};
    double ao[3], bo[3];
    for (int d = 0, e = 0; d < dim; ++d)
        if (d != k) { ao[e] = a[d]; bo[e] = b[d]; ++e; }
    std::vector<double> Xo, Wo;
    vol_rec(dim - 1, outer, ao, bo, G, depth, Xo, Wo);
    std::vector<double> c(P.n + 1);
    for (size_t i = 0; i < Wo.size(); ++i) {
        double tl[3], xo[3];
        for (int d = 0, e = 0; d < dim; ++d)
            if (d != k) {
                xo[d] = Xo[i * (dim - 1) + e];
                tl[d] = (xo[d] - a[d]) / (b[d] - a[d]);
                ++e;
            }
        tl[k] = 0;
        fiber(P, k, tl, c.data());
        std::vector<double> rts;
        if (all_roots) roots_rec(c.data(), P.n, 0.0, 1.0, 0, rts);
        else if (c[0] * c[P.n] <= 0) rts.push_back(mono_root(c.data(), P.n));
        for (double rt : rts) {
            tl[k] = rt;
            double gl[3], gp[3], gn = 0, wphys = Wo[i];
            grad(P, tl, gl);
            for (int d = 0; d < dim; ++d) {
                gp[d] = gl[d] / (h[d] * (b[d] - a[d]));
                gn += gp[d] * gp[d];
                if (d != k) wphys *= h[d];
            }
            gn = std::sqrt(gn);
            if (!(gn > 0) || gp[k] == 0) continue;
            for (int d = 0; d < dim; ++d) X.push_back(d == k ? a[k] + (b[k] - a[k]) * tl[k] : xo[d]);
            W.push_back(wphys * gn / std::fabs(gp[k]));
            for (int d = 0; d < dim; ++d) Nn.push_back(gp[d] / gn);
        }
    }
}

void to_rule(const double* lo, const double* hi, const int* act, int dim, const std::vector<double>& X,
             const std::vector<double>& W, bool scale_w, QRule& out) {
    out.x.clear();
    out.w.clear();
    out.nrm.clear();
    const size_t n = W.size();
    double vol = 1;
    for (int e = 0; e < dim; ++e) vol *= hi[act[e]] - lo[act[e]];
    for (size_t i = 0; i < n; ++i) {
        double x[3] = {lo[0], lo[1], lo[2]};
        for (int e = 0; e < dim; ++e) x[act[e]] = lo[act[e]] + (hi[act[e]] - lo[act[e]]) * X[i * dim + e];
        out.x.insert(out.x.end(), x, x + 3);
        out.w.push_back(scale_w ? W[i] * vol : W[i]);
    }
}

}  // namespace

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

int poly_classify(const PolyLS& L, const double* lo, const double* hi) {
    int act[3];
    const BPoly P = from_poly(L, lo, hi, act);
    double a, b;
    range(P, a, b);
    if (b < 0) return -1;
    if (a > 0) return +1;
    return 0;
}

void poly_volume_rule(const PolyLS& L, const double* lo, const double* hi, int species, int q, QRule& out) {
    int act[3];
    const BPoly P = from_poly(L, lo, hi, act);
    const Gauss01 G(q);
    const double a[3] = {0, 0, 0}, b[3] = {1, 1, 1};
    std::vector<double> X, W;
    vol_rec(P.dim, {{P, species}}, a, b, G, 0, X, W);
    to_rule(lo, hi, act, P.dim, X, W, true, out);
}

void poly_face_rule(const PolyLS& L, const double* lo, const double* hi, int dir, int species, int q, QRule& out) {
    double fl[3], fh[3];
    for (int d = 0; d < 3; ++d) { fl[d] = lo[d]; fh[d] = hi[d]; }
    fh[dir] = fl[dir];
    poly_volume_rule(L, fl, fh, species, q, out);
}

void poly_interface_rule(const PolyLS& L, const double* lo, const double* hi, int q, QRule& out) {
    int act[3];
    const BPoly P = from_poly(L, lo, hi, act);
    const Gauss01 G(q);
    const double a[3] = {0, 0, 0}, b[3] = {1, 1, 1};
    double h[3];
    for (int e = 0; e < P.dim; ++e) h[e] = hi[act[e]] - lo[act[e]];
    std::vector<double> X, W, Nn;
    surf_rec(P, a, b, h, G, 0, X, W, Nn);
    to_rule(lo, hi, act, P.dim, X, W, false, out);
    out.nrm.clear();
    for (size_t i = 0; i < W.size(); ++i) {
        double nv[3] = {0, 0, 0};
        for (int e = 0; e < P.dim; ++e) nv[act[e]] = Nn[i * P.dim + e];
        out.nrm.insert(out.nrm.end(), nv, nv + 3);
    }
}

}  // namespace bh

// Cut-cell quadrature for a polynomial level set on axis-aligned boxes, by Saye's
// dimension reduction (R. Saye, "High-order quadrature methods for implicitly
// defined surfaces and volumes in hyperrectangles", SIAM J. Sci. Comput. 37 (2015)
// A993-A1019), the algorithm of the reference's default cut-cell rules
// (X/Quadrature/SayeGaussRule_Algorithm.cs, SayeGaussRule_Cube.cs).  One
// implementation for the host set-up and for the device (one thread per rule):
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
//    Casteljau subdivision of the Bernstein form), at most MAXDEPTH times; beyond,
//    the direction of the largest derivative with every root of each fibre (counted
//    as a fallback);
//  * the interface rule takes the root of phi above every node of the reduced
//    rule, weight |grad phi| / |d phi / dx_k|, normal grad phi / |grad phi|.
// Polynomials are tensor Bernstein forms of degree n <= NM per direction.  The
// recursions use no call stack: the dimension levels are nested templates (3 -> 2
// -> 1), and the halvings of a level are an explicit depth-first stack of boxes
// whose Bernstein forms are recomputed from the level's input by replaying the
// recorded halvings (the same operations as carrying the halves down).  Nodes go
// to an emitter e(x[3], w, nrm or nullptr) in a fixed order, so host and device
// give the same rules bit for bit when both round alike (IEEE add, multiply,
// divide, square root; no contraction into fused multiply-adds).
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define BH_HD __host__ __device__
#else
#define BH_HD
#endif
// called at many sites: one frame each instead of a copy of its arrays per site
#define SAYE_OUTLINE __attribute__((noinline))

namespace bh {
namespace saye {

constexpr int MAXDEPTH = 10;    // halvings per box
constexpr int ROOTDEPTH = 60;   // bisections of a univariate root search
constexpr int DEGMAX = 8;       // polynomial degree per direction
constexpr int QMAX = 32;        // 1D Gauss points

BH_HD constexpr int ipow(int b, int e) { return e == 0 ? 1 : b * ipow(b, e - 1); }
// functions per level: one at the top, two (the faces of each kept one) per level down
BH_HD constexpr int nfun(int dim) { return dim >= 3 ? 1 : (dim == 2 ? 2 : 4); }

// rule context: Gauss points on [0, 1], failure flag, fallback count
struct Ctx {
    const double* gt;
    const double* gw;
    int q;
    int fallbacks;
    bool bad;
};

// Bernstein polynomial of degree n per direction in DIM local coordinates:
// b[i0 + (n+1) i1 + (n+1)^2 i2]
template <int NM, int DIM>
struct BP {
    int n;
    double b[ipow(NM + 1, DIM)];
    BH_HD int stride(int k) const {
        int s = 1;
        for (int j = 0; j < k; ++j) s *= n + 1;
        return s;
    }
    BH_HD int size() const { return stride(DIM); }
};

template <int NM, int DIM>
struct Fns {
    BP<NM, DIM> p[nfun(DIM)];
    int s[nfun(DIM)];   // required sign (-1 / +1), 0: splits the domain only
    int nf;
};

BH_HD inline double binom_d(int n, int k) {
    double r = 1;
    for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
    return r;
}

BH_HD inline double ipowd(double x, int e) {
    double r = 1;
    for (int i = 0; i < e; ++i) r *= x;
    return r;
}

// value (and derivative) of a univariate Bernstein polynomial at t
BH_HD inline double casteljau(const double* c, int n, double t, double* deriv = nullptr) {
    double beta[DEGMAX + 1];
    for (int i = 0; i <= n; ++i) beta[i] = c[i];
    double d = 0;
    for (int r = 1; r <= n; ++r) {
        if (r == n) d = n * (beta[1] - beta[0]);
        for (int i = 0; i <= n - r; ++i) beta[i] = (1 - t) * beta[i] + t * beta[i + 1];
    }
    if (deriv) *deriv = d;
    return beta[0];
}

// univariate coefficients along k at the other coordinates t (t[k] ignored): the
// other directions are evaluated away, last direction first, in place
template <int NM, int DIM>
BH_HD SAYE_OUTLINE void fiber(const BP<NM, DIM>& P, int k, const double* t, double* c) {
    double a[ipow(NM + 1, DIM)];
    const int n = P.n, m = n + 1, N = P.size();
    for (int i = 0; i < N; ++i) a[i] = P.b[i];
    int dim = DIM;
    for (int d = DIM - 1; d >= 0; --d) {
        if (d == k) continue;
        // remove direction d of the current (dim-direction) array: only directions
        // above d have been removed, so d is still its local index
        int sd = 1, so = 1;
        for (int j = 0; j < d; ++j) sd *= m;
        for (int j = d + 1; j < dim; ++j) so *= m;
        double col[DEGMAX + 1];
        for (int o = 0; o < so; ++o)
            for (int i = 0; i < sd; ++i) {
                for (int j = 0; j <= n; ++j) col[j] = a[(o * m + j) * sd + i];
                a[o * sd + i] = casteljau(col, n, t[d]);
            }
        --dim;
    }
    for (int j = 0; j <= n; ++j) c[j] = a[j];
}

template <int NM, int DIM>
BH_HD inline void grad(const BP<NM, DIM>& P, const double* t, double* g) {
    double c[DEGMAX + 1];
    for (int k = 0; k < DIM; ++k) {
        fiber(P, k, t, c);
        casteljau(c, P.n, t[k], &g[k]);
    }
}

template <int NM, int DIM>
BH_HD inline void range(const BP<NM, DIM>& P, double& lo, double& hi) {
    lo = hi = P.b[0];
    const int N = P.size();
    for (int i = 1; i < N; ++i) {
        lo = fmin(lo, P.b[i]);
        hi = fmax(hi, P.b[i]);
    }
}

// phi is monotone in direction k on the box, with a margin: a derivative that is
// zero on a face (a surface perpendicular to the face) leaves Bernstein coefficients
// at rounding level whose sign means nothing, and taking such a direction would put
// 1 / (d phi / dx_k) -> infinity into the surface weights
template <int NM, int DIM>
BH_HD inline bool monotone(const BP<NM, DIM>& P, int k) {
    const int m = P.n + 1, sk = P.stride(k), N = P.size();
    double lo = 1e300, hi = -1e300;
    if (P.n == 0) lo = hi = 0;
    else
        for (int i = 0; i < N; ++i) {
            if ((i / sk) % m == P.n) continue;
            const double v = P.n * (P.b[i + sk] - P.b[i]);
            lo = fmin(lo, v);
            hi = fmax(hi, v);
        }
    const double tol = 1e-9 * fmax(fabs(lo), fabs(hi));
    return lo > tol || hi < -tol;
}

// the face t_k = side of the box
template <int NM, int DIM>
BH_HD inline void restrict_face(const BP<NM, DIM>& P, int k, int side, BP<NM, DIM - 1>& R) {
    R.n = P.n;
    const int m = P.n + 1, sk = P.stride(k), N = P.size();
    int o = 0;
    for (int i = 0; i < N; ++i)
        if ((i / sk) % m == (side ? P.n : 0)) R.b[o++] = P.b[i];
}

// in place: the half t_k in [0, 1/2] (half = 0) or [1/2, 1] (half = 1), reparametrised to [0, 1]
template <int NM, int DIM>
BH_HD inline void split(BP<NM, DIM>& P, int k, int half) {
    const int m = P.n + 1, sk = P.stride(k), N = P.size(), n = P.n;
    double beta[DEGMAX + 1], out[DEGMAX + 1];
    for (int i = 0; i < N; ++i) {
        if ((i / sk) % m != 0) continue;
        for (int j = 0; j <= n; ++j) beta[j] = P.b[i + j * sk];
        for (int r = 0; r <= n; ++r) {
            if (half == 0) out[r] = beta[0];
            else out[n - r] = beta[n - r];
            for (int j = 0; j < n - r; ++j) beta[j] = 0.5 * (beta[j] + beta[j + 1]);
        }
        for (int j = 0; j <= n; ++j) P.b[i + j * sk] = out[j];
    }
}

// root of a monotone univariate Bernstein polynomial with c0 c_n <= 0 in [0, 1]:
// Newton's iteration, safeguarded by the bisection bracket
BH_HD inline double mono_root(const double* c, int n) {
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
        if (fabs(tn - t) <= 1e-16 || hi - lo <= 1e-16) return tn;
        t = tn;
    }
    return t;
}

// breakpoints of [0, 1]
struct Pts {
    static constexpr int CAP = 4 * DEGMAX + 4;
    double v[CAP];
    int n = 0;
    BH_HD void push(double x, Ctx& ctx) {
        if (n < CAP) v[n++] = x;
        else ctx.bad = true;
    }
};

// all roots in [0, 1] of a univariate Bernstein polynomial, ascending: bisection of
// the Bernstein form, depth first (left half first); a stack entry keeps its interval
// and the halvings that lead to it, replayed from c
BH_HD SAYE_OUTLINE inline void roots_all(const double* c, int n, Pts& out, Ctx& ctx) {
    struct E {
        double u, v;
        uint64_t path;
        int depth;
    };
    E st[ROOTDEPTH + 2];
    int sp = 0;
    st[sp++] = E{0.0, 1.0, 0, 0};
    while (sp > 0) {
        const E e = st[--sp];
        double cc[DEGMAX + 1];
        for (int j = 0; j <= n; ++j) cc[j] = c[j];
        for (int s = 0; s < e.depth; ++s) {
            const int half = (int)((e.path >> s) & 1);
            double beta[DEGMAX + 1], o[DEGMAX + 1];
            for (int j = 0; j <= n; ++j) beta[j] = cc[j];
            for (int r = 0; r <= n; ++r) {
                if (half == 0) o[r] = beta[0];
                else o[n - r] = beta[n - r];
                for (int j = 0; j < n - r; ++j) beta[j] = 0.5 * (beta[j] + beta[j + 1]);
            }
            for (int j = 0; j <= n; ++j) cc[j] = o[j];
        }
        bool pos = true, neg = true;
        for (int i = 0; i <= n; ++i) {
            if (cc[i] <= 0) pos = false;
            if (cc[i] >= 0) neg = false;
        }
        if (pos || neg) continue;
        bool inc = true, dec = true;
        for (int i = 0; i < n; ++i) {
            if (cc[i + 1] < cc[i]) inc = false;
            if (cc[i + 1] > cc[i]) dec = false;
        }
        if (inc || dec || e.depth >= ROOTDEPTH) {
            if (cc[0] * cc[n] <= 0) out.push(e.u + (e.v - e.u) * mono_root(cc, n), ctx);
            continue;
        }
        if (sp + 2 > ROOTDEPTH + 2) { ctx.bad = true; return; }
        const double mid = 0.5 * (e.u + e.v);
        st[sp++] = E{mid, e.v, e.path | ((uint64_t)1 << e.depth), e.depth + 1};
        st[sp++] = E{e.u, mid, e.path, e.depth + 1};
    }
}

// split [0, 1] at the breakpoints, keep the pieces whose midpoint satisfies ok(t)
template <class OK, class EMIT>
BH_HD inline void pieces01(Pts& bp, Ctx& ctx, OK ok, EMIT emit) {
    bp.push(0.0, ctx);
    bp.push(1.0, ctx);
    for (int i = 1; i < bp.n; ++i) {   // insertion sort, ascending
        const double x = bp.v[i];
        int j = i;
        while (j > 0 && bp.v[j - 1] > x) { bp.v[j] = bp.v[j - 1]; --j; }
        bp.v[j] = x;
    }
    for (int i = 0; i + 1 < bp.n; ++i) {
        const double u = fmax(0.0, bp.v[i]), v = fmin(1.0, bp.v[i + 1]);
        if (!(v - u > 1e-14)) continue;
        if (ok(0.5 * (u + v))) emit(u, v);
    }
}

// directions 0..dim-1 ordered by decreasing key (ties: lower index first)
BH_HD inline void order_desc(int dim, const double* key, int* order) {
    for (int i = 0; i < dim; ++i) order[i] = i;
    for (int i = 1; i < dim; ++i) {
        const int x = order[i];
        int j = i;
        while (j > 0 && key[x] > key[order[j - 1]]) { order[j] = order[j - 1]; --j; }
        order[j] = x;
    }
}

// explicit depth-first stack of the halvings of one level: box, depth and the halvings
// from the level's input box (2 bits direction, 1 bit half each)
template <int DIM>
struct Box {
    double a[DIM], b[DIM];
    uint32_t path;
    int plen, depth;
};

// Nodes (DIM coordinates each, in the top box's t-space) and weights (t-measure) of
// {t in [a, b]: s_i psi_i(t) > 0 for s_i != 0}, the psi_i given on [a, b] as local
// Bernstein forms; emit(x, w)
template <int NM, int DIM, class EMIT>
BH_HD SAYE_OUTLINE void vol(const Fns<NM, DIM>& top, const double* a0, const double* b0, int depth0, Ctx& ctx, EMIT&& emit) {
    const int q = ctx.q;
    Box<DIM> st[MAXDEPTH + 2];
    int sp = 0;
    {
        Box<DIM> r;
        for (int d = 0; d < DIM; ++d) { r.a[d] = a0[d]; r.b[d] = b0[d]; }
        r.path = 0;
        r.plen = 0;
        r.depth = depth0;
        st[sp++] = r;
    }
    Fns<NM, DIM> cur;
    while (sp > 0 && !ctx.bad) {
        const Box<DIM> bx = st[--sp];
        const double* a = bx.a;
        const double* b = bx.b;
        // the functions on this box, those with a root kept (in order)
        cur.nf = 0;
        bool empty = false;
        for (int f = 0; f < top.nf && !empty; ++f) {
            BP<NM, DIM>& P = cur.p[cur.nf];
            P.n = top.p[f].n;
            const int N = P.size();
            for (int i = 0; i < N; ++i) P.b[i] = top.p[f].b[i];
            for (int s = 0; s < bx.plen; ++s) split(P, (int)((bx.path >> (3 * s)) & 3), (int)((bx.path >> (3 * s + 2)) & 1));
            double lo, hi;
            range(P, lo, hi);
            if (lo > 0 || hi < 0) {
                if (top.s[f] != 0 && (lo > 0 ? 1 : -1) != top.s[f]) empty = true;
                continue;
            }
            cur.s[cur.nf++] = top.s[f];
        }
        if (empty) continue;
        if (cur.nf == 0) {   // tensor Gauss rule of the box
            int tot = 1;
            for (int d = 0; d < DIM; ++d) tot *= q;
            for (int i = 0; i < tot; ++i) {
                double x[DIM], w = 1;
                int r = i;
                for (int d = 0; d < DIM; ++d) {
                    const int id = r % q;
                    r /= q;
                    x[d] = a[d] + (b[d] - a[d]) * ctx.gt[id];
                    w *= (b[d] - a[d]) * ctx.gw[id];
                }
                emit(x, w);
            }
            continue;
        }
        if constexpr (DIM == 1) {
            Pts bp;
            for (int f = 0; f < cur.nf; ++f) roots_all(cur.p[f].b, cur.p[f].n, bp, ctx);
            pieces01(
                bp, ctx,
                [&](double tm) {
                    for (int f = 0; f < cur.nf; ++f)
                        if (cur.s[f] != 0 && (casteljau(cur.p[f].b, cur.p[f].n, tm) > 0 ? 1 : -1) != cur.s[f]) return false;
                    return true;
                },
                [&](double u, double v) {
                    for (int i = 0; i < q; ++i) {
                        double x[1] = {a[0] + (b[0] - a[0]) * (u + (v - u) * ctx.gt[i])};
                        emit(x, (b[0] - a[0]) * (v - u) * ctx.gw[i]);
                    }
                });
        } else {
            // height direction: largest |d psi / dt_k| at the centre first, monotone for every psi
            double tc[DIM], g0[DIM], key[DIM];
            for (int d = 0; d < DIM; ++d) tc[d] = 0.5;
            grad(cur.p[0], tc, g0);
            for (int d = 0; d < DIM; ++d) key[d] = fabs(g0[d]);
            int order[DIM];
            order_desc(DIM, key, order);
            int k = -1;
            for (int oi = 0; oi < DIM && k < 0; ++oi) {
                bool ok = true;
                for (int f = 0; f < cur.nf; ++f)
                    if (!monotone(cur.p[f], order[oi])) { ok = false; break; }
                if (ok) k = order[oi];
            }
            bool all_roots = false;
            if (k < 0 && bx.depth >= MAXDEPTH) {
                ++ctx.fallbacks;
                k = order[0];
                all_roots = true;
            }
            if (k < 0) {   // halve the longest side; the lower half is taken first
                int sd = 0;
                for (int d = 1; d < DIM; ++d)
                    if (b[d] - a[d] > b[sd] - a[sd]) sd = d;
                if (sp + 2 > MAXDEPTH + 2) { ctx.bad = true; return; }
                const double mid = 0.5 * (a[sd] + b[sd]);
                for (int half = 1; half >= 0; --half) {
                    Box<DIM> c = bx;
                    if (half == 0) c.b[sd] = mid;
                    else c.a[sd] = mid;
                    c.path = bx.path | ((uint32_t)(sd | (half << 2)) << (3 * bx.plen));
                    c.plen = bx.plen + 1;
                    c.depth = bx.depth + 1;
                    st[sp++] = c;
                }
                continue;
            }
            // reduced problem on the box without k
            Fns<NM, DIM - 1> outer;
            outer.nf = 0;
            for (int f = 0; f < cur.nf; ++f)
                for (int side = 0; side < 2; ++side) {
                    restrict_face(cur.p[f], k, side, outer.p[outer.nf]);
                    outer.s[outer.nf++] = 0;
                }
            double ao[DIM - 1], bo[DIM - 1];
            for (int d = 0, e = 0; d < DIM; ++d)
                if (d != k) { ao[e] = a[d]; bo[e] = b[d]; ++e; }
            double fib[nfun(DIM)][DEGMAX + 1];
            vol<NM, DIM - 1>(outer, ao, bo, bx.depth, ctx, [&](const double* xe, double wo) {
                double tl[DIM], xo[DIM];
                for (int d = 0, e = 0; d < DIM; ++d)
                    if (d != k) {
                        xo[d] = xe[e];
                        tl[d] = (xo[d] - a[d]) / (b[d] - a[d]);
                        ++e;
                    }
                tl[k] = 0;
                xo[k] = 0;
                Pts bp;
                for (int f = 0; f < cur.nf; ++f) {
                    fiber(cur.p[f], k, tl, fib[f]);
                    const int n = cur.p[f].n;
                    if (all_roots) roots_all(fib[f], n, bp, ctx);
                    else if (fib[f][0] * fib[f][n] <= 0) bp.push(mono_root(fib[f], n), ctx);
                }
                pieces01(
                    bp, ctx,
                    [&](double tm) {
                        for (int f = 0; f < cur.nf; ++f)
                            if (cur.s[f] != 0 && (casteljau(fib[f], cur.p[f].n, tm) > 0 ? 1 : -1) != cur.s[f]) return false;
                        return true;
                    },
                    [&](double u, double v) {
                        for (int gq = 0; gq < q; ++gq) {
                            double x[DIM];
                            for (int d = 0; d < DIM; ++d) x[d] = d == k ? a[k] + (b[k] - a[k]) * (u + (v - u) * ctx.gt[gq]) : xo[d];
                            emit(x, wo * (b[k] - a[k]) * (v - u) * ctx.gw[gq]);
                        }
                    });
            });
        }
    }
}

// interface {phi = 0} in the box [a, b] (top-box t-space): emit(x, w, nrm) with
// physical weights and unit normals grad phi / |grad phi| (physical); h: physical
// sizes of the top box
template <int NM, int DIM, class EMIT>
BH_HD SAYE_OUTLINE void surf(const BP<NM, DIM>& top, const double* h, Ctx& ctx, EMIT&& emit) {
    Box<DIM> st[MAXDEPTH + 2];
    int sp = 0;
    {
        Box<DIM> r;
        for (int d = 0; d < DIM; ++d) { r.a[d] = 0; r.b[d] = 1; }
        r.path = 0;
        r.plen = 0;
        r.depth = 0;
        st[sp++] = r;
    }
    BP<NM, DIM> P;
    while (sp > 0 && !ctx.bad) {
        const Box<DIM> bx = st[--sp];
        const double* a = bx.a;
        const double* b = bx.b;
        P.n = top.n;
        const int N = P.size();
        for (int i = 0; i < N; ++i) P.b[i] = top.b[i];
        for (int s = 0; s < bx.plen; ++s) split(P, (int)((bx.path >> (3 * s)) & 3), (int)((bx.path >> (3 * s + 2)) & 1));
        double lo, hi;
        range(P, lo, hi);
        if (lo > 0 || hi < 0) continue;
        double tc[DIM], g0[DIM], key[DIM];
        for (int d = 0; d < DIM; ++d) tc[d] = 0.5;
        grad(P, tc, g0);
        for (int d = 0; d < DIM; ++d) key[d] = fabs(g0[d] / (h[d] * (b[d] - a[d])));   // physical derivatives
        int order[DIM];
        order_desc(DIM, key, order);
        int k = -1;
        for (int oi = 0; oi < DIM && k < 0; ++oi)
            if (monotone(P, order[oi])) k = order[oi];
        bool all_roots = false;
        if (k < 0 && bx.depth >= MAXDEPTH) {   // as in vol: every root of the fibre
            ++ctx.fallbacks;
            k = order[0];
            all_roots = true;
        }
        if (k < 0) {
            int sd = 0;
            for (int d = 1; d < DIM; ++d)
                if ((b[d] - a[d]) * h[d] > (b[sd] - a[sd]) * h[sd]) sd = d;
            if (sp + 2 > MAXDEPTH + 2) { ctx.bad = true; return; }
            const double mid = 0.5 * (a[sd] + b[sd]);
            for (int half = 1; half >= 0; --half) {
                Box<DIM> c = bx;
                if (half == 0) c.b[sd] = mid;
                else c.a[sd] = mid;
                c.path = bx.path | ((uint32_t)(sd | (half << 2)) << (3 * bx.plen));
                c.plen = bx.plen + 1;
                c.depth = bx.depth + 1;
                st[sp++] = c;
            }
            continue;
        }
        Fns<NM, DIM - 1> outer;
        outer.nf = 2;
        restrict_face(P, k, 0, outer.p[0]);
        restrict_face(P, k, 1, outer.p[1]);
        outer.s[0] = outer.s[1] = 0;
        double ao[DIM - 1], bo[DIM - 1];
        for (int d = 0, e = 0; d < DIM; ++d)
            if (d != k) { ao[e] = a[d]; bo[e] = b[d]; ++e; }
        double c[DEGMAX + 1];
        vol<NM, DIM - 1>(outer, ao, bo, bx.depth, ctx, [&](const double* xe, double wo) {
            double tl[DIM], xo[DIM];
            for (int d = 0, e = 0; d < DIM; ++d)
                if (d != k) {
                    xo[d] = xe[e];
                    tl[d] = (xo[d] - a[d]) / (b[d] - a[d]);
                    ++e;
                }
            tl[k] = 0;
            xo[k] = 0;
            fiber(P, k, tl, c);
            Pts rts;
            if (all_roots) roots_all(c, P.n, rts, ctx);
            else if (c[0] * c[P.n] <= 0) rts.push(mono_root(c, P.n), ctx);
            for (int i = 0; i < rts.n; ++i) {
                tl[k] = rts.v[i];
                double gl[DIM], gp[DIM], gn = 0, wphys = wo;
                grad(P, tl, gl);
                for (int d = 0; d < DIM; ++d) {
                    gp[d] = gl[d] / (h[d] * (b[d] - a[d]));
                    gn += gp[d] * gp[d];
                    if (d != k) wphys *= h[d];
                }
                gn = sqrt(gn);
                if (!(gn > 0) || gp[k] == 0) continue;
                double x[DIM], nv[DIM];
                for (int d = 0; d < DIM; ++d) {
                    x[d] = d == k ? a[k] + (b[k] - a[k]) * tl[k] : xo[d];
                    nv[d] = gp[d] / gn;
                }
                emit(x, wphys * gn / fabs(gp[k]), nv);
            }
        });
    }
}

// phi (power basis c[i0 + m i1 + m^2 i2], m = n + 1, in D physical coordinates) on the
// box [lo, hi] (lo[d] == hi[d]: that coordinate fixed) as a Bernstein form in the
// box's active directions: out[(n+1)^dim], act[] the physical direction of each local
// one; returns dim
template <int NM>
BH_HD SAYE_OUTLINE int from_poly(const double* coef, int D, int n, const double* lo, const double* hi, double* out, int* act) {
    constexpr int CAP = ipow(NM + 1, 3);
    const int m = n + 1;
    int N = 1;
    for (int d = 0; d < D; ++d) N *= m;
    double a[CAP], nb[CAP];
    for (int i = 0; i < N; ++i) a[i] = coef[i];
    const int S[3] = {1, m, m * m};
    // x_d = lo_d + h_d t_d: power basis in t_d
    for (int d = 0; d < D; ++d) {
        const double h = hi[d] - lo[d], x0 = lo[d];
        for (int i = 0; i < N; ++i) nb[i] = 0;
        for (int i = 0; i < N; ++i) {
            if ((i / S[d]) % m != 0) continue;
            for (int j = 0; j <= n; ++j) {
                double s = 0;
                for (int ii = j; ii <= n; ++ii) s += a[i + ii * S[d]] * binom_d(ii, j) * ipowd(x0, ii - j);
                nb[i + j * S[d]] = s * ipowd(h, j);
            }
        }
        for (int i = 0; i < N; ++i) a[i] = nb[i];
    }
    // power -> Bernstein in the active directions; fixed ones keep only degree 0
    int dim = 0;
    for (int d = 0; d < D; ++d)
        if (hi[d] > lo[d]) act[dim++] = d;
    for (int e = 0; e < dim; ++e) {
        const int d = act[e];
        for (int i = 0; i < N; ++i) nb[i] = 0;
        for (int i = 0; i < N; ++i) {
            if ((i / S[d]) % m != 0) continue;
            for (int j = 0; j <= n; ++j) {
                double s = 0;
                for (int ii = 0; ii <= j; ++ii) s += binom_d(j, ii) / binom_d(n, ii) * a[i + ii * S[d]];
                nb[i + j * S[d]] = s;
            }
        }
        for (int i = 0; i < N; ++i) a[i] = nb[i];
    }
    int outn = 1;
    for (int e = 0; e < dim; ++e) outn *= m;
    for (int i = 0; i < outn; ++i) out[i] = 0;
    for (int i = 0; i < N; ++i) {
        bool keep = true;
        int idx = 0, mul = 1;
        for (int d = 0; d < D; ++d) {
            const int id = (i / S[d]) % m;
            if (hi[d] > lo[d]) { idx += id * mul; mul *= m; }
            else if (id != 0) keep = false;
        }
        if (keep) out[idx] = a[i];
    }
    return dim;
}

// -1 phi < 0 on the box, +1 phi > 0, 0 the box may be cut (Bernstein bounds)
template <int NM>
BH_HD inline int classify(const double* coef, int D, int n, const double* lo, const double* hi) {
    double b[ipow(NM + 1, 3)];
    int act[3];
    const int dim = from_poly<NM>(coef, D, n, lo, hi, b, act);
    int N = 1;
    for (int e = 0; e < dim; ++e) N *= n + 1;
    double l = b[0], u = b[0];
    for (int i = 1; i < N; ++i) { l = fmin(l, b[i]); u = fmax(u, b[i]); }
    if (u < 0) return -1;
    if (l > 0) return +1;
    return 0;
}

// Volume rule of species s (-1: phi < 0, +1: phi > 0) in the box [lo, hi] (a face when
// lo[d] == hi[d]), or (iface) the interface rule; nodes to e(x[3], w, nrm): physical
// coordinates, physical weights, unit normals for the interface; ctx.bad when a
// capacity was exceeded.
template <int NM, int DIM, class E>
BH_HD SAYE_OUTLINE void rule_dim(const double* bern, int n, const int* act, const double* lo, const double* hi,
                                 int species, bool iface, Ctx& ctx, E& e) {
    double meas = 1, h[DIM];
    for (int i = 0; i < DIM; ++i) {
        h[i] = hi[act[i]] - lo[act[i]];
        meas *= h[i];
    }
    if (iface) {
        if constexpr (DIM == 1) {   // no interface rule on a segment
            ctx.bad = true;
        } else {
            BP<NM, DIM> P;
            P.n = n;
            for (int i = 0; i < P.size(); ++i) P.b[i] = bern[i];
            surf<NM, DIM>(P, h, ctx, [&](const double* X, double w, const double* N) {
                double x[3] = {lo[0], lo[1], lo[2]}, nv[3] = {0, 0, 0};
                for (int i = 0; i < DIM; ++i) {
                    x[act[i]] = lo[act[i]] + (hi[act[i]] - lo[act[i]]) * X[i];
                    nv[act[i]] = N[i];
                }
                e(x, w, nv);
            });
        }
    } else {
        Fns<NM, DIM> F;
        F.nf = 1;
        F.s[0] = species;
        F.p[0].n = n;
        for (int i = 0; i < F.p[0].size(); ++i) F.p[0].b[i] = bern[i];
        double a[DIM], b[DIM];
        for (int i = 0; i < DIM; ++i) { a[i] = 0; b[i] = 1; }
        vol<NM, DIM>(F, a, b, 0, ctx, [&](const double* X, double w) {
            double x[3] = {lo[0], lo[1], lo[2]};
            for (int i = 0; i < DIM; ++i) x[act[i]] = lo[act[i]] + (hi[act[i]] - lo[act[i]]) * X[i];
            e(x, w * meas, nullptr);
        });
    }
}

// the rule of one box; false when the degree exceeds NM or a capacity was exceeded
template <int NM, class E>
BH_HD inline bool rule(const double* coef, int D, int n, const double* lo, const double* hi, int species, bool iface,
                       Ctx& ctx, E& e) {
    if (n > NM) return false;
    double b[ipow(NM + 1, 3)];
    int act[3];
    const int dim = from_poly<NM>(coef, D, n, lo, hi, b, act);
    if (dim == 3) rule_dim<NM, 3>(b, n, act, lo, hi, species, iface, ctx, e);
    else if (dim == 2) rule_dim<NM, 2>(b, n, act, lo, hi, species, iface, ctx, e);
    else if (dim == 1 && !iface) rule_dim<NM, 1>(b, n, act, lo, hi, species, iface, ctx, e);
    else return false;
    return !ctx.bad;
}

}  // namespace saye
}  // namespace bh

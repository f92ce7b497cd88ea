// Cut-cell, interface and cut-face quadrature rules of a sphere level set
// phi = |x - c| - r, one implementation for the host set-up and for the device
// (one thread per rule).  The method is Saye's dimension reduction
// (X/Quadrature/SayeGaussRule_Algorithm.cs, SayeGaussRule_Cube.cs) with the
// sphere's closed-form roots: in a box the height direction k is the one of the
// largest radial component at the box centre; when r^2 - |x' - c'|^2 keeps its sign
// over the box's projection (no silhouette inside), the box is integrated as
// outer directions x' (breakpoints where the circles of the box's k-faces cross,
// and graded nodes next to the silhouette's square-root end points) times the
// height direction (pieces between the roots c_k +- sqrt(r^2 - |x' - c'|^2), a
// species sign test at each piece's midpoint); otherwise the box is split into 2^D
// children (depth <= 6), depth first, children in index order.  Interface nodes sit
// on the roots with weight w' r / |x_k - c_k| (dS = |grad phi| / |d phi / dx_k| dx')
// and the normal (x - c) / r.  Nodes are handed to an emitter e(x, w, nrm) in a fixed
// order (nrm = nullptr for volume and face nodes).
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define BH_HD __host__ __device__
#else
#define BH_HD
#endif

namespace bh {
namespace sph {

constexpr int MAXB = 24;   // breakpoints per direction and call
constexpr int QMAX = 32;   // 1D Gauss points

struct Ball {
    double c[3];
    double r;
};

BH_HD inline double ball_phi(const Ball& S, const double* x, int D) {
    double s = 0;
    for (int d = 0; d < D; ++d) s += (x[d] - S.c[d]) * (x[d] - S.c[d]);
    return sqrt(s) - S.r;
}

struct Pts {
    double v[MAXB];
    int n = 0;
    bool overflow = false;
    BH_HD void push(double x) {
        if (n < MAXB) v[n++] = x;
        else overflow = true;
    }
};

// roots c0 +- sqrt(R2) of (t - c0)^2 = R2
BH_HD inline void add_roots(Pts& v, double c0, double R2) {
    if (R2 > 0) {
        const double s = sqrt(R2);
        v.push(c0 - s);
        v.push(c0 + s);
    }
}

// [a, b] split at the breakpoints inside it (sorted, near-duplicates dropped)
BH_HD inline void pieces(double a, double b, const Pts& bp, Pts& out) {
    double s[MAXB];
    const int n = bp.n;
    for (int i = 0; i < n; ++i) {
        const double x = bp.v[i];
        int j = i;
        while (j > 0 && s[j - 1] > x) { s[j] = s[j - 1]; --j; }
        s[j] = x;
    }
    out.n = 0;
    out.overflow = bp.overflow;
    out.push(a);
    const double eps = 1e-14 * fmax(1.0, fabs(b - a));
    for (int i = 0; i < n; ++i) {
        const double v = s[i];
        if (v > a + eps && v < b - eps && v > out.v[out.n - 1] + eps) out.push(v);
    }
    out.push(b);
}

// Outer-direction nodes of a piece [a, b] whose integrand may carry a square-root
// singularity at an anchor e outside the piece or at its end (the extreme
// x_o = co +- R of a circle, where two breakpoints of the inner direction meet).  If
// the nearest anchor lies within one piece length of the piece, the piece is
// integrated in t with x_o = e + (f - e) t^2 (f the far end), which turns
// sqrt(x_o - e) into a linear function of t; otherwise plain Gauss.  X, W: 2 q slots.
BH_HD inline int outer_nodes(double a, double b, const Pts& anchors, const double* gx, const double* gw, int q, double* X,
                             double* W) {
    const double L = b - a;
    const double tol = 1e-12 * fmax(fabs(L), 1e-300);
    double el = -1e300, er = 1e300;
    for (int i = 0; i < anchors.n; ++i) {
        const double e = anchors.v[i];
        if (e <= a + tol) el = fmax(el, e);
        if (e >= b - tol) er = fmin(er, e);
    }
    const bool use_l = a - el <= L, use_r = er - b <= L;
    int n = 0;
    auto graded = [&](double e, double xn, double f) {
        const double t0 = sqrt(fmax(0.0, (xn - e) / (f - e)));
        for (int i = 0; i < q; ++i) {
            const double t = t0 + 0.5 * (gx[i] + 1) * (1.0 - t0);
            X[n] = e + (f - e) * t * t;
            W[n] = fabs(f - e) * 2.0 * t * 0.5 * (1.0 - t0) * gw[i];
            ++n;
        }
    };
    if (use_l && use_r) {
        const double m = 0.5 * (a + b);
        graded(el, a, m);
        graded(er, b, m);
    } else if (use_l) {
        graded(el, a, b);
    } else if (use_r) {
        graded(er, b, a);
    } else {
        for (int i = 0; i < q; ++i) {
            X[n] = a + 0.5 * (gx[i] + 1) * (b - a);
            W[n] = 0.5 * (b - a) * gw[i];
            ++n;
        }
    }
    return n;
}

// 2D integration over [lo_o, hi_o] x [lo_k, hi_k] of a function whose smoothness
// breaks along circles |x' - c'|^2 = R2_b (b over radii): outer breakpoints where a
// circle crosses x_k = lo / hi and at the circles' extremes x_o = co +- R (also the
// anchors of outer_nodes), inner breakpoints at the circle crossings.
// visit(x_o, x_k, w); returns false on a breakpoint overflow.
template <typename V>
BH_HD bool circle_2d(double lo_o, double hi_o, double lo_k, double hi_k, double co, double ck, const double* radii, int nr,
                     const double* gx, const double* gw, int q, V& visit) {
    Pts bo, anchors;
    for (int i = 0; i < nr; ++i) {
        const double R2 = radii[i];
        if (R2 <= 0) continue;
        add_roots(bo, co, R2 - (lo_k - ck) * (lo_k - ck));
        add_roots(bo, co, R2 - (hi_k - ck) * (hi_k - ck));
        add_roots(bo, co, R2);
        add_roots(anchors, co, R2);
    }
    Pts po;
    pieces(lo_o, hi_o, bo, po);
    bool ok = !po.overflow && !anchors.overflow;
    double X[2 * QMAX], W[2 * QMAX];
    for (int a = 0; a + 1 < po.n; ++a) {
        const int nx = outer_nodes(po.v[a], po.v[a + 1], anchors, gx, gw, q, X, W);
        for (int i = 0; i < nx; ++i) {
            const double xo = X[i], wo = W[i];
            Pts bk;
            for (int r = 0; r < nr; ++r) add_roots(bk, ck, radii[r] - (xo - co) * (xo - co));
            Pts pk;
            pieces(lo_k, hi_k, bk, pk);
            ok = ok && !pk.overflow;
            for (int b = 0; b + 1 < pk.n; ++b) {
                const double ka = pk.v[b], kb = pk.v[b + 1];
                for (int j = 0; j < q; ++j) visit(xo, ka + 0.5 * (gx[j] + 1) * (kb - ka), wo * 0.5 * (kb - ka) * gw[j]);
            }
        }
    }
    return ok;
}

BH_HD inline int argmax_abs(const double* v, const int* dims, int n) {
    int best = dims[0];
    for (int i = 1; i < n; ++i)
        if (fabs(v[dims[i]]) > fabs(v[best])) best = dims[i];
    return best;
}

// Volume (mode 0, species < 0: phi < 0) or interface (mode 1) rule of the box.
// Returns false on a breakpoint overflow or a subdivision stack overflow.
template <typename E>
BH_HD bool volume_or_interface(int D, const double* lo0, const double* hi0, const Ball& S, int species, int mode,
                               const double* gx, const double* gw, int q, E& emit) {
    constexpr int MAXDEPTH = 6, STACK = MAXDEPTH * 7 + 1;
    struct Box {
        double lo[3], hi[3];
        int depth;
    };
    Box stack[STACK];
    int top = 0;
    bool ok = true;
    for (int d = 0; d < 3; ++d) { stack[0].lo[d] = lo0[d]; stack[0].hi[d] = hi0[d]; }
    stack[0].depth = 0;
    top = 1;
    const double r2 = S.r * S.r;
    while (top > 0) {
        const Box bx = stack[--top];
        const double* lo = bx.lo;
        const double* hi = bx.hi;
        {
            // a box the sphere does not cut: tensor Gauss rule or nothing
            double dmin2 = 0, dmax2 = 0;
            for (int d = 0; d < D; ++d) {
                const double a = lo[d] - S.c[d], b = hi[d] - S.c[d];
                dmin2 += (a <= 0 && b >= 0) ? 0.0 : fmin(a * a, b * b);
                dmax2 += fmax(a * a, b * b);
            }
            const bool inside = dmax2 <= r2, outside = dmin2 >= r2;
            if (inside || outside) {
                if (mode == 1 || inside != (species < 0)) continue;
                int n = 1;
                for (int d = 0; d < D; ++d) n *= q;
                for (int t = 0; t < n; ++t) {
                    double x[3] = {0, 0, 0}, w = 1;
                    int r = t;
                    for (int d = 0; d < D; ++d) {
                        const int id = r % q;
                        r /= q;
                        x[d] = lo[d] + 0.5 * (gx[id] + 1) * (hi[d] - lo[d]);
                        w *= 0.5 * (hi[d] - lo[d]) * gw[id];
                    }
                    emit(x, w, (const double*)nullptr);
                }
                continue;
            }
        }
        double xc[3], rad[3];
        for (int d = 0; d < D; ++d) { xc[d] = 0.5 * (lo[d] + hi[d]); rad[d] = xc[d] - S.c[d]; }
        const int all[3] = {0, 1, 2};
        const int k = argmax_abs(rad, all, D);
        int od[2] = {0, 0}, no = 0;
        for (int d = 0; d < D; ++d)
            if (d != k) od[no++] = d;
        // silhouette check: r^2 - |x' - c'|^2 must not change sign over the face box
        double mn = 0, mx = 0;
        for (int i = 0; i < no; ++i) {
            const int d = od[i];
            const double a = lo[d] - S.c[d], b = hi[d] - S.c[d];
            const double m1 = (a <= 0 && b >= 0) ? 0.0 : fmin(a * a, b * b);
            mn += m1;
            mx += fmax(a * a, b * b);
        }
        const bool ok_sil = (r2 - mx > 0) || (r2 - mn <= 0);
        if (!ok_sil && bx.depth < MAXDEPTH) {
            // children in index order: pushed in reverse, so child 0 is processed first
            const int nc = 1 << D;
            if (top + nc > STACK) { ok = false; continue; }
            for (int cidx = nc - 1; cidx >= 0; --cidx) {
                Box& ch = stack[top++];
                for (int d = 0; d < 3; ++d) { ch.lo[d] = lo[d]; ch.hi[d] = hi[d]; }
                for (int d = 0; d < D; ++d) {
                    const bool up = (cidx >> d) & 1;
                    ch.lo[d] = up ? xc[d] : lo[d];
                    ch.hi[d] = up ? hi[d] : xc[d];
                }
                ch.depth = bx.depth + 1;
            }
            continue;
        }
        auto line_s2 = [&](const double* xp) {
            double s2 = r2;
            for (int i = 0; i < no; ++i) s2 -= (xp[od[i]] - S.c[od[i]]) * (xp[od[i]] - S.c[od[i]]);
            return s2;
        };
        // nodes along the height direction k above the outer point xp (weight w2)
        auto column = [&](const double* xp, double w2) {
            const double s2 = line_s2(xp);
            if (mode == 0) {
                Pts bt;
                add_roots(bt, S.c[k], s2);
                Pts pt;
                pieces(lo[k], hi[k], bt, pt);
                for (int b = 0; b + 1 < pt.n; ++b) {
                    const double ta = pt.v[b], tb = pt.v[b + 1];
                    double xm[3] = {xp[0], xp[1], xp[2]};
                    xm[k] = 0.5 * (ta + tb);
                    if ((ball_phi(S, xm, D) < 0) != (species < 0)) continue;
                    for (int j = 0; j < q; ++j) {
                        double x[3] = {xp[0], xp[1], xp[2]};
                        x[k] = ta + 0.5 * (gx[j] + 1) * (tb - ta);
                        emit(x, w2 * 0.5 * (tb - ta) * gw[j], (const double*)nullptr);
                    }
                }
            } else if (s2 > 0) {
                const double s = sqrt(s2);
                const double ts[2] = {S.c[k] - s, S.c[k] + s};
                for (int i = 0; i < 2; ++i) {
                    const double t = ts[i];
                    if (t <= lo[k] || t >= hi[k]) continue;
                    double x[3] = {xp[0], xp[1], xp[2]};
                    x[k] = t;
                    double nrm[3];
                    for (int d = 0; d < 3; ++d) nrm[d] = d < D ? (x[d] - S.c[d]) / S.r : 0.0;
                    emit(x, w2 * S.r / fabs(t - S.c[k]), (const double*)nrm);
                }
            }
        };
        if (D == 3) {
            const int d0 = od[0], d1 = od[1];
            // inner 2D direction: larger radial component
            const int k2 = fabs(rad[d0]) >= fabs(rad[d1]) ? d0 : d1;
            const int o = k2 == d0 ? d1 : d0;
            const double radii[2] = {r2 - (lo[k] - S.c[k]) * (lo[k] - S.c[k]), r2 - (hi[k] - S.c[k]) * (hi[k] - S.c[k])};
            auto visit = [&](double xo, double xk2, double w2) {
                double xp[3];
                xp[o] = xo;
                xp[k2] = xk2;
                xp[k] = 0;
                column(xp, w2);
            };
            ok = circle_2d(lo[o], hi[o], lo[k2], hi[k2], S.c[o], S.c[k2], radii, 2, gx, gw, q, visit) && ok;
        } else {
            // D == 2: one outer direction o, height direction k
            const int o = od[0];
            Pts bo;
            add_roots(bo, S.c[o], r2 - (lo[k] - S.c[k]) * (lo[k] - S.c[k]));
            add_roots(bo, S.c[o], r2 - (hi[k] - S.c[k]) * (hi[k] - S.c[k]));
            Pts po;
            pieces(lo[o], hi[o], bo, po);
            Pts anchors;
            add_roots(anchors, S.c[o], r2);   // silhouette: the height roots meet there
            ok = ok && !po.overflow;
            double XO[2 * QMAX], WO[2 * QMAX];
            for (int a = 0; a + 1 < po.n; ++a) {
                const int nxo = outer_nodes(po.v[a], po.v[a + 1], anchors, gx, gw, q, XO, WO);
                for (int i = 0; i < nxo; ++i) {
                    double xp[3] = {0, 0, 0};
                    xp[o] = XO[i];
                    column(xp, WO[i]);
                }
            }
        }
    }
    return ok;
}

// Rule of the species (species < 0: phi < 0) on the face x_dir = lo[dir] of the box
// (lo[dir] == hi[dir]): the level set on the face is a circle (3D) / two points (2D).
template <typename E>
BH_HD bool face(int D, const double* lo, const double* hi, int dir, const Ball& S, int species, const double* gx,
                const double* gw, int q, E& emit) {
    const double xd = lo[dir];
    const double R2 = S.r * S.r - (xd - S.c[dir]) * (xd - S.c[dir]);
    if (D == 2) {
        const int o = 1 - dir;
        Pts bp;
        add_roots(bp, S.c[o], R2);
        Pts p;
        pieces(lo[o], hi[o], bp, p);
        for (int a = 0; a + 1 < p.n; ++a) {
            double xm[3] = {0, 0, 0};
            xm[dir] = xd;
            xm[o] = 0.5 * (p.v[a] + p.v[a + 1]);
            if ((ball_phi(S, xm, 2) < 0) != (species < 0)) continue;
            for (int i = 0; i < q; ++i) {
                double x[3] = {0, 0, 0};
                x[dir] = xd;
                x[o] = p.v[a] + 0.5 * (gx[i] + 1) * (p.v[a + 1] - p.v[a]);
                emit(x, 0.5 * (p.v[a + 1] - p.v[a]) * gw[i], (const double*)nullptr);
            }
        }
        return !p.overflow;
    }
    int od[2], no = 0;
    for (int d = 0; d < 3; ++d)
        if (d != dir) od[no++] = d;
    const double rad0 = 0.5 * (lo[od[0]] + hi[od[0]]) - S.c[od[0]];
    const double rad1 = 0.5 * (lo[od[1]] + hi[od[1]]) - S.c[od[1]];
    const int k2 = fabs(rad0) >= fabs(rad1) ? od[0] : od[1];
    const int o = k2 == od[0] ? od[1] : od[0];
    const double radii[1] = {R2};
    // pieces never straddle the circle: keep the nodes of the requested species
    auto visit = [&](double xo, double xk, double w) {
        double x[3];
        x[dir] = xd;
        x[o] = xo;
        x[k2] = xk;
        if ((ball_phi(S, x, 3) < 0) != (species < 0)) return;
        emit(x, w, (const double*)nullptr);
    };
    return circle_2d(lo[o], hi[o], lo[k2], hi[k2], S.c[o], S.c[k2], radii, 1, gx, gw, q, visit);
}

}  // namespace sph
}  // namespace bh

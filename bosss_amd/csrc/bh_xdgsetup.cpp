// XDG setup on the host (FP64): cut-cell quadrature for a spherical level
// set, species metrics and length scales, agglomeration, element lists.
//
// Cut-cell rules: the dimension-reduction construction of Saye
// ("High-order quadrature methods for implicitly defined surfaces and volumes
// in hyperrectangles", SIAM J. Sci. Comput. 37 (2015)), which the reference
// offers as its default cut-cell rule (X/Quadrature/SayeGaussRule_*.cs,
// AppControl.cs:1323), specialised to phi = |x - c| - r: every root needed by
// the construction is a closed-form square root, so no Newton iteration is
// required.  In a box, the height direction k is the largest component of
// the radial direction at the box centre; the remaining coordinates are
// integrated with Gauss pieces split at the curves where the interface meets
// the faces x_k = lo_k / hi_k (kinks of the height integrand), and along k
// the Gauss pieces are split at the interface.  Boxes where a height
// direction is not uniformly valid (silhouette or tangency inside) are
// subdivided.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <exception>
#include <functional>
#include <mutex>
#include <stdexcept>
#include <thread>

#include "bh_sphrule.hpp"
#include "bh_xdg.hpp"

namespace bh {

namespace {

// rules through the shared implementation (bh_sphrule.hpp): the emitter appends to a QRule
struct QRuleEmit {
    QRule& R;
    void operator()(const double* x, double w, const double* nrm) {
        R.x.insert(R.x.end(), x, x + 3);
        R.w.push_back(w);
        if (nrm) R.nrm.insert(R.nrm.end(), nrm, nrm + 3);
    }
};

struct Gauss1 {
    std::vector<double> x, w;
    explicit Gauss1(int q) : x(q), w(q) { gauss_rule(q, x.data(), w.data()); }
};

sph::Ball ball_of(const Sphere& S) { return sph::Ball{{S.c[0], S.c[1], S.c[2]}, S.r}; }

}  // namespace

// The rules of different cells and faces are independent; every result lands
// in its own slot and is consumed afterwards in index order, so the output does
// not depend on the thread count.
void parallel_for(int n, const std::function<void(int)>& f) {
    const int nt = std::max(1, std::min({(int)std::thread::hardware_concurrency(), 16, n / 8}));
    if (nt <= 1) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int> next{0};
    std::mutex em;
    std::exception_ptr err;   // the first exception of any thread, rethrown on the caller's
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t)
        th.emplace_back([&] {
            try {
                for (int i = next.fetch_add(1); i < n; i = next.fetch_add(1)) f(i);
            } catch (...) {
                std::lock_guard<std::mutex> lk(em);
                if (!err) err = std::current_exception();
                next.store(n);
            }
        });
    for (auto& x : th) x.join();
    if (err) std::rethrow_exception(err);
}

void sphere_volume_rule(int D, const double* lo, const double* hi, const Sphere& S, int species, int q, QRule& out) {
    BH_CHECK(q <= sph::QMAX, "XDG cut rules: at most 32 Gauss points per direction");
    Gauss1 G(q);
    QRuleEmit e{out};
    BH_CHECK(sph::volume_or_interface(D, lo, hi, ball_of(S), species, 0, G.x.data(), G.w.data(), q, e),
             "XDG cut rules: breakpoint or subdivision capacity exceeded");
}

void sphere_interface_rule(int D, const double* lo, const double* hi, const Sphere& S, int q, QRule& out) {
    BH_CHECK(q <= sph::QMAX, "XDG cut rules: at most 32 Gauss points per direction");
    Gauss1 G(q);
    QRuleEmit e{out};
    BH_CHECK(sph::volume_or_interface(D, lo, hi, ball_of(S), 0, 1, G.x.data(), G.w.data(), q, e),
             "XDG cut rules: breakpoint or subdivision capacity exceeded");
}

// Face rule: the face is a (D-1)-box in the plane x_dir = lo[dir]; the level
// set restricted to it is a circle (3D) / two points (2D).
void sphere_face_rule(int D, const double* lo, const double* hi, int dir, const Sphere& S, int species, int q, QRule& out) {
    BH_CHECK(q <= sph::QMAX, "XDG cut rules: at most 32 Gauss points per direction");
    Gauss1 G(q);
    QRuleEmit e{out};
    BH_CHECK(sph::face(D, lo, hi, dir, ball_of(S), species, G.x.data(), G.w.data(), q, e),
             "XDG cut rules: breakpoint capacity exceeded");
}

// ------------------------------------------------------------------ setup
void XdgProblem::setup() {
    D = g->D;
    Np = basis_count(D, k);
    const double dd = D, p = k;
    penalty_deg = std::max((p + 1) * (p + dd) / dd, (p + 1.0) * (p + 1.0));   // XLaplace.cs CoefficientUpdate
    const int J = g->J;
    nspec.assign(J, 1);
    spmask.assign(J, 0);
    for (int s = 0; s < 2; ++s) { vol[s].assign(J, 0.0); len[s].assign(J, 0.0); frac[s].assign(J, 0.0); agg_target[s].assign(J, -1); }
    iface_area.assign(J, 0.0);
    elems.clear();
    nx.clear(); nw.clear(); nn.clear();
    auto box = [&](int c, double* lo, double* hi) {
        int idx[3];
        g->cell_ijk(c, idx);
        for (int d = 0; d < 3; ++d) { lo[d] = d < D ? g->nodes[d][idx[d]] : 0; hi[d] = d < D ? g->nodes[d][idx[d] + 1] : 0; }
    };
    auto push_rule = [&](const QRule& R, int& r0, int& rn) {
        r0 = (int)nw.size();
        rn = R.n();
        nx.insert(nx.end(), R.x.begin(), R.x.end());
        nw.insert(nw.end(), R.w.begin(), R.w.end());
        if (!R.nrm.empty()) nn.insert(nn.end(), R.nrm.begin(), R.nrm.end());
        else nn.insert(nn.end(), R.x.size(), 0.0);
    };
    // the level set: a sphere (closed-form roots) or a polynomial (Saye's algorithm);
    // classify: -1 phi < 0 on the box, +1 phi > 0, 0 possibly cut
    auto classify = [&](const double* lo, const double* hi) {
        if (ls_kind == 1) return poly_classify(L, lo, hi);
        double dmin2 = 0, dmax2 = 0;
        for (int d = 0; d < D; ++d) {
            const double a = lo[d] - S.c[d], b = hi[d] - S.c[d];
            dmin2 += (a <= 0 && b >= 0) ? 0.0 : std::min(a * a, b * b);
            dmax2 += std::max(a * a, b * b);
        }
        const double r2 = S.r * S.r;
        return dmax2 <= r2 ? -1 : (dmin2 >= r2 ? +1 : 0);
    };
    auto vrule = [&](const double* lo, const double* hi, int sp, QRule& R) {
        if (ls_kind == 1) poly_volume_rule(L, lo, hi, sp, q, R);
        else sphere_volume_rule(D, lo, hi, S, sp, q, R);
    };
    auto irule = [&](const double* lo, const double* hi, QRule& R) {
        if (ls_kind == 1) poly_interface_rule(L, lo, hi, q, R);
        else sphere_interface_rule(D, lo, hi, S, q, R);
    };
    auto frule_of = [&](const double* lo, const double* hi, int dir, int sp, QRule& R) {
        if (ls_kind == 1) poly_face_rule(L, lo, hi, dir, sp, q, R);
        else sphere_face_rule(D, lo, hi, dir, S, sp, q, R);
    };
    std::vector<int> cut(J, 0);
    std::vector<int> vol_rule0[2], vol_rulen[2], if_r0(J, 0), if_rn(J, 0);
    for (int s = 0; s < 2; ++s) { vol_rule0[s].assign(J, 0); vol_rulen[s].assign(J, 0); }
    ncut = 0;
    // classification of many boxes at once: a polynomial's on the device (degree <= 4),
    // otherwise on the host threads
    auto classify_batch = [&](const std::vector<SphereRuleItem>& bx, std::vector<int>& out) {
        if (ls_kind == 1 && ctx != nullptr && poly_rules_on_device(L)) {
            poly_classify_device(ctx, L, bx, out);
            return;
        }
        out.assign(bx.size(), 0);
        parallel_for((int)bx.size(), [&](int i) { out[i] = classify(bx[i].lo, bx[i].hi); });
    };
    // the boxes of the faces, in the order the face pass below visits them: per cell,
    // per direction, the '-' side on the boundary only, then the '+' side
    auto face_box = [&](int c, int d, int s, double* fl, double* fh) {
        double lo[3], hi[3];
        box(c, lo, hi);
        for (int e = 0; e < 3; ++e) { fl[e] = lo[e]; fh[e] = hi[e]; }
        fl[d] = fh[d] = s ? hi[d] : lo[d];
    };
    std::vector<int> ccl, fcl;
    {
        std::vector<SphereRuleItem> bx(J);
        for (int c = 0; c < J; ++c) box(c, bx[c].lo, bx[c].hi);
        classify_batch(bx, ccl);
        bx.clear();
        for (int c = 0; c < J; ++c) {
            int idx[3];
            g->cell_ijk(c, idx);
            for (int d = 0; d < D; ++d)
                for (int s = 0; s < 2; ++s) {
                    const bool bnd = s == 0 ? idx[d] == 0 : idx[d] == g->n[d] - 1;
                    if (!bnd && s == 0) continue;
                    SphereRuleItem it{};
                    face_box(c, d, s, it.lo, it.hi);
                    bx.push_back(it);
                }
        }
        classify_batch(bx, fcl);
    }
    // pass 1: classify; candidates get their species and interface rules computed in parallel
    std::vector<int> cand_of(J, -1), cand;
    for (int c = 0; c < J; ++c) {
        double lo[3], hi[3];
        box(c, lo, hi);
        double V = 1;
        for (int d = 0; d < D; ++d) V *= hi[d] - lo[d];
        const int cl = ccl[c];
        if (cl < 0) { spmask[c] = 1; vol[0][c] = V; }
        else if (cl > 0) { spmask[c] = 2; vol[1][c] = V; }
        else { cand_of[c] = (int)cand.size(); cand.push_back(c); }
    }
    // rules kind-major (species A volumes, species B volumes, interfaces): on the device
    // neighbouring threads then run the same kind of construction
    const size_t nc = cand.size();
    std::vector<QRule> crule(3 * nc);
    // cut rules on the device: the sphere's, and a polynomial's of degree <= 4
    const bool dev_rules = ctx != nullptr && (ls_kind == 0 || poly_rules_on_device(L));
    auto rules_device = [&](const std::vector<SphereRuleItem>& items, std::vector<QRule>& out) {
        if (ls_kind == 1) poly_rules_device(ctx, L, q, items, out);
        else sphere_rules_device(ctx, D, q, S, items, out);
    };
    if (dev_rules) {
        std::vector<SphereRuleItem> items(3 * cand.size());
        for (size_t i = 0; i < cand.size(); ++i)
            for (int t = 0; t < 3; ++t) {
                SphereRuleItem& it = items[t * nc + i];
                box(cand[i], it.lo, it.hi);
                it.kind = t;
                it.dir = 0;
            }
        rules_device(items, crule);
    } else {
        parallel_for((int)cand.size(), [&](int i) {
            double lo[3], hi[3];
            box(cand[i], lo, hi);
            vrule(lo, hi, -1, crule[i]);
            vrule(lo, hi, +1, crule[nc + i]);
            irule(lo, hi, crule[2 * nc + i]);
        });
    }
    // pass 2, in cell order: species volumes, masks and rule slots
    for (int c = 0; c < J; ++c) {
        double lo[3], hi[3];
        box(c, lo, hi);
        double V = 1;
        for (int d = 0; d < D; ++d) V *= hi[d] - lo[d];
        if (cand_of[c] >= 0) {
            const QRule& RA = crule[cand_of[c]];
            const QRule& RB = crule[nc + cand_of[c]];
            const QRule& RI = crule[2 * nc + cand_of[c]];
            double va = 0, vb = 0, ai = 0;
            for (double w : RA.w) va += w;
            for (double w : RB.w) vb += w;
            for (double w : RI.w) ai += w;
            vol[0][c] = va;
            vol[1][c] = vb;
            iface_area[c] = ai;
            int m = (va > 0 ? 1 : 0) | (vb > 0 ? 2 : 0);
            spmask[c] = m;
            if (m == 3) {
                cut[c] = 1;
                ++ncut;
                push_rule(RA, vol_rule0[0][c], vol_rulen[0][c]);
                push_rule(RB, vol_rule0[1][c], vol_rulen[1][c]);
                push_rule(RI, if_r0[c], if_rn[c]);
            }
        }
        nspec[c] = (spmask[c] == 3) ? 2 : 1;
        for (int s = 0; s < 2; ++s) frac[s][c] = vol[s][c] / V;
    }
    // faces: per species, area and rule; (cell, dir) '+' faces and boundary faces
    struct FaceInfo { int a, b, dir, side; double area[2]; int r0[2], rn[2]; bool tensor[2]; };
    std::vector<FaceInfo> faces;
    struct CutFace { int face; double lo[3], hi[3]; };
    std::vector<CutFace> cut_faces;
    std::vector<double> edge_sum[2];
    for (int s = 0; s < 2; ++s) edge_sum[s].assign(J, 0.0);
    size_t fpos = 0;
    for (int c = 0; c < J; ++c) {
        int idx[3];
        g->cell_ijk(c, idx);
        double lo[3], hi[3];
        box(c, lo, hi);
        for (int d = 0; d < D; ++d)
            for (int s = 0; s < 2; ++s) {
                const bool bnd = s == 0 ? idx[d] == 0 : idx[d] == g->n[d] - 1;
                if (!bnd && s == 0) continue;   // interior faces once, from the '-' cell
                FaceInfo F{};
                F.a = c;
                F.dir = d;
                F.side = s;
                if (bnd) F.b = -1;
                else { int nb[3] = {idx[0], idx[1], idx[2]}; nb[d]++; F.b = g->cell_index(nb[0], nb[1], nb[2]); }
                double fl[3], fh[3];
                face_box(c, d, s, fl, fh);
                double A = 1;
                for (int e = 0; e < D; ++e) if (e != d) A *= fh[e] - fl[e];
                // face cut?
                const int cl = fcl[fpos++];
                for (int sp = 0; sp < 2; ++sp) { F.area[sp] = 0; F.rn[sp] = 0; F.tensor[sp] = false; }
                if (cl < 0) { F.area[0] = A; F.tensor[0] = true; }
                else if (cl > 0) { F.area[1] = A; F.tensor[1] = true; }
                else {
                    CutFace cf;
                    cf.face = (int)faces.size();
                    for (int e = 0; e < 3; ++e) { cf.lo[e] = fl[e]; cf.hi[e] = fh[e]; }
                    cut_faces.push_back(cf);
                }
                faces.push_back(F);
            }
    }
    // species rules of the cut faces, in parallel, then areas and rule slots in face order
    const size_t nfc = cut_faces.size();   // species-major, as the cell rules
    std::vector<QRule> frule(2 * nfc);
    if (dev_rules) {
        std::vector<SphereRuleItem> items(2 * cut_faces.size());
        for (size_t i = 0; i < cut_faces.size(); ++i)
            for (int sp = 0; sp < 2; ++sp) {
                SphereRuleItem& it = items[sp * nfc + i];
                for (int e = 0; e < 3; ++e) { it.lo[e] = cut_faces[i].lo[e]; it.hi[e] = cut_faces[i].hi[e]; }
                it.kind = 3 + sp;
                it.dir = faces[cut_faces[i].face].dir;
            }
        rules_device(items, frule);
    } else {
        parallel_for((int)cut_faces.size(), [&](int i) {
            const CutFace& cf = cut_faces[i];
            for (int sp = 0; sp < 2; ++sp)
                frule_of(cf.lo, cf.hi, faces[cf.face].dir, sp == 0 ? -1 : +1, frule[sp * nfc + i]);
        });
    }
    for (size_t i = 0; i < cut_faces.size(); ++i) {
        FaceInfo& F = faces[cut_faces[i].face];
        for (int sp = 0; sp < 2; ++sp) {
            const QRule& R = frule[sp * nfc + i];
            double a = 0;
            for (double w : R.w) a += w;
            F.area[sp] = a;
            if (a > 0) push_rule(R, F.r0[sp], F.rn[sp]);
        }
    }
    for (const FaceInfo& F : faces)
        for (int sp = 0; sp < 2; ++sp) {
            edge_sum[sp][F.a] += F.area[sp];
            if (F.b >= 0) edge_sum[sp][F.b] += F.area[sp];
        }
    // face index of every (cell, local face 2d+s)
    std::vector<int> cell_face((size_t)J * 6, -1);
    for (size_t fi = 0; fi < faces.size(); ++fi) {
        const FaceInfo& F = faces[fi];
        cell_face[(size_t)F.a * 6 + 2 * F.dir + F.side] = (int)fi;
        if (F.b >= 0) cell_face[(size_t)F.b * 6 + 2 * F.dir + 0] = (int)fi;
    }
    // agglomeration, per species (X/AgglomerationAlgorithm.cs, one rank, static
    // interface): sources :393-460; "sources cannot be targets" :1721-1726; direct
    // targets SearchDirectAgglomeration_Mk3 :2064-2200; chains
    // SearchChainAgglomeration_Mk3 :1815-2055; groups SearchGroupAgglomeration_Mk3
    // :2232-2330 with AgglomerationGroup.AddNeighbors :72-111; self pairs removed
    agg_pairs.clear();
    n_failed = 0;
    const double EmptyEdge = 1e-10;
    auto center = [&](int c, double* x) {
        double lo[3], hi[3];
        box(c, lo, hi);
        for (int d = 0; d < 3; ++d) x[d] = 0.5 * (lo[d] + hi[d]);
    };
    for (int sp = 0; sp < 2; ++sp) {
        std::vector<char> source(J, 0), cand(J, 0);
        for (int c = 0; c < J; ++c) {
            double lo[3], hi[3];
            box(c, lo, hi);
            double V = 1;
            for (int d = 0; d < D; ++d) V *= hi[d] - lo[d];
            const bool present = (spmask[c] >> sp) & 1;
            const double f = std::min(1.0, std::max(0.0, std::max(vol[sp][c], 0.0) / V));
            source[c] = cut[c] && present && f <= agg_threshold;
            cand[c] = present && !source[c];
        }
        // neighbours and species cut-edge areas per face, per direction '-' then '+'
        auto neighbours = [&](int c, int* nbs, double* areas) {
            int idx[3];
            g->cell_ijk(c, idx);
            int t = 0;
            for (int d = 0; d < D; ++d)
                for (int s = 0; s < 2; ++s, ++t) {
                    int nb[3] = {idx[0], idx[1], idx[2]};
                    nb[d] += s ? 1 : -1;
                    const bool inside = nb[d] >= 0 && nb[d] < g->n[d];
                    nbs[t] = inside ? g->cell_index(nb[0], nb[1], nb[2]) : -1;
                    const int fi = cell_face[(size_t)c * 6 + 2 * d + s];
                    areas[t] = fi >= 0 ? faces[fi].area[sp] : 0.0;
                }
            return t;
        };
        struct Pair { int src, tgt; double frac; double pos[3]; };
        std::vector<Pair> pairs;
        std::vector<int> pair_of(J, -1);          // pair index per source, visible pairs only
        auto add_pairs = [&](const std::vector<Pair>& ps) {
            for (const Pair& q : ps) {
                if (q.src != q.tgt && pair_of[q.src] >= 0) throw std::runtime_error("agglomeration: a cell is mapped to two targets");
                pair_of[q.src] = (int)pairs.size();
                pairs.push_back(q);
            }
        };
        std::vector<int> need;
        int failed = 0;
        // direct pass
        for (int c = 0; c < J; ++c) {
            if (!source[c]) continue;
            int nbs[6];
            double areas[6];
            const int nn = neighbours(c, nbs, areas);
            bool nonEmpty = false;
            for (int i = 0; i < nn; ++i) if (nbs[i] >= 0 && areas[i] > EmptyEdge) nonEmpty = true;
            if (vol[sp][c] <= 0) nonEmpty = false;
            double fmax = -1.0;
            int tgt = -1;
            for (int i = 0; i < nn; ++i) {
                const int nb = nbs[i];
                if (nb < 0) continue;
                if (areas[i] <= EmptyEdge && nonEmpty) continue;
                if (!cand[nb]) continue;
                if (frac[sp][nb] > fmax) { fmax = frac[sp][nb]; tgt = nb; }
            }
            if (tgt < 0) {
                bool chainable = false;
                for (int i = 0; i < nn; ++i) if (nbs[i] >= 0 && source[nbs[i]]) chainable = true;
                if (chainable) need.push_back(c);
                else ++failed;
                continue;
            }
            Pair q{c, tgt, fmax, {0, 0, 0}};
            center(tgt, q.pos);
            add_pairs({q});
        }
        auto remove_need = [&](int c) { need.erase(std::find(need.begin(), need.end(), c)); };
        // chain pass: connect waiting sources to the final target of a paired neighbour
        auto chain = [&]() {
            struct WEdge { double dist, frac; int cell, pair; };
            auto before = [](const WEdge& a, const WEdge& b) { return a.dist < b.dist || (a.dist == b.dist && a.frac > b.frac); };
            bool update = true;
            while (!need.empty() && update) {
                update = false;
                std::vector<Pair> loop;
                std::vector<WEdge> weighted;
                std::vector<int> immediate;
                for (int c : need) {
                    int nbs[6];
                    double areas[6], cc[3];
                    const int nn = neighbours(c, nbs, areas);
                    center(c, cc);
                    int nsrc = 0, npaired = 0;
                    std::vector<WEdge> wc;
                    for (int i = 0; i < nn; ++i) {
                        const int nb = nbs[i];
                        if (nb < 0) continue;
                        if (source[nb]) ++nsrc;
                        const int pi = pair_of[nb];
                        if (pi < 0) continue;
                        double s2 = 0;
                        for (int d = 0; d < D; ++d) s2 += (pairs[pi].pos[d] - cc[d]) * (pairs[pi].pos[d] - cc[d]);
                        wc.push_back({std::sqrt(s2), pairs[pi].frac, c, pi});
                        ++npaired;
                    }
                    if (npaired == nsrc && npaired > 0) {
                        std::stable_sort(wc.begin(), wc.end(), before);
                        const Pair& tp = pairs[wc[0].pair];
                        Pair q{c, tp.tgt, tp.frac, {tp.pos[0], tp.pos[1], tp.pos[2]}};
                        loop.push_back(q);
                        immediate.push_back(c);
                    } else {
                        weighted.insert(weighted.end(), wc.begin(), wc.end());
                    }
                }
                for (int c : immediate) { remove_need(c); update = true; }
                std::stable_sort(weighted.begin(), weighted.end(), before);
                if (!weighted.empty() && !need.empty()) {
                    const Pair& tp = pairs[weighted[0].pair];
                    Pair q{weighted[0].cell, tp.tgt, tp.frac, {tp.pos[0], tp.pos[1], tp.pos[2]}};
                    loop.push_back(q);
                    remove_need(weighted[0].cell);
                    update = true;
                }
                add_pairs(loop);
            }
        };
        // neighbours that join a group: interior faces with a non-empty cut edge,
        // or all interior faces when every cut edge is empty
        auto group_neighbours = [&](int c, std::vector<int>& out) {
            int nbs[6];
            double areas[6];
            const int nn = neighbours(c, nbs, areas);
            bool nonEmpty = false;
            for (int i = 0; i < nn; ++i) if (nbs[i] >= 0 && areas[i] > EmptyEdge) nonEmpty = true;
            for (int i = 0; i < nn; ++i)
                if (nbs[i] >= 0 && (areas[i] > EmptyEdge || !nonEmpty)) out.push_back(nbs[i]);
        };
        chain();
        while (!need.empty()) {
            int gt = need[0];
            for (int c : need)
                if (vol[sp][c] > vol[sp][gt] || (vol[sp][c] == vol[sp][gt] && c < gt)) gt = c;
            std::vector<int> members{gt}, nbl;
            group_neighbours(gt, nbl);
            remove_need(gt);
            for (;;) {
                std::vector<int> inter;
                for (int c : nbl)
                    if (std::find(inter.begin(), inter.end(), c) == inter.end() &&
                        std::find(need.begin(), need.end(), c) != need.end()) inter.push_back(c);
                if (inter.empty()) break;
                for (int c : inter) {
                    if (std::find(members.begin(), members.end(), c) == members.end()) {
                        members.push_back(c);
                        group_neighbours(c, nbl);
                    }
                    remove_need(c);
                }
            }
            std::vector<Pair> gp;
            for (int c : members) {
                Pair q{c, gt, vol[sp][gt], {0, 0, 0}};
                center(gt, q.pos);
                gp.push_back(q);
            }
            add_pairs(gp);
            chain();
        }
        n_failed += failed;
        for (const Pair& q : pairs) {
            if (q.src == q.tgt) continue;
            agg_target[sp][q.src] = q.tgt;
            agg_pairs.push_back(q.src);
            agg_pairs.push_back(q.tgt);
            agg_pairs.push_back(sp);
        }
    }
    // length scales (MultiphaseAgglomerator.cs:870-1030): Volume / (interface area +
    // cut-edge areas excluding agglomeration edges), accumulated over agglomeration pairs
    std::vector<double> surf[2], volA[2];
    for (int sp = 0; sp < 2; ++sp) {
        surf[sp].assign(J, 0.0);
        volA[sp] = vol[sp];
        for (int c = 0; c < J; ++c) surf[sp][c] = ((spmask[c] >> sp) & 1) ? iface_area[c] : 0.0;
        // agglomeration edges: faces between two cells of one group, target and
        // sources (CellAgglomerator.cs:418-446)
        std::vector<int> gid(J, -1);
        for (int c = 0; c < J; ++c)
            if (agg_target[sp][c] >= 0) { gid[c] = agg_target[sp][c]; gid[agg_target[sp][c]] = agg_target[sp][c]; }
        for (auto& F : faces) {
            const bool aggEdge = F.b >= 0 && gid[F.a] >= 0 && gid[F.a] == gid[F.b];
            if (aggEdge) continue;
            surf[sp][F.a] += F.area[sp];
            if (F.b >= 0) surf[sp][F.b] += F.area[sp];
        }
        for (int c = 0; c < J; ++c) {
            const int t = agg_target[sp][c];
            if (t >= 0) { surf[sp][t] += surf[sp][c]; volA[sp][t] += volA[sp][c]; }
        }
        for (int c = 0; c < J; ++c) {
            const int t = agg_target[sp][c];
            if (t >= 0) { surf[sp][c] = surf[sp][t]; volA[sp][c] = volA[sp][t]; }
        }
        for (int c = 0; c < J; ++c) len[sp][c] = surf[sp][c] > 1e-25 ? volA[sp][c] / surf[sp][c] : 0.0;
    }
    // elements
    const double pb = penalty_base * penalty_deg;
    for (int c = 0; c < J; ++c)
        for (int sp = 0; sp < 2; ++sp) {
            if (!((spmask[c] >> sp) & 1)) continue;
            XElem E{};
            E.type = 0; E.cellP = E.cellQ = c; E.spP = E.spQ = sp; E.dir = -1; E.side = 0;
            E.r0 = cut[c] ? vol_rule0[sp][c] : 0;
            E.rn = cut[c] ? vol_rulen[sp][c] : 0;
            if (cut[c] && E.rn == 0) continue;
            E.nuP = E.nuQ = nu[sp];
            E.mu = 0;
            elems.push_back(E);
        }
    for (int c = 0; c < J; ++c)
        if (cut[c] && if_rn[c] > 0) {
            XElem E{};
            E.type = 3; E.cellP = E.cellQ = c; E.spP = 0; E.spQ = 1; E.dir = -1; E.side = 0;
            E.r0 = if_r0[c]; E.rn = if_rn[c];
            E.nuP = nu[0]; E.nuQ = nu[1];
            // XLaplace_Interface.GetPenalty (SIP mode): max(1/l_A, 1/l_B) * max(|muA|,|muB|)*sign(muA)
            E.mu = penalty_base * penalty_deg * std::max(1.0 / len[0][c], 1.0 / len[1][c]) *
                   std::max(std::fabs(nu[0]), std::fabs(nu[1])) * (nu[0] >= 0 ? 1.0 : -1.0);
            elems.push_back(E);
        }
    for (auto& F : faces)
        for (int sp = 0; sp < 2; ++sp) {
            if (F.area[sp] <= 0) continue;
            XElem E{};
            E.dir = F.dir;
            E.spP = E.spQ = sp;
            E.cellP = F.a;
            E.nuP = E.nuQ = nu[sp];
            E.r0 = F.tensor[sp] ? 0 : F.r0[sp];
            E.rn = F.tensor[sp] ? 0 : F.rn[sp];
            if (F.b < 0) {
                E.type = 2; E.cellQ = F.a; E.side = F.side;
                E.mu = 2.0 * pb / len[sp][F.a];                            // SIPLaplace boundary: 2 * penalty(jIn)
            } else {
                if (!((spmask[F.b] >> sp) & 1) || !((spmask[F.a] >> sp) & 1)) continue;
                E.type = 1; E.cellQ = F.b; E.side = 1;
                E.mu = pb * std::max(1.0 / len[sp][F.a], 1.0 / len[sp][F.b]) * nu[sp];   // XLaplace_Bulk.GetPenalty
            }
            elems.push_back(E);
        }
    // DOF maps
    off_full.assign(J + 1, 0);
    bsz_full.assign(J, 0);
    off_red.assign(J + 1, 0);
    bsz_red.assign(J, 0);
    for (int c = 0; c < J; ++c) {
        int nf = 0, nr = 0;
        for (int sp = 0; sp < 2; ++sp)
            if ((spmask[c] >> sp) & 1) { ++nf; if (agg_target[sp][c] < 0) ++nr; }
        bsz_full[c] = nf * Np;
        bsz_red[c] = nr * Np;
        off_full[c + 1] = off_full[c] + bsz_full[c];
        off_red[c + 1] = off_red[c] + bsz_red[c];
    }
    N_full = off_full[J];
    N_red = off_red[J];
}

}  // namespace bh

// XDG (extended DG) two-phase Poisson with a spherical level set: cut-cell
// quadrature, agglomeration, element lists and the variable-block matrix.
#pragma once
#include <functional>
#include <vector>

#include "bh_internal.hpp"

namespace bh {

// Quadrature rule with physical nodes (n x 3), weights and (interface only) normals
struct QRule {
    std::vector<double> x, w, nrm;
    int n() const { return (int)w.size(); }
};

struct Sphere {
    double c[3];
    double r;
    double phi(const double* x, int D) const {
        double s = 0;
        for (int d = 0; d < D; ++d) s += (x[d] - c[d]) * (x[d] - c[d]);
        return std::sqrt(s) - r;
    }
};

// Polynomial level set phi(x) = sum over i0, i1, i2 <= n of
// c[i0 + (n + 1) (i1 + (n + 1) i2)] x^i0 y^i1 z^i2 (2D: i2 = 0); species A is phi < 0
struct PolyLS {
    int D = 3, n = 0;
    std::vector<double> c;
    double eval(const double* x) const;
};
// Cut-cell rules of a polynomial level set on axis-aligned boxes (bh_saye.cpp):
// poly_classify: -1 phi < 0 on the box, +1 phi > 0, 0 the box may be cut (Bernstein
// bounds); a face has lo[d] == hi[d] in its normal direction.  species -1 = A, +1 = B;
// q Gauss points per 1D piece.  poly_fallbacks: boxes where no height direction was
// found at the subdivision limit (a low-order rule was used there).
int poly_classify(const PolyLS& L, const double* lo, const double* hi);
void poly_volume_rule(const PolyLS& L, const double* lo, const double* hi, int species, int q, QRule& out);
void poly_interface_rule(const PolyLS& L, const double* lo, const double* hi, int q, QRule& out);
void poly_face_rule(const PolyLS& L, const double* lo, const double* hi, int dir, int species, int q, QRule& out);
long poly_fallbacks();

// Cut-cell rules for a sphere and an axis-aligned box (bh_xdgsetup.cpp).
// species: -1 = inside (phi < 0, "A"), +1 = outside ("B").
void sphere_volume_rule(int D, const double* lo, const double* hi, const Sphere& S, int species, int q, QRule& out);
// f(i) for i in [0, n) on up to 16 host threads (f must write disjoint slots
// and allocate little: page faults of freshly allocated memory serialise)
void parallel_for(int n, const std::function<void(int)>& f);
void sphere_interface_rule(int D, const double* lo, const double* hi, const Sphere& S, int q, QRule& out);
// face in the plane x_dir = lo[dir] (lo[dir] == hi[dir])
void sphere_face_rule(int D, const double* lo, const double* hi, int dir, const Sphere& S, int species, int q, QRule& out);
// The same rules on the device, one thread per item (bh_xdgrule.hip); out[i] is the
// rule of items[i].  kind: 0 / 1 volume of species A / B, 2 interface, 3 / 4 face
// of species A / B (face in the plane x_dir = lo[dir])
struct SphereRuleItem {
    double lo[3], hi[3];
    int kind, dir;
};
void sphere_rules_device(bh_ctx* ctx, int D, int q, const Sphere& S, const std::vector<SphereRuleItem>& items,
                         std::vector<QRule>& out);
// the same items for a polynomial level set (bh_sayerule.hpp on the device, degree <= 4)
bool poly_rules_on_device(const PolyLS& L);
void poly_rules_device(bh_ctx* ctx, const PolyLS& L, int q, const std::vector<SphereRuleItem>& items, std::vector<QRule>& out);
// poly_classify of the boxes [lo, hi] of the items on the device (kind, dir unused)
void poly_classify_device(bh_ctx* ctx, const PolyLS& L, const std::vector<SphereRuleItem>& items, std::vector<int>& out);
// Gauss-Legendre points and weights on [0, 1]
void gauss01(int q, double* t, double* w);
void poly_fallbacks_add(long n);

// element of the XDG bilinear form (one wave of the assembly kernel each)
struct XElem {
    int type;        // 0 volume, 1 interior face, 2 boundary face, 3 interface
    int cellP, cellQ;
    int spP, spQ;    // species index (0 = A, 1 = B) of each side
    int dir;         // face normal direction (faces), -1 otherwise
    int side;        // boundary side / +1 if normal = +e_dir from P to Q
    int r0, rn;      // rule slice [r0, r0+rn) in the node arrays, rn = 0: tensor rule
    double nuP, nuQ, mu;
};

struct XdgProblem {
    bh_ctx* ctx;
    bh_grid* g;
    int D, k, Np, q;
    Sphere S;
    int ls_kind = 0;                 // 0: the sphere S (closed-form roots), 1: the polynomial L
    PolyLS L;
    double nu[2];
    double penalty_base, penalty_deg, agg_threshold;
    // per cell
    std::vector<int> nspec;          // 1 or 2
    std::vector<int> spmask;         // bit 0: A present, bit 1: B present
    std::vector<double> vol[2];      // species volumes
    std::vector<double> iface_area;  // interface area per cell
    std::vector<double> len[2];      // agglomerated length scales per species
    std::vector<double> frac[2];     // volume fractions
    // elements and node storage
    std::vector<XElem> elems;
    std::vector<double> nx, nw, nn;  // nodes (3 per), weights, normals (3 per)
    // agglomeration: per species, source cell -> target cell (-1 none)
    std::vector<int> agg_target[2];
    std::vector<int> agg_pairs;      // flattened (source, target, species) triples, in discovery order
    int n_failed = 0;
    // DOF maps: full (every present species block) and reduced (agglomeration sources removed)
    std::vector<int64_t> off_full, off_red;
    std::vector<int> bsz_full, bsz_red;   // row block sizes
    int64_t N_full = 0, N_red = 0;
    int ncut = 0;
    void setup();
};

}  // namespace bh

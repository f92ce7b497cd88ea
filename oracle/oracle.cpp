// CPU oracle (test infrastructure; see oracle.hpp).  FP64 restatement of the
// BoSSS operator-evaluate / assembly / block-solver path.  Exposed through a
// plain C API (bottom of this file) that tests load with ctypes.
#include "oracle.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// =================================================================== basis ==
// Binomial(k+D, D) modes (F/Basis.cs Polynomials up to degree k).
int basis_count(int D, int k) {
    int n = 1;
    for (int i = 1; i <= D; ++i) n = n * (k + i) / i;
    return n;
}

// Mode order of the reference tables: degree-graded; in 2D within degree d the
// x-exponent runs 0..d (F/Square.cs:197-260: mode 1 = {0,1} i.e. eta, mode 2 =
// {1,0} i.e. xi); in 3D x-exponent 0..d, then y-exponent 0..d-ex
// (F/Cube_polynomials.cs:20-60: modes 1..3 = z, y, x).
void basis_exponents(int D, int k, std::vector<int>& ex) {
    ex.clear();
    for (int deg = 0; deg <= k; ++deg) {
        if (D == 1) {
            ex.push_back(deg);
        } else if (D == 2) {
            for (int a = 0; a <= deg; ++a) { ex.push_back(a); ex.push_back(deg - a); }
        } else {
            for (int a = 0; a <= deg; ++a)
                for (int b = 0; b <= deg - a; ++b) {
                    ex.push_back(a); ex.push_back(b); ex.push_back(deg - a - b);
                }
        }
    }
}

// Normalised Legendre Ptilde_n(x) = sqrt((2n+1)/2) P_n(x) and derivative, by
// the three-term recurrence.
static void legendre(int nmax, double x, double* p, double* dp) {
    p[0] = 1.0; dp[0] = 0.0;
    if (nmax >= 1) { p[1] = x; dp[1] = 1.0; }
    for (int n = 2; n <= nmax; ++n) {
        p[n] = ((2 * n - 1) * x * p[n - 1] - (n - 1) * p[n - 2]) / n;
        dp[n] = dp[n - 2] + (2 * n - 1) * p[n - 1];
    }
    for (int n = 0; n <= nmax; ++n) {
        double s = std::sqrt((2.0 * n + 1.0) / 2.0);
        p[n] *= s; dp[n] *= s;
    }
}

void basis_eval(int D, int k, const double* xi, double* val, double* grad) {
    std::vector<int> ex;
    basis_exponents(D, k, ex);
    int Np = (int)ex.size() / D;
    double p[3][32], dp[3][32];
    for (int d = 0; d < D; ++d) legendre(k, xi[d], p[d], dp[d]);
    for (int n = 0; n < Np; ++n) {
        double v = 1.0;
        for (int d = 0; d < D; ++d) v *= p[d][ex[n * D + d]];
        if (val) val[n] = v;
        if (grad)
            for (int d = 0; d < D; ++d) {
                double g = 1.0;
                for (int e = 0; e < D; ++e) g *= (e == d) ? dp[e][ex[n * D + e]] : p[e][ex[n * D + e]];
                grad[n * D + d] = g;
            }
    }
}

// Gauss-Legendre nodes by Newton iteration in long double (ascending order, as
// the reference tables F/LineQuadRules/Line{n}.txt list them).
void gauss_rule(int n, double* x, double* w) {
    for (int i = 0; i < n; ++i) {
        long double z = std::cos(M_PI * (i + 0.75L) / (n + 0.5L));
        long double pp = 0;
        for (int it = 0; it < 100; ++it) {
            long double p1 = 1, p2 = 0;
            for (int j = 1; j <= n; ++j) {
                long double p3 = p2; p2 = p1;
                p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
            }
            pp = n * (z * p1 - p2) / (z * z - 1);
            long double z1 = z;
            z = z1 - p1 / pp;
            if (std::fabs((double)(z - z1)) < 1e-19) break;
        }
        // recompute derivative at the converged node
        long double p1 = 1, p2 = 0;
        for (int j = 1; j <= n; ++j) {
            long double p3 = p2; p2 = p1;
            p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
        }
        pp = n * (z * p1 - p2) / (z * z - 1);
        x[n - 1 - i] = (double)z;
        w[n - 1 - i] = (double)(2.0L / ((1 - z * z) * pp * pp));
    }
}

// ==================================================================== grid ==
// Cell numbering of the reference generators: 2D GlobalID = i*nY + j
// (BoSSS.Foundation.Grid/Grid2D.cs:586), 3D GlobalID = i + j*nX + k*nX*nY
// (BoSSS.Foundation.Grid/Grid3D.cs:261).
int Grid::cell_index(int i, int j, int k) const {
    if (D == 2) return i * n[1] + j;
    return i + j * n[0] + k * n[0] * n[1];
}

Grid* grid_create(int D, const double* x, int nx1, const double* y, int ny1, const double* z, int nz1) {
    Grid* g = new Grid();
    g->D = D;
    const double* src[3] = {x, y, z};
    int len[3] = {nx1, ny1, nz1};
    for (int d = 0; d < D; ++d) {
        g->nodes[d].assign(src[d], src[d] + len[d]);
        g->n[d] = len[d] - 1;
    }
    g->J = g->n[0] * g->n[1] * (D == 3 ? g->n[2] : 1);
    g->ijk.assign(3 * g->J, 0);
    g->lo.assign(D * g->J, 0);
    g->h.assign(D * g->J, 0);
    int nz = D == 3 ? g->n[2] : 1;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                int c = g->cell_index(i, j, k);
                int idx[3] = {i, j, k};
                for (int d = 0; d < 3; ++d) g->ijk[3 * c + d] = idx[d];
                for (int d = 0; d < D; ++d) {
                    g->lo[D * c + d] = g->nodes[d][idx[d]];
                    g->h[D * c + d] = g->nodes[d][idx[d] + 1] - g->nodes[d][idx[d]];
                }
            }
    // edges: per cell (index order), per direction: '-' face if on boundary,
    // '+' face (interior or boundary)
    for (int c = 0; c < g->J; ++c) {
        for (int d = 0; d < D; ++d) {
            int idx[3] = {g->ijk[3 * c], g->ijk[3 * c + 1], g->ijk[3 * c + 2]};
            if (idx[d] == 0) {
                g->e_in.push_back(c); g->e_out.push_back(-1); g->e_dir.push_back(d); g->e_side.push_back(2 * d);
            }
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
    g->E = (int)g->e_in.size();
    // cj = (CellSurfaceArea + bndArea) * 0.5 / volume, where CellSurfaceArea
    // already counts boundary edges once (F/GridData_CellData.cs:666-700).
    std::vector<double> surf(g->J, 0.0), bnd(g->J, 0.0);
    for (int e = 0; e < g->E; ++e) {
        int c = g->e_in[e], d = g->e_dir[e];
        double area = 1.0;
        for (int dd = 0; dd < D; ++dd) if (dd != d) area *= g->h[D * c + dd];
        surf[c] += area;
        if (g->e_out[e] >= 0) surf[g->e_out[e]] += area; else bnd[c] += area;
    }
    g->cj.resize(g->J);
    for (int c = 0; c < g->J; ++c) {
        double vol = 1.0;
        for (int d = 0; d < D; ++d) vol *= g->h[D * c + d];
        g->cj[c] = (surf[c] + bnd[c]) * 0.5 / vol;
    }
    return g;
}

// ===================================================================== SIP ==
// Tensor Gauss rule on [-1,1]^m (F/Square.cs:110-128 builds the 2D rule as a
// tensor product; node index i*NN+j with x <- node j, y <- node i).
static void tensor_rule(int m, int n1, std::vector<double>& X, std::vector<double>& W) {
    std::vector<double> x(n1), w(n1);
    gauss_rule(n1, x.data(), w.data());
    int nq = 1;
    for (int i = 0; i < m; ++i) nq *= n1;
    X.assign((size_t)nq * std::max(m, 1), 0.0);
    W.assign(nq, 1.0);
    for (int q = 0; q < nq; ++q) {
        int r = q;
        for (int d = 0; d < m; ++d) {
            int id = r % n1; r /= n1;
            X[q * m + d] = x[id];
            W[q] *= w[id];
        }
    }
}

struct NodeData {  // physical basis values at one node of one cell
    double phi[64];
    double gphi[64 * 3];
};

// Reference-element basis values at the volume and face nodes of the
// operator's rules, cached once per operator as BasisData caches them per
// quadrature rule (F/BasisData.cs).  Scaled to the physical cell on use:
// phi = psi / sqrt(detJ) (F/BasisData.cs:43-50), gradients through the
// inverse (diagonal) Jacobian 2/h_d.
static void face_node_xi(int D, int d, int side, const double* fx, double* xi);
static void tensor_rule(int m, int n1, std::vector<double>& X, std::vector<double>& W);

static void build_cache(Sip& s) {
    int D = s.D, Np = s.Np;
    std::vector<double> VX, VW, FX, FW;
    tensor_rule(D, s.nq1, VX, VW);
    tensor_rule(D - 1, s.nq1, FX, FW);
    s.vpsi.assign((size_t)s.nqv * Np, 0.0);
    s.vgpsi.assign((size_t)s.nqv * Np * D, 0.0);
    for (int q = 0; q < s.nqv; ++q) basis_eval(D, s.k, &VX[q * D], &s.vpsi[(size_t)q * Np], &s.vgpsi[(size_t)q * Np * D]);
    s.fpsi.assign((size_t)2 * D * s.nqf * Np, 0.0);
    s.fgpsi.assign((size_t)2 * D * s.nqf * Np * D, 0.0);
    for (int f = 0; f < 2 * D; ++f)
        for (int q = 0; q < s.nqf; ++q) {
            double xi[3];
            face_node_xi(D, f / 2, f % 2, &FX[q * (D - 1 > 0 ? D - 1 : 1)], xi);
            size_t o = (size_t)f * s.nqf + q;
            basis_eval(D, s.k, xi, &s.fpsi[o * Np], &s.fgpsi[o * Np * D]);
        }
}

static void phys_scale(const Sip& s, int cell, const double* v, const double* gv, NodeData& nd) {
    const Grid& g = *s.g;
    int D = s.D;
    double detJ = 1.0;
    for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * cell + d];
    double sc = 1.0 / std::sqrt(detJ);
    for (int n = 0; n < s.Np; ++n) {
        nd.phi[n] = v[n] * sc;
        for (int d = 0; d < D; ++d) nd.gphi[n * D + d] = gv[n * D + d] * sc * 2.0 / g.h[D * cell + d];
    }
}
static void phys_vol(const Sip& s, int cell, int q, NodeData& nd) {
    phys_scale(s, cell, &s.vpsi[(size_t)q * s.Np], &s.vgpsi[(size_t)q * s.Np * s.D], nd);
}
static void phys_face(const Sip& s, int cell, int f, int q, NodeData& nd) {
    size_t o = (size_t)f * s.nqf + q;
    phys_scale(s, cell, &s.fpsi[o * s.Np], &s.fgpsi[o * s.Np * s.D], nd);
}

// --- the SIP forms, restated from SIPLaplace
// (L3-solution/BoSSS.Solution.NSECommon/ipLaplace.cs:100-250) -------------
static double penalty(const Sip& s, int jIn, int jOut) {
    // GetPenalty: mu = base*deg*cj_in, max with cj_out (ipLaplace.cs:131-144)
    double mu = s.penalty_base * s.penalty_deg * s.g->cj[jIn];
    if (jOut >= 0) mu = std::max(mu, s.penalty_base * s.penalty_deg * s.g->cj[jOut]);
    return mu;
}

static double volume_form(const Sip& s, int cell, const double* gradU, double V, const double* gradV) {
    (void)V;
    double acc = 0;
    for (int d = 0; d < s.D; ++d) acc -= gradU[d] * gradV[d] * s.nu[cell];
    return acc;
}

static double inner_edge_form(const Sip& s, int jIn, int jOut, const double* n, double uA, double uB,
                              const double* gA, const double* gB, double vA, double vB,
                              const double* gvA, const double* gvB) {
    double pn = penalty(s, jIn, jOut);
    double nuA = s.nu[jIn], nuB = s.nu[jOut];
    double acc = 0;
    for (int d = 0; d < s.D; ++d) {
        acc += 0.5 * (nuA * gA[d] + nuB * gB[d]) * (vA - vB) * n[d];
        acc += 0.5 * (nuA * gvA[d] + nuB * gvB[d]) * (uA - uB) * n[d];
    }
    double nuMax = (std::fabs(nuA) > std::fabs(nuB)) ? nuA : nuB;
    acc -= (uA - uB) * (vA - vB) * pn * nuMax;
    return acc;
}

static double boundary_edge_form(const Sip& s, int jIn, int side, const double* n, double gval,
                                 double uA, const double* gA, double vA, const double* gvA) {
    double acc = 0;
    double pn = 2.0 * penalty(s, jIn, -1);
    double nuA = s.nu[jIn];
    if (s.bc[side] == 0) {
        for (int d = 0; d < s.D; ++d) {
            acc += nuA * gA[d] * vA * n[d];
            acc += nuA * gvA[d] * (uA - gval) * n[d];
        }
        acc -= nuA * (uA - gval) * vA * pn;
    } else {
        acc += nuA * gval * vA;
    }
    return acc;
}

// Face nodes of a cell: face normal direction d, side s (0: xi_d=-1, 1: +1);
// the remaining coordinates run over the (D-1)-dim tensor rule in ascending
// dimension order.
static void face_node_xi(int D, int d, int side, const double* fx, double* xi) {
    int m = 0;
    for (int dd = 0; dd < D; ++dd) {
        if (dd == d) xi[dd] = side ? 1.0 : -1.0;
        else xi[dd] = fx[m++];
    }
}

Sip* sip_create(const Grid* g, int k, double penalty_base, const double* nu, const int* bc, int nq1) {
    Sip* s = new Sip();
    s->g = g; s->D = g->D; s->k = k; s->Np = basis_count(g->D, k);
    s->nq1 = nq1 > 0 ? nq1 : k + 1;
    s->penalty_base = penalty_base;
    double D = g->D, p = k;
    // CoefficientUpdate (ipLaplace.cs:100-110)
    s->penalty_deg = std::max((p + 1) * (p + D) / D, (p + 1.0) * (p + 1.0));
    s->nu.assign(nu, nu + g->J);
    for (int i = 0; i < 2 * g->D; ++i) s->bc[i] = bc ? bc[i] : 0;
    s->nqv = 1; s->nqf = 1;
    for (int d = 0; d < g->D; ++d) s->nqv *= s->nq1;
    for (int d = 0; d < g->D - 1; ++d) s->nqf *= s->nq1;
    s->bnd_edge_slot.assign(g->E, -1);
    int slot = 0;
    for (int e = 0; e < g->E; ++e)
        if (g->e_out[e] < 0) { s->bnd_edge_slot[e] = slot; slot += s->nqf; }
    s->gbnd.assign(slot, 0.0);
    build_cache(*s);
    return s;
}

// field value and gradient at a node (F/DGField_Evaluation.cs:61-125, 258-311)
static void eval_field(const Sip& s, const double* u, int cell, const NodeData& nd, double& val, double* grad) {
    const double* c = u + (size_t)cell * s.Np;
    val = 0;
    for (int d = 0; d < s.D; ++d) grad[d] = 0;
    for (int n = 0; n < s.Np; ++n) {
        val += c[n] * nd.phi[n];
        for (int d = 0; d < s.D; ++d) grad[d] += c[n] * nd.gphi[n * s.D + d];
    }
}

// Matrix-free evaluate: y <- beta*y + alpha*Op(u), Op including the affine
// (boundary-value and RHS-source) parts.  Volume part follows
// NECQuadratureVolume.EvaluateEx (F/NECQuadratureVolume.cs:377-870) with the
// flux vectorised per node (F/NonlinFormVectorizer.cs:153-262); edge part
// follows NECQuadratureEdge (F/NECQuadratureEdge.cs:561-1400, scatter to both
// cells :1790-1850).  Chunks of <= s.chunk items (F/Quadrature.cs:43,738-745).
void sip_evaluate(const Sip& s, double alpha, const double* u, double beta, double* y) {
    const Grid& g = *s.g;
    int D = s.D, Np = s.Np;
    size_t N = (size_t)g.J * Np;
    for (size_t i = 0; i < N; ++i) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
    std::vector<double> VX, VW, FX, FW;
    tensor_rule(D, s.nq1, VX, VW);
    tensor_rule(D - 1, s.nq1, FX, FW);
    int nchunk = (g.J + s.chunk - 1) / s.chunk;
    // ---- volume
#pragma omp parallel for schedule(dynamic)
    for (int ch = 0; ch < nchunk; ++ch) {
        int j0 = ch * s.chunk, j1 = std::min(g.J, j0 + s.chunk);
        NodeData nd;
        double res[64];
        for (int j = j0; j < j1; ++j) {
            double detJ = 1.0;
            for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * j + d];
            std::fill(res, res + Np, 0.0);
            for (int q = 0; q < s.nqv; ++q) {
                phys_vol(s, j, q, nd);
                double uval, gu[3];
                eval_field(s, u, j, nd, uval, gu);
                // vectorizer: call the form with V = 0, GradV = e_d
                double fgv[3], zero3[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) {
                    double ed[3] = {0, 0, 0};
                    ed[d] = 1.0;
                    fgv[d] = volume_form(s, j, gu, 0.0, ed);
                }
                (void)zero3;
                // RHS source: -rhs * V (SipPoissonMain.cs:405-436)
                double fv = 0;
                if (!s.f_rhs.empty()) {
                    double fval, fg[3];
                    eval_field(s, s.f_rhs.data(), j, nd, fval, fg);
                    fv = -fval;
                }
                double wq = VW[q] * detJ;
                for (int m = 0; m < Np; ++m) {
                    double a = fv * nd.phi[m];
                    for (int d = 0; d < D; ++d) a += fgv[d] * nd.gphi[m * D + d];
                    res[m] += wq * a;
                }
            }
            for (int m = 0; m < Np; ++m) y[(size_t)j * Np + m] += alpha * res[m];
        }
    }
    // ---- edges: results per edge, then scatter in edge order
    std::vector<double> eres((size_t)g.E * 2 * Np, 0.0);
    int nechunk = (g.E + s.chunk - 1) / s.chunk;
#pragma omp parallel for schedule(dynamic)
    for (int ch = 0; ch < nechunk; ++ch) {
        int e0 = ch * s.chunk, e1 = std::min(g.E, e0 + s.chunk);
        NodeData ndA, ndB;
        for (int e = e0; e < e1; ++e) {
            int jA = g.e_in[e], jB = g.e_out[e], d = g.e_dir[e];
            double area = 1.0;
            for (int dd = 0; dd < D; ++dd) if (dd != d) area *= 0.5 * g.h[D * jA + dd];
            double* rA = &eres[(size_t)e * 2 * Np];
            double* rB = rA + Np;
            double nrm[3] = {0, 0, 0};
            int sideA;
            if (jB >= 0) { nrm[d] = 1.0; sideA = 1; }
            else { sideA = g.e_side[e] & 1; nrm[d] = sideA ? 1.0 : -1.0; }
            for (int q = 0; q < s.nqf; ++q) {
                phys_face(s, jA, 2 * d + sideA, q, ndA);
                double uA, gA[3];
                eval_field(s, u, jA, ndA, uA, gA);
                double wq = FW[q] * area;
                double z3[3] = {0, 0, 0};
                if (jB >= 0) {
                    phys_face(s, jB, 2 * d, q, ndB);
                    double uB, gB[3];
                    eval_field(s, u, jB, ndB, uB, gB);
                    double fvA = inner_edge_form(s, jA, jB, nrm, uA, uB, gA, gB, 1, 0, z3, z3);
                    double fvB = inner_edge_form(s, jA, jB, nrm, uA, uB, gA, gB, 0, 1, z3, z3);
                    double fgA[3], fgB[3];
                    for (int dd = 0; dd < D; ++dd) {
                        double ed[3] = {0, 0, 0};
                        ed[dd] = 1;
                        fgA[dd] = inner_edge_form(s, jA, jB, nrm, uA, uB, gA, gB, 0, 0, ed, z3);
                        fgB[dd] = inner_edge_form(s, jA, jB, nrm, uA, uB, gA, gB, 0, 0, z3, ed);
                    }
                    for (int m = 0; m < Np; ++m) {
                        double a = fvA * ndA.phi[m], b = fvB * ndB.phi[m];
                        for (int dd = 0; dd < D; ++dd) {
                            a += fgA[dd] * ndA.gphi[m * D + dd];
                            b += fgB[dd] * ndB.gphi[m * D + dd];
                        }
                        rA[m] += wq * a;
                        rB[m] += wq * b;
                    }
                } else {
                    int side = g.e_side[e];
                    double gv = s.gbnd[s.bnd_edge_slot[e] + q];
                    double fvA = boundary_edge_form(s, jA, side, nrm, gv, uA, gA, 1, z3);
                    double fgA[3];
                    for (int dd = 0; dd < D; ++dd) {
                        double ed[3] = {0, 0, 0};
                        ed[dd] = 1;
                        fgA[dd] = boundary_edge_form(s, jA, side, nrm, gv, uA, gA, 0, ed);
                    }
                    for (int m = 0; m < Np; ++m) {
                        double a = fvA * ndA.phi[m];
                        for (int dd = 0; dd < D; ++dd) a += fgA[dd] * ndA.gphi[m * D + dd];
                        rA[m] += wq * a;
                    }
                }
            }
        }
    }
    for (int e = 0; e < g.E; ++e) {
        int jA = g.e_in[e], jB = g.e_out[e];
        for (int m = 0; m < Np; ++m) y[(size_t)jA * Np + m] += alpha * eres[(size_t)e * 2 * Np + m];
        if (jB >= 0)
            for (int m = 0; m < Np; ++m) y[(size_t)jB * Np + m] += alpha * eres[(size_t)e * 2 * Np + Np + m];
    }
}

// Matrix assembly: blocks of the bilinear forms plus the affine vector, as
// LECVolumeQuadrature2 / LECEdgeQuadrature2 (F/LECVolumeQuadrature2.cs:643-1140,
// F/LECEdgeQuadrature2.cs:1370-1560) accumulate them with AccBlock
// (P/BlockMsrMatrix.cs:1631).  The linear part is the form with trial
// function phi_n minus the form at u = 0; the affine part is the form at u = 0.
Bsr* sip_assemble(const Sip& s, double* affine) {
    const Grid& g = *s.g;
    int D = s.D, Np = s.Np, J = g.J;
    Bsr* M = new Bsr();
    M->nrows = J; M->bs = Np;
    // sparsity: each cell couples to itself and its face neighbours, columns ascending
    std::vector<std::vector<int>> cols(J);
    for (int j = 0; j < J; ++j) cols[j].push_back(j);
    for (int e = 0; e < g.E; ++e)
        if (g.e_out[e] >= 0) { cols[g.e_in[e]].push_back(g.e_out[e]); cols[g.e_out[e]].push_back(g.e_in[e]); }
    M->rowptr.assign(J + 1, 0);
    for (int j = 0; j < J; ++j) {
        std::sort(cols[j].begin(), cols[j].end());
        M->rowptr[j + 1] = M->rowptr[j] + (int64_t)cols[j].size();
    }
    M->col.resize(M->rowptr[J]);
    for (int j = 0; j < J; ++j) std::copy(cols[j].begin(), cols[j].end(), M->col.begin() + M->rowptr[j]);
    M->val.assign((size_t)M->rowptr[J] * Np * Np, 0.0);
    auto blk = [&](int r, int c) -> double* {
        auto b = std::lower_bound(M->col.begin() + M->rowptr[r], M->col.begin() + M->rowptr[r + 1], c);
        return &M->val[(size_t)(b - M->col.begin()) * Np * Np];
    };
    size_t N = (size_t)J * Np;
    std::vector<double> aff(N, 0.0);
    std::vector<double> VX, VW, FX, FW;
    tensor_rule(D, s.nq1, VX, VW);
    tensor_rule(D - 1, s.nq1, FX, FW);
    NodeData nd, ndA, ndB;
    double z3[3] = {0, 0, 0};
    for (int j = 0; j < J; ++j) {
        double detJ = 1.0;
        for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * j + d];
        double* B = blk(j, j);
        for (int q = 0; q < s.nqv; ++q) {
            phys_vol(s, j, q, nd);
            double wq = VW[q] * detJ;
            for (int m = 0; m < Np; ++m)
                for (int n = 0; n < Np; ++n)
                    B[m * Np + n] += wq * volume_form(s, j, &nd.gphi[n * D], nd.phi[m], &nd.gphi[m * D]);
            if (!s.f_rhs.empty()) {
                double fval, fg[3];
                eval_field(s, s.f_rhs.data(), j, nd, fval, fg);
                for (int m = 0; m < Np; ++m) aff[(size_t)j * Np + m] += wq * (-fval) * nd.phi[m];
            }
        }
    }
    for (int e = 0; e < g.E; ++e) {
        int jA = g.e_in[e], jB = g.e_out[e], d = g.e_dir[e];
        double area = 1.0;
        for (int dd = 0; dd < D; ++dd) if (dd != d) area *= 0.5 * g.h[D * jA + dd];
        double nrm[3] = {0, 0, 0};
        int sideA;
        if (jB >= 0) { nrm[d] = 1.0; sideA = 1; }
        else { sideA = g.e_side[e] & 1; nrm[d] = sideA ? 1.0 : -1.0; }
        for (int q = 0; q < s.nqf; ++q) {
            phys_face(s, jA, 2 * d + sideA, q, ndA);
            double wq = FW[q] * area;
            if (jB >= 0) {
                phys_face(s, jB, 2 * d, q, ndB);
                double* AA = blk(jA, jA); double* AB = blk(jA, jB);
                double* BA = blk(jB, jA); double* BB = blk(jB, jB);
                for (int m = 0; m < Np; ++m)
                    for (int n = 0; n < Np; ++n) {
                        // trial on A
                        double uA = ndA.phi[n];
                        const double* gA = &ndA.gphi[n * D];
                        AA[m * Np + n] += wq * inner_edge_form(s, jA, jB, nrm, uA, 0, gA, z3, ndA.phi[m], 0, &ndA.gphi[m * D], z3);
                        BA[m * Np + n] += wq * inner_edge_form(s, jA, jB, nrm, uA, 0, gA, z3, 0, ndB.phi[m], z3, &ndB.gphi[m * D]);
                        double uB = ndB.phi[n];
                        const double* gB = &ndB.gphi[n * D];
                        AB[m * Np + n] += wq * inner_edge_form(s, jA, jB, nrm, 0, uB, z3, gB, ndA.phi[m], 0, &ndA.gphi[m * D], z3);
                        BB[m * Np + n] += wq * inner_edge_form(s, jA, jB, nrm, 0, uB, z3, gB, 0, ndB.phi[m], z3, &ndB.gphi[m * D]);
                    }
            } else {
                int side = g.e_side[e];
                double gv = s.gbnd[s.bnd_edge_slot[e] + q];
                double* AA = blk(jA, jA);
                for (int m = 0; m < Np; ++m) {
                    double f0 = boundary_edge_form(s, jA, side, nrm, gv, 0, z3, ndA.phi[m], &ndA.gphi[m * D]);
                    aff[(size_t)jA * Np + m] += wq * f0;
                    for (int n = 0; n < Np; ++n) {
                        double f1 = boundary_edge_form(s, jA, side, nrm, gv, ndA.phi[n], &ndA.gphi[n * D], ndA.phi[m], &ndA.gphi[m * D]);
                        AA[m * Np + n] += wq * (f1 - f0);
                    }
                }
            }
        }
    }
    if (affine) std::copy(aff.begin(), aff.end(), affine);
    return M;
}

// ===================================================================== BSR ==
// y <- beta*y + alpha*A*x over block rows (P/BlockMsrMatrix.cs:2970-3260;
// ParallelFor over block rows :3239, GEMV inner loops :3162-3194).
void bsr_spmv(const Bsr& A, double alpha, const double* x, double beta, double* y) {
    int bs = A.bs;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < A.nrows; ++r) {
        double acc[64];
        for (int i = 0; i < bs; ++i) acc[i] = 0;
        for (int64_t p = A.rowptr[r]; p < A.rowptr[r + 1]; ++p) {
            const double* B = &A.val[(size_t)p * bs * bs];
            const double* xc = x + (size_t)A.col[p] * bs;
            for (int i = 0; i < bs; ++i) {
                double s = 0;
                for (int j = 0; j < bs; ++j) s += B[i * bs + j] * xc[j];
                acc[i] += s;
            }
        }
        double* yr = y + (size_t)r * bs;
        for (int i = 0; i < bs; ++i) yr[i] = (beta == 0.0 ? 0.0 : beta * yr[i]) + alpha * acc[i];
    }
}

// In-place inverse with partial pivoting (IL/IMatrixExtensions.cs:2221-2335,
// DGETRF/DGETRI semantics).
static bool invert_block(double* A, int n) {
    std::vector<double> M(A, A + n * n), I(n * n, 0.0);
    for (int i = 0; i < n; ++i) I[i * n + i] = 1;
    for (int c = 0; c < n; ++c) {
        int p = c;
        for (int r = c + 1; r < n; ++r) if (std::fabs(M[r * n + c]) > std::fabs(M[p * n + c])) p = r;
        if (M[p * n + c] == 0.0) return false;
        if (p != c)
            for (int k = 0; k < n; ++k) { std::swap(M[p * n + k], M[c * n + k]); std::swap(I[p * n + k], I[c * n + k]); }
        double piv = 1.0 / M[c * n + c];
        for (int k = 0; k < n; ++k) { M[c * n + k] *= piv; I[c * n + k] *= piv; }
        for (int r = 0; r < n; ++r) {
            if (r == c) continue;
            double f = M[r * n + c];
            if (f == 0.0) continue;
            for (int k = 0; k < n; ++k) { M[r * n + k] -= f * M[c * n + k]; I[r * n + k] -= f * I[c * n + k]; }
        }
    }
    std::copy(I.begin(), I.end(), A);
    return true;
}

// BlockJacobi (AS/BlockJacobi.cs:100-255): Init inverts the diagonal blocks;
// Solve runs sweeps while iter <= NoOfIterations:
//   q = b - A x;  q += D x;  x = omega * Dinv q + (1 - omega) x.
struct BlockJacobi {
    const Bsr* A = nullptr;
    std::vector<double> Dg, Dinv;
    double omega = 1.0;
    int NoOfIterations = 1;
    void init(const Bsr* M) {
        A = M;
        int bs = M->bs;
        Dg.assign((size_t)M->nrows * bs * bs, 0.0);
        for (int r = 0; r < M->nrows; ++r)
            for (int64_t p = M->rowptr[r]; p < M->rowptr[r + 1]; ++p)
                if (M->col[p] == r) std::copy(&M->val[(size_t)p * bs * bs], &M->val[(size_t)(p + 1) * bs * bs], &Dg[(size_t)r * bs * bs]);
        Dinv = Dg;
        for (int r = 0; r < M->nrows; ++r)
            if (!invert_block(&Dinv[(size_t)r * bs * bs], bs)) throw std::runtime_error("singular diagonal block");
    }
    void solve(double* x, const double* b) const {
        int bs = A->bs;
        size_t N = (size_t)A->nrows * bs;
        std::vector<double> q(N);
        for (int it = 0; true; ++it) {
            std::copy(b, b + N, q.begin());
            bsr_spmv(*A, -1.0, x, 1.0, q.data());
            if (!(it <= NoOfIterations)) return;
#pragma omp parallel for schedule(static)
            for (int r = 0; r < A->nrows; ++r) {
                const double* Dr = &Dg[(size_t)r * bs * bs];
                const double* Di = &Dinv[(size_t)r * bs * bs];
                double* xr = x + (size_t)r * bs;
                double* qr = &q[(size_t)r * bs];
                double t[64];
                for (int i = 0; i < bs; ++i) {
                    double s = 0;
                    for (int j = 0; j < bs; ++j) s += Dr[i * bs + j] * xr[j];
                    t[i] = qr[i] + s;
                }
                for (int i = 0; i < bs; ++i) {
                    double s = 0;
                    for (int j = 0; j < bs; ++j) s += Di[i * bs + j] * t[j];
                    xr[i] = omega * s + (1.0 - omega) * xr[i];
                }
            }
        }
    }
};

static double dot(const double* a, const double* b, size_t n) {
    double s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (size_t i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

// Termination of IterativeSolverConfig.DefaultTermination
// (AS/IterativeSolverConfig.cs:48-58).
struct Termination {
    double tol = 1e-10;
    int MinIter = 0, MaxIter = 10000;
    double MachineEpsilon = 2.2204460492503131e-16 * 1000;
    bool keep_running(int iter, double R0, double R, bool& success) const {
        success = false;
        if (iter <= MinIter) return true;
        if (R < R0 * tol + tol) { success = true; return false; }
        if (iter > MaxIter) return false;
        if (R0 < MachineEpsilon || R < MachineEpsilon) { success = true; return false; }
        return true;
    }
};

// Givens parameters (AS/SoftGMRES.cs:208-222)
static void rotmat(double& c, double& s, double a, double b) {
    if (b == 0.0) { c = 1.0; s = 0.0; }
    else if (std::fabs(b) > std::fabs(a)) { double t = a / b; s = 1.0 / std::sqrt(1.0 + t * t); c = t * s; }
    else { double t = b / a; c = 1.0 / std::sqrt(1.0 + t * t); s = t * c; }
}

// Back substitution of the (already rotated) upper-triangular H(1:i,1:i).
static void solve_upper(const std::vector<double>& H, int ld, int i, const double* s, double* y) {
    for (int r = i - 1; r >= 0; --r) {
        double a = s[r];
        for (int c = r + 1; c < i; ++c) a -= H[r * ld + c] * y[c];
        y[r] = a / H[r * ld + r];
    }
}

using Precond = std::function<void(double* x, const double* b)>;  // x <- M^-1 b (x zeroed by caller)

// SoftGMRES.Solve (AS/SoftGMRES.cs:227-455): left-preconditioned restarted
// GMRES(m) with modified Gram-Schmidt and Givens rotations; termination on
// the unpreconditioned residual norm at restarts / the preconditioned
// estimate inside the cycle, exactly as the reference passes error2.
struct GmresStats { int iters = 0; bool converged = false; double res0 = 0, res = 0; };

GmresStats soft_gmres(const Bsr& A, const Precond& P, double* X, const double* B, int m, const Termination& T) {
    size_t N = (size_t)A.nrows * A.bs;
    GmresStats st;
    std::vector<double> r(N), z(N), w(N);
    auto applyP = [&](const double* zin, double* out) {
        if (P) { std::fill(out, out + N, 0.0); P(out, zin); }
        else std::copy(zin, zin + N, out);
    };
    std::copy(B, B + N, z.begin());
    bsr_spmv(A, -1.0, X, 1.0, z.data());
    applyP(z.data(), r.data());
    double error2 = std::sqrt(dot(z.data(), z.data(), N));
    double iter0 = error2;
    st.res0 = iter0;
    bool succ;
    if (!T.keep_running(0, error2, error2, succ)) { st.converged = succ; st.res = error2; return st; }
    std::vector<std::vector<double>> V(m + 1, std::vector<double>(N));
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), s(m + 2), y;
    int tot = 0;
    while (true) {
        std::copy(B, B + N, z.begin());
        bsr_spmv(A, -1.0, X, 1.0, z.data());
        error2 = std::sqrt(dot(z.data(), z.data(), N));
        applyP(z.data(), r.data());
        double nr = std::sqrt(dot(r.data(), r.data(), N));
        for (size_t i = 0; i < N; ++i) V[0][i] = r[i] / nr;
        std::fill(s.begin(), s.end(), 0.0);
        s[0] = nr;
        std::fill(H.begin(), H.end(), 0.0);
        int i;
        bool done = false;
        for (i = 1; i <= m; ++i) {
            st.iters++;
            bsr_spmv(A, 1.0, V[i - 1].data(), 0.0, z.data());
            applyP(z.data(), w.data());
            for (int k = 0; k < i; ++k) {
                double h = dot(w.data(), V[k].data(), N);
                H[k * m + (i - 1)] = h;
                for (size_t q = 0; q < N; ++q) w[q] -= h * V[k][q];
            }
            double nw = std::sqrt(dot(w.data(), w.data(), N));
            H[i * m + (i - 1)] = nw;
            for (size_t q = 0; q < N; ++q) V[i][q] = w[q] / nw;
            for (int k = 1; k <= i - 1; ++k) {
                double t = cs[k - 1] * H[(k - 1) * m + i - 1] + sn[k - 1] * H[k * m + i - 1];
                H[k * m + i - 1] = -sn[k - 1] * H[(k - 1) * m + i - 1] + cs[k - 1] * H[k * m + i - 1];
                H[(k - 1) * m + i - 1] = t;
            }
            rotmat(cs[i - 1], sn[i - 1], H[(i - 1) * m + i - 1], H[i * m + i - 1]);
            double t = cs[i - 1] * s[i - 1];
            H[(i - 1) * m + i - 1] = cs[i - 1] * H[(i - 1) * m + i - 1] + sn[i - 1] * H[i * m + i - 1];
            H[i * m + i - 1] = 0.0;
            s[i] = -sn[i - 1] * s[i - 1];
            s[i - 1] = t;
            error2 = std::fabs(s[i]);
            if (!T.keep_running(tot, iter0, error2, succ)) {
                y.assign(i, 0.0);
                solve_upper(H, m, i, s.data(), y.data());
                for (int ii = 0; ii < i; ++ii)
                    for (size_t q = 0; q < N; ++q) X[q] += y[ii] * V[ii][q];
                st.converged = succ;
                done = true;
                break;
            }
            tot++;
        }
        if (done) break;
        if (!T.keep_running(tot, iter0, error2, succ)) { st.converged = succ; break; }
        y.assign(m, 0.0);
        solve_upper(H, m, m, s.data(), y.data());
        for (int ii = 0; ii < m; ++ii)
            for (size_t q = 0; q < N; ++q) X[q] += y[ii] * V[ii][q];
    }
    std::copy(B, B + N, z.begin());
    bsr_spmv(A, -1.0, X, 1.0, z.data());
    st.res = std::sqrt(dot(z.data(), z.data(), N));
    return st;
}


// ============================================================ multigrid ==
// p/h multigrid restated from the reference's p-multigrid over the
// hierarchical orthonormal basis (AS/PmgConfig.cs:80-200, AS/LevelPmg.cs) and
// aggregation coarsening (AS/AggregationGridBasis.cs,
// F/CoarseningAlgorithms.cs:555-611) with Galerkin operators
// (AS/MultigridOperator.cs:573-640): lower degree = leading modes of each
// cell (R = truncation, P = R^T); the last p-level is aggregated 2^D cells at
// a time, the coarse space being the same-degree polynomials on the aggregate
// box, injected by the L2 projection P_i; A_c = P^T A P; block-Jacobi
// smoothing (AS/BlockJacobi.cs:196-255); dense solve on the coarsest level.
struct MgLev {
    Bsr A;
    std::vector<double> Dinv;
    int n[3] = {1, 1, 1};
    int deg = 0;
    std::vector<double> lo, h;  // cell boxes J x D
    int kind = 0;               // transfer to next: 0 truncation, 1 aggregation
    std::vector<int> parent;
    std::vector<double> P;      // J x bs x bs, [fine mode][coarse mode]
    // additive Schwarz over the aggregates (AS/Schwarz.cs with multigrid-aggregate
    // blocks): cells per block (the aggregate's children and, with overlap, their
    // neighbour layers; ascending cell index) and the inverse of A restricted to
    // their DOFs, padded to cmax cells with identity
    int sw_cmax = 0;
    std::vector<std::vector<int>> sw_children;
    std::vector<double> sw_inv;  // nagg x (cmax*bs)^2, row-major
    std::vector<int> sw_mult;    // with overlap scaling: blocks holding each cell
};

struct Pmg {
    int D = 3;
    std::vector<MgLev> lv;
    std::vector<double> Ainv;   // coarsest dense inverse (row-major)
    int nc = 0;
    double omega = 1.0;
    int pre = 1, post = 1;
    // LevelPmg on the finest level: per cell the inverse of the diagonal block's
    // high-order part (modes lp_lo.. of the cell), row-major
    bool level_pmg = false;
    int lp_lo = 0;
    std::vector<double> lp_inv;
};

// P[m][n] = int_f phi^f_m phi^c_n dx, evaluated with the fine cell's Gauss rule
static void agg_projection(int D, int deg, const double* lof, const double* hf, const double* loc, const double* hc, double* P) {
    int Np = basis_count(D, deg), n1 = deg + 1;
    std::vector<double> X, W;
    tensor_rule(D, n1, X, W);
    double dJf = 1, dJc = 1;
    for (int d = 0; d < D; ++d) { dJf *= 0.5 * hf[d]; dJc *= 0.5 * hc[d]; }
    std::vector<double> pf(Np), pc(Np);
    std::fill(P, P + Np * Np, 0.0);
    for (size_t q = 0; q < W.size(); ++q) {
        double xc[3];
        for (int d = 0; d < D; ++d) {
            double x = lof[d] + 0.5 * (X[q * D + d] + 1.0) * hf[d];
            xc[d] = 2.0 * (x - loc[d]) / hc[d] - 1.0;
        }
        basis_eval(D, deg, &X[q * D], pf.data(), nullptr);
        basis_eval(D, deg, xc, pc.data(), nullptr);
        for (int m = 0; m < Np; ++m)
            for (int n = 0; n < Np; ++n) P[m * Np + n] += W[q] * pf[m] * pc[n] * std::sqrt(dJf / dJc);
    }
}

Pmg* pmg_build(const Bsr& Af, const Grid& g, const int* deg, int ndeg, int min_coarse, double omega, int pre, int post) {
    Pmg* P = new Pmg();
    P->omega = omega; P->pre = pre; P->post = post;
    int D = g.D;
    P->D = D;
    for (int l = 0; l < ndeg; ++l) {
        MgLev L;
        int bs = basis_count(D, deg[l]);
        L.deg = deg[l];
        L.A.nrows = Af.nrows; L.A.bs = bs; L.A.rowptr = Af.rowptr; L.A.col = Af.col;
        L.A.val.resize((size_t)Af.rowptr[Af.nrows] * bs * bs);
        for (int64_t p = 0; p < Af.rowptr[Af.nrows]; ++p)
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j)
                    L.A.val[(size_t)p * bs * bs + i * bs + j] = Af.val[(size_t)p * Af.bs * Af.bs + i * Af.bs + j];
        for (int d = 0; d < D; ++d) L.n[d] = g.n[d];
        P->lv.push_back(std::move(L));
    }
    {
        MgLev& L0 = P->lv.back();
        L0.lo = g.lo;
        L0.h = g.h;
    }
    auto idx_of = [&](const int* n, int c, int* idx) {
        if (D == 2) { idx[0] = c / n[1]; idx[1] = c % n[1]; idx[2] = 0; }
        else { idx[0] = c % n[0]; idx[1] = (c / n[0]) % n[1]; idx[2] = c / (n[0] * n[1]); }
    };
    auto num = [&](const int* n, int i, int j, int k) { return D == 2 ? i * n[1] + j : i + j * n[0] + k * n[0] * n[1]; };
    while (P->lv.back().A.nrows > min_coarse) {
        MgLev& F = P->lv.back();
        int nc[3] = {1, 1, 1};
        bool can = false;
        for (int d = 0; d < D; ++d) { nc[d] = (F.n[d] + 1) / 2; can |= F.n[d] > 1; }
        if (!can) break;
        int Jf = F.A.nrows, Jc = nc[0] * nc[1] * (D == 3 ? nc[2] : 1), bs = F.A.bs;
        F.parent.resize(Jf);
        std::vector<double> loc((size_t)Jc * D, 1e300), hic((size_t)Jc * D, -1e300), hc((size_t)Jc * D);
        for (int c = 0; c < Jf; ++c) {
            int idx[3];
            idx_of(F.n, c, idx);
            int I = num(nc, idx[0] / 2, idx[1] / 2, idx[2] / 2);
            F.parent[c] = I;
            for (int d = 0; d < D; ++d) {
                loc[(size_t)I * D + d] = std::min(loc[(size_t)I * D + d], F.lo[(size_t)c * D + d]);
                hic[(size_t)I * D + d] = std::max(hic[(size_t)I * D + d], F.lo[(size_t)c * D + d] + F.h[(size_t)c * D + d]);
            }
        }
        for (size_t t = 0; t < hc.size(); ++t) hc[t] = hic[t] - loc[t];
        F.P.resize((size_t)Jf * bs * bs);
        for (int c = 0; c < Jf; ++c)
            agg_projection(D, F.deg, &F.lo[(size_t)c * D], &F.h[(size_t)c * D], &loc[(size_t)F.parent[c] * D],
                           &hc[(size_t)F.parent[c] * D], &F.P[(size_t)c * bs * bs]);
        F.kind = 1;
        std::vector<std::vector<std::pair<int, std::vector<double>>>> rows(Jc);
        std::vector<double> t1(bs * bs), t2(bs * bs);
        for (int i = 0; i < Jf; ++i) {
            const double* Pi = &F.P[(size_t)i * bs * bs];
            for (int64_t p = F.A.rowptr[i]; p < F.A.rowptr[i + 1]; ++p) {
                int j = F.A.col[p];
                const double* Aij = &F.A.val[(size_t)p * bs * bs];
                const double* Pj = &F.P[(size_t)j * bs * bs];
                for (int r = 0; r < bs; ++r)
                    for (int n = 0; n < bs; ++n) {
                        double s = 0;
                        for (int c = 0; c < bs; ++c) s += Aij[r * bs + c] * Pj[c * bs + n];
                        t1[r * bs + n] = s;
                    }
                for (int m = 0; m < bs; ++m)
                    for (int n = 0; n < bs; ++n) {
                        double s = 0;
                        for (int r = 0; r < bs; ++r) s += Pi[r * bs + m] * t1[r * bs + n];
                        t2[m * bs + n] = s;
                    }
                auto& row = rows[F.parent[i]];
                bool found = false;
                for (auto& e : row)
                    if (e.first == F.parent[j]) { for (int t = 0; t < bs * bs; ++t) e.second[t] += t2[t]; found = true; break; }
                if (!found) row.push_back({F.parent[j], t2});
            }
        }
        MgLev C;
        C.A.nrows = Jc; C.A.bs = bs;
        C.A.rowptr.assign(Jc + 1, 0);
        for (int I = 0; I < Jc; ++I) {
            std::sort(rows[I].begin(), rows[I].end(), [](const auto& a, const auto& b) { return a.first < b.first; });
            for (auto& e : rows[I]) { C.A.col.push_back(e.first); C.A.val.insert(C.A.val.end(), e.second.begin(), e.second.end()); }
            C.A.rowptr[I + 1] = (int64_t)C.A.col.size();
        }
        for (int d = 0; d < 3; ++d) C.n[d] = nc[d];
        C.deg = F.deg;
        C.lo = loc;
        C.h = hc;
        P->lv.push_back(std::move(C));
    }
    for (size_t l = 0; l + 1 < P->lv.size(); ++l) {
        BlockJacobi bj;
        bj.init(&P->lv[l].A);
        P->lv[l].Dinv = bj.Dinv;
    }
    MgLev& Cc = P->lv.back();
    int bs = Cc.A.bs;
    P->nc = Cc.A.nrows * bs;
    P->Ainv.assign((size_t)P->nc * P->nc, 0.0);
    for (int r = 0; r < Cc.A.nrows; ++r)
        for (int64_t p = Cc.A.rowptr[r]; p < Cc.A.rowptr[r + 1]; ++p)
            for (int i = 0; i < bs; ++i)
                for (int j = 0; j < bs; ++j)
                    P->Ainv[(size_t)(r * bs + i) * P->nc + Cc.A.col[p] * bs + j] = Cc.A.val[(size_t)p * bs * bs + i * bs + j];
    if (!invert_block(P->Ainv.data(), P->nc)) throw std::runtime_error("singular coarse matrix");
    return P;
}

// Schwarz blocks (AS/Schwarz.cs:683-730): the children of every aggregate, grown by
// `overlap` sweeps over the face neighbours of the cells present at the start of the
// sweep, then sorted by cell index; `scale`: the block sum of a DOF divided by the
// number of blocks that hold it (EnableOverlapScaling, :791-815).
void pmg_set_schwarz(Pmg& P, bool on, int overlap, bool scale) {
    if (overlap < 0 || overlap > 2) throw std::runtime_error("Schwarz overlap must be 0, 1 or 2");
    const int D = P.D;
    for (size_t l = 0; l + 1 < P.lv.size(); ++l) {
        MgLev& L = P.lv[l];
        L.sw_cmax = 0;
        L.sw_children.clear();
        L.sw_inv.clear();
        L.sw_mult.clear();
        if (!on || L.kind != 1) continue;
        const int bs = L.A.bs, J = L.A.nrows;
        const int Jc = P.lv[l + 1].A.nrows;
        std::vector<std::vector<int>> blocks(Jc);
        for (int c = 0; c < J; ++c) blocks[L.parent[c]].push_back(c);
        if (overlap > 0) {
            std::vector<char> in(J, 0);
            for (auto& blk : blocks) {
                for (int c : blk) in[c] = 1;
                for (int sweep = 0; sweep < overlap; ++sweep) {
                    const size_t n0 = blk.size();
                    for (size_t q = 0; q < n0; ++q) {
                        const int c = blk[q];
                        int ijk[3];
                        if (D == 2) { ijk[0] = c / L.n[1]; ijk[1] = c % L.n[1]; ijk[2] = 0; }
                        else { ijk[0] = c % L.n[0]; ijk[1] = (c / L.n[0]) % L.n[1]; ijk[2] = c / (L.n[0] * L.n[1]); }
                        for (int d = 0; d < D; ++d)
                            for (int sg : {-1, 1}) {
                                int nb[3] = {ijk[0], ijk[1], ijk[2]};
                                nb[d] += sg;
                                if (nb[d] < 0 || nb[d] >= L.n[d]) continue;
                                const int cn = D == 2 ? nb[0] * L.n[1] + nb[1] : nb[0] + nb[1] * L.n[0] + nb[2] * L.n[0] * L.n[1];
                                if (!in[cn]) { in[cn] = 1; blk.push_back(cn); }
                            }
                    }
                }
                std::sort(blk.begin(), blk.end());
                for (int c : blk) in[c] = 0;
            }
        }
        int cmax = overlap > 0 ? 0 : 1 << D;
        for (auto& blk : blocks) cmax = std::max(cmax, (int)blk.size());
        const int BA = cmax * bs;
        if (BA > (overlap > 0 ? 128 : 64)) continue;   // as the product: such levels keep block-Jacobi
        L.sw_children = std::move(blocks);
        if (scale) {
            L.sw_mult.assign(J, 0);
            for (auto& blk : L.sw_children)
                for (int c : blk) ++L.sw_mult[c];
        }
        L.sw_inv.assign((size_t)Jc * BA * BA, 0.0);
        for (int I = 0; I < Jc; ++I) {
            const std::vector<int>& ch = L.sw_children[I];
            double* B = &L.sw_inv[(size_t)I * BA * BA];
            for (int s = 0; s < cmax; ++s)
                for (int u = 0; u < cmax; ++u)
                    for (int i = 0; i < bs; ++i)
                        for (int j = 0; j < bs; ++j) {
                            double v = (s == u && i == j) ? 1.0 : 0.0;
                            if (s < (int)ch.size() && u < (int)ch.size()) {
                                v = 0.0;
                                for (int64_t p = L.A.rowptr[ch[s]]; p < L.A.rowptr[ch[s] + 1]; ++p)
                                    if (L.A.col[p] == ch[u]) v = L.A.val[(size_t)p * bs * bs + i * bs + j];
                            }
                            B[(size_t)(s * bs + i) * BA + u * bs + j] = v;
                        }
            if (!invert_block(B, BA)) throw std::runtime_error("singular Schwarz block");
        }
        L.sw_cmax = cmax;
    }
}

// x (+)= omega y with y = sum_I R_I^T A_I^-1 R_I r (additive: a DOF in several blocks
// sums their local solutions, AS/Schwarz.cs:1056-1085, no overlap scaling) or, without
// Schwarz blocks, y = D^-1 r
static void bgemv(const MgLev& L, const double* r, double* x, double omega, bool add) {
    int bs = L.A.bs;
    if (L.sw_cmax > 0) {
        const int BA = L.sw_cmax * bs;
        std::vector<double> rr(BA), y((size_t)L.A.nrows * bs, 0.0);
        for (size_t I = 0; I < L.sw_children.size(); ++I) {
            const std::vector<int>& ch = L.sw_children[I];
            std::fill(rr.begin(), rr.end(), 0.0);
            for (size_t s = 0; s < ch.size(); ++s)
                for (int i = 0; i < bs; ++i) rr[s * bs + i] = r[(size_t)ch[s] * bs + i];
            const double* B = &L.sw_inv[I * BA * BA];
            for (size_t s = 0; s < ch.size(); ++s)
                for (int i = 0; i < bs; ++i) {
                    double v = 0;
                    for (int c = 0; c < BA; ++c) v += B[(s * bs + i) * BA + c] * rr[c];
                    y[(size_t)ch[s] * bs + i] += v;
                }
        }
        if (!L.sw_mult.empty())
            for (size_t t = 0; t < y.size(); ++t) y[t] = y[t] / L.sw_mult[t / bs];
        for (size_t t = 0; t < y.size(); ++t) x[t] = (add ? x[t] : 0.0) + omega * y[t];
        return;
    }
    for (int row = 0; row < L.A.nrows; ++row) {
        const double* Di = &L.Dinv[(size_t)row * bs * bs];
        for (int i = 0; i < bs; ++i) {
            double s = 0;
            for (int j = 0; j < bs; ++j) s += Di[i * bs + j] * r[(size_t)row * bs + j];
            x[(size_t)row * bs + i] = (add ? x[(size_t)row * bs + i] : 0.0) + omega * s;
        }
    }
}

// LevelPmg (AS/LevelPmg.cs:232-330 set-up with UseDiagonalPmg, :397-470 CoarseSolve and
// FineSolve, :561-600 Solve): the low-order modes (degree <= the next level's) form the
// coarse system; the high-order modes of each cell are solved with the inverse of
// their part of the cell's diagonal block.
int pmg_set_level_pmg(Pmg& P, bool on) {
    P.level_pmg = false;
    P.lp_inv.clear();
    if (!on) return 0;
    if (P.lv.size() < 2 || P.lv[0].kind != 0) return -1;
    const Bsr& A = P.lv[0].A;
    const int bs = A.bs, lo = P.lv[1].A.bs, nh = bs - lo;
    if (nh <= 0) return -1;
    P.lp_lo = lo;
    P.lp_inv.assign((size_t)A.nrows * nh * nh, 0.0);
    for (int c = 0; c < A.nrows; ++c) {
        const double* Dc = nullptr;
        for (int64_t p = A.rowptr[c]; p < A.rowptr[c + 1]; ++p)
            if (A.col[p] == c) Dc = &A.val[(size_t)p * bs * bs];
        if (!Dc) return -1;
        double* H = &P.lp_inv[(size_t)c * nh * nh];
        for (int i = 0; i < nh; ++i)
            for (int j = 0; j < nh; ++j) H[i * nh + j] = Dc[(lo + i) * bs + lo + j];
        if (!invert_block(H, nh)) return -1;
    }
    P.level_pmg = true;
    return 0;
}

void pmg_vcycle(const Pmg& P, int l, const double* b, double* x) {
    const MgLev& L = P.lv[l];
    size_t N = (size_t)L.A.nrows * L.A.bs;
    if (l + 1 == (int)P.lv.size()) {
        for (int i = 0; i < P.nc; ++i) {
            double s = 0;
            for (int j = 0; j < P.nc; ++j) s += P.Ainv[(size_t)i * P.nc + j] * b[j];
            x[i] = s;
        }
        return;
    }
    std::vector<double> r(N);
    auto resid = [&]() { std::copy(b, b + N, r.begin()); bsr_spmv(L.A, -1.0, x, 1.0, r.data()); };
    if (l == 0 && P.level_pmg) {
        // x = coarse correction from zero (low-order modes), then the high-order modes
        const int bs = L.A.bs, lo = P.lp_lo, nh = bs - lo;
        const size_t Nc = (size_t)L.A.nrows * lo;
        std::vector<double> bc(Nc), xc(Nc, 0.0);
        for (int j = 0; j < L.A.nrows; ++j)
            for (int m = 0; m < lo; ++m) bc[(size_t)j * lo + m] = b[(size_t)j * bs + m];
        pmg_vcycle(P, 1, bc.data(), xc.data());
        std::fill(x, x + N, 0.0);
        for (int j = 0; j < L.A.nrows; ++j)
            for (int m = 0; m < lo; ++m) x[(size_t)j * bs + m] = xc[(size_t)j * lo + m];
        resid();
        for (int c = 0; c < L.A.nrows; ++c) {
            const double* H = &P.lp_inv[(size_t)c * nh * nh];
            for (int i = 0; i < nh; ++i) {
                double s = 0;
                for (int j = 0; j < nh; ++j) s += H[i * nh + j] * r[(size_t)c * bs + lo + j];
                x[(size_t)c * bs + lo + i] += s;
            }
        }
        return;
    }
    bgemv(L, b, x, P.omega, false);
    for (int s = 1; s < P.pre; ++s) { resid(); bgemv(L, r.data(), x, P.omega, true); }
    resid();
    const MgLev& C = P.lv[l + 1];
    size_t Nc = (size_t)C.A.nrows * C.A.bs;
    std::vector<double> bc(Nc, 0.0), xc(Nc, 0.0);
    int bs = L.A.bs;
    if (L.kind == 0) {
        for (int j = 0; j < L.A.nrows; ++j)
            for (int m = 0; m < C.A.bs; ++m) bc[(size_t)j * C.A.bs + m] = r[(size_t)j * L.A.bs + m];
    } else {
        for (int i = 0; i < L.A.nrows; ++i)
            for (int n = 0; n < bs; ++n)
                for (int m = 0; m < bs; ++m) bc[(size_t)L.parent[i] * bs + n] += L.P[(size_t)i * bs * bs + m * bs + n] * r[(size_t)i * bs + m];
    }
    pmg_vcycle(P, l + 1, bc.data(), xc.data());
    if (L.kind == 0) {
        for (int j = 0; j < L.A.nrows; ++j)
            for (int m = 0; m < C.A.bs; ++m) x[(size_t)j * L.A.bs + m] += xc[(size_t)j * C.A.bs + m];
    } else {
        for (int i = 0; i < L.A.nrows; ++i)
            for (int m = 0; m < bs; ++m)
                for (int n = 0; n < bs; ++n) x[(size_t)i * bs + m] += L.P[(size_t)i * bs * bs + m * bs + n] * xc[(size_t)L.parent[i] * bs + n];
    }
    for (int s = 0; s < P.post; ++s) { resid(); bgemv(L, r.data(), x, P.omega, true); }
}


// ================================================================ Euler ==
// Compressible Euler, explicit DG on a periodic Cartesian grid.  State layout
// [cell][variable][mode] with variables (rho, m_1..m_D, E) (the cell-major,
// field-major coordinate map of F/UnsetteledCoordinateMapping.cs:696-711).
// dU/dt = R(U) = int F(U) . grad(phi) - oint Fhat(U-, U+, n) phi, with the
// HLLC flux of OptimizedHLLC{Density,Momentum,Energy}Flux.InnerEdgeFlux
// (L3-solution/BoSSS.Solution.CompressibleFlowCommon/Convection/
// OptimizedHLLCDensityFlux.cs:45-167 and the Momentum/Energy variants) and the
// volume fluxes of the same classes' Flux(); non-dimensionalisation with
// MachScaling = gamma * Mach^2 as there.
struct Euler {
    const Grid* g = nullptr;
    int D = 3, k = 2, Np = 10, nq1 = 4, nqv = 64, nqf = 16, nv = 5;
    double gamma = 1.4, MS = 1.0;
    std::vector<double> VX, VW, FX, FW;
    std::vector<double> vpsi, vgpsi, fpsi;   // reference basis at volume nodes / face f nodes
};

static void hllc_flux(int D, double gamma, double MS, const double* Ui, const double* Uo, const double* n, double* F) {
    double rIn = Ui[0], rOut = Uo[0];
    double m2In = 0, m2Out = 0, vnIn = 0, vnOut = 0;
    for (int d = 0; d < D; ++d) {
        m2In += Ui[d + 1] * Ui[d + 1];
        m2Out += Uo[d + 1] * Uo[d + 1];
        vnIn += Ui[d + 1] * n[d];
        vnOut += Uo[d + 1] * n[d];
    }
    vnIn /= rIn;
    vnOut /= rOut;
    double EIn = Ui[D + 1], EOut = Uo[D + 1];
    double pIn = (gamma - 1.0) * (EIn - MS * 0.5 * m2In / rIn);
    double pOut = (gamma - 1.0) * (EOut - MS * 0.5 * m2Out / rOut);
    double Mach = std::sqrt(MS / gamma);
    double aIn = std::sqrt(pIn / rIn) / Mach, aOut = std::sqrt(pOut / rOut) / Mach;
    double rMean = 0.5 * (rIn + rOut), pMean = 0.5 * (pIn + pOut), aMean = 0.5 * (aIn + aOut);
    double jump = vnOut - vnIn;
    double pStar = std::max(0.0, pMean - 0.5 * MS * jump * aMean * rMean);
    double qIn = pStar > pIn ? std::sqrt(1.0 + 0.5 * (gamma + 1.0) * (pStar / pIn - 1.0) / gamma) : 1.0;
    double qOut = pStar > pOut ? std::sqrt(1.0 + 0.5 * (gamma + 1.0) * (pStar / pOut - 1.0) / gamma) : 1.0;
    double sIn = vnIn - aIn * qIn, sOut = vnOut + aOut * qOut;
    double cIn = rIn * (sIn - vnIn), cOut = rOut * (sOut - vnOut);
    double sStar = (cOut * vnOut - cIn * vnIn + (pIn - pOut) / MS) / (cOut - cIn);
    if (sStar > 0.0) {
        F[0] = vnIn * rIn;
        for (int c = 0; c < D; ++c) F[c + 1] = vnIn * Ui[c + 1] + pIn / MS * n[c];
        F[D + 1] = vnIn * (EIn + pIn);
        if (sIn <= 0.0) {
            double f = (sIn - vnIn) / (sIn - sStar);
            F[0] += sIn * (rIn * f - rIn);
            for (int c = 0; c < D; ++c)
                F[c + 1] += sIn * (rIn * f * (Ui[c + 1] / rIn + (sStar - vnIn) * n[c]) - Ui[c + 1]);
            double factor = rIn * sStar * MS + pIn / (sIn - vnIn);
            F[D + 1] += sIn * (f * (EIn + factor * (sStar - vnIn)) - EIn);
        }
    } else {
        F[0] = vnOut * rOut;
        for (int c = 0; c < D; ++c) F[c + 1] = vnOut * Uo[c + 1] + pOut / MS * n[c];
        F[D + 1] = vnOut * (EOut + pOut);
        if (sOut >= 0.0) {
            double f = (sOut - vnOut) / (sOut - sStar);
            F[0] += sOut * (rOut * f - rOut);
            for (int c = 0; c < D; ++c)
                F[c + 1] += sOut * (rOut * f * (Uo[c + 1] / rOut + (sStar - vnOut) * n[c]) - Uo[c + 1]);
            double factor = rOut * sStar * MS + pOut / (sOut - vnOut);
            F[D + 1] += sOut * (f * (EOut + factor * (sStar - vnOut)) - EOut);
        }
    }
}

// physical flux F[var][d] (OptimizedHLLC*Flux.Flux)
static void euler_flux(int D, double gamma, double MS, const double* U, double* F) {
    double r = U[0], E = U[D + 1], m2 = 0;
    for (int d = 0; d < D; ++d) m2 += U[d + 1] * U[d + 1];
    double p = (gamma - 1.0) * (E - MS * 0.5 * m2 / r);
    for (int d = 0; d < D; ++d) F[0 * D + d] = U[d + 1];
    for (int c = 0; c < D; ++c)
        for (int d = 0; d < D; ++d) F[(c + 1) * D + d] = U[c + 1] * U[d + 1] / r + (c == d ? p / MS : 0.0);
    for (int d = 0; d < D; ++d) F[(D + 1) * D + d] = U[d + 1] / r * (E + p);
}

Euler* euler_create(const Grid* g, int k, int nq1, double gamma, double mach) {
    Euler* e = new Euler();
    e->g = g; e->D = g->D; e->k = k; e->Np = basis_count(g->D, k); e->nq1 = nq1; e->nv = g->D + 2;
    e->gamma = gamma; e->MS = gamma * mach * mach;
    tensor_rule(e->D, nq1, e->VX, e->VW);
    tensor_rule(e->D - 1, nq1, e->FX, e->FW);
    e->nqv = (int)e->VW.size(); e->nqf = (int)e->FW.size();
    int D = e->D, Np = e->Np;
    e->vpsi.resize((size_t)e->nqv * Np); e->vgpsi.resize((size_t)e->nqv * Np * D);
    for (int q = 0; q < e->nqv; ++q) basis_eval(D, k, &e->VX[q * D], &e->vpsi[(size_t)q * Np], &e->vgpsi[(size_t)q * Np * D]);
    e->fpsi.resize((size_t)2 * D * e->nqf * Np);
    std::vector<double> gtmp((size_t)Np * D);
    for (int f = 0; f < 2 * D; ++f)
        for (int q = 0; q < e->nqf; ++q) {
            double xi[3];
            face_node_xi(D, f / 2, f % 2, &e->FX[q * (D - 1 > 0 ? D - 1 : 1)], xi);
            basis_eval(D, k, xi, &e->fpsi[((size_t)f * e->nqf + q) * Np], gtmp.data());
        }
    return e;
}

static int periodic_nb(const Grid& g, int c, int d, int s) {
    int idx[3] = {g.ijk[3 * c], g.ijk[3 * c + 1], g.ijk[3 * c + 2]};
    idx[d] = (idx[d] + (s ? 1 : -1) + g.n[d]) % g.n[d];
    return g.cell_index(idx[0], idx[1], idx[2]);
}

void euler_rhs(const Euler& e, const double* U, double* R) {
    const Grid& g = *e.g;
    int D = e.D, Np = e.Np, nv = e.nv, blk = nv * Np;
#pragma omp parallel for schedule(dynamic, 16)
    for (int j = 0; j < g.J; ++j) {
        double res[5 * 64];
        std::fill(res, res + blk, 0.0);
        double detJ = 1;
        for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * j + d];
        double sc = 1.0 / std::sqrt(detJ);
        const double* Uj = U + (size_t)j * blk;
        for (int q = 0; q < e.nqv; ++q) {
            const double* psi = &e.vpsi[(size_t)q * Np];
            const double* gpsi = &e.vgpsi[(size_t)q * Np * D];
            double Uq[5] = {0, 0, 0, 0, 0};
            for (int c = 0; c < nv; ++c)
                for (int n = 0; n < Np; ++n) Uq[c] += Uj[c * Np + n] * psi[n] * sc;
            double F[5 * 3];
            euler_flux(D, e.gamma, e.MS, Uq, F);
            double w = e.VW[q] * detJ;
            for (int c = 0; c < nv; ++c)
                for (int m = 0; m < Np; ++m) {
                    double a = 0;
                    for (int d = 0; d < D; ++d) a += F[c * D + d] * gpsi[m * D + d] * sc * 2.0 / g.h[D * j + d];
                    res[c * Np + m] += w * a;
                }
        }
        for (int f = 0; f < 2 * D; ++f) {
            int d = f / 2, s = f % 2;
            int nb = periodic_nb(g, j, d, s);
            double detN = 1;
            for (int dd = 0; dd < D; ++dd) detN *= 0.5 * g.h[D * nb + dd];
            double scN = 1.0 / std::sqrt(detN);
            double dA = detJ / (0.5 * g.h[D * j + d]);
            double nrm[3] = {0, 0, 0};
            nrm[d] = s ? 1.0 : -1.0;
            const double* Un = U + (size_t)nb * blk;
            int fo = 2 * d + (1 - s);
            for (int q = 0; q < e.nqf; ++q) {
                const double* pi = &e.fpsi[((size_t)f * e.nqf + q) * Np];
                const double* po = &e.fpsi[((size_t)fo * e.nqf + q) * Np];
                double Ui[5] = {0, 0, 0, 0, 0}, Uo[5] = {0, 0, 0, 0, 0};
                for (int c = 0; c < nv; ++c)
                    for (int n = 0; n < Np; ++n) { Ui[c] += Uj[c * Np + n] * pi[n] * sc; Uo[c] += Un[c * Np + n] * po[n] * scN; }
                double Fh[5];
                hllc_flux(D, e.gamma, e.MS, Ui, Uo, nrm, Fh);
                double w = e.FW[q] * dA;
                for (int c = 0; c < nv; ++c)
                    for (int m = 0; m < Np; ++m) res[c * Np + m] -= w * Fh[c] * pi[m] * sc;
            }
        }
        std::copy(res, res + blk, R + (size_t)j * blk);
    }
}

// Butcher form of RungeKuttaScheme.TVD3 (L3-solution/BoSSS.Solution/TimeStepping/
// RungeKuttaScheme.cs:145-167): a21 = 1, a31 = a32 = 1/4, b = (1/6, 1/6, 2/3).
void euler_tvd3(const Euler& e, double* U, double dt, int steps) {
    size_t n = (size_t)e.g->J * e.nv * e.Np;
    std::vector<double> k1(n), k2(n), k3(n), t(n);
    for (int s = 0; s < steps; ++s) {
        euler_rhs(e, U, k1.data());
        for (size_t i = 0; i < n; ++i) t[i] = U[i] + dt * k1[i];
        euler_rhs(e, t.data(), k2.data());
        for (size_t i = 0; i < n; ++i) t[i] = U[i] + dt * 0.25 * (k1[i] + k2[i]);
        euler_rhs(e, t.data(), k3.data());
        for (size_t i = 0; i < n; ++i) U[i] += dt * (k1[i] / 6.0 + k2[i] / 6.0 + 2.0 * k3[i] / 3.0);
    }
}

}  // namespace orc

// ================================================================== C API ==
using namespace orc;
extern "C" {

int orc_basis_count(int D, int k) { return basis_count(D, k); }
void orc_basis_exponents(int D, int k, int* out) {
    std::vector<int> ex;
    basis_exponents(D, k, ex);
    std::copy(ex.begin(), ex.end(), out);
}
void orc_basis_eval(int D, int k, int npts, const double* xi, double* val, double* grad) {
    int Np = basis_count(D, k);
    for (int p = 0; p < npts; ++p)
        basis_eval(D, k, xi + (size_t)p * D, val ? val + (size_t)p * Np : nullptr, grad ? grad + (size_t)p * Np * D : nullptr);
}
void orc_gauss(int n, double* x, double* w) { gauss_rule(n, x, w); }

void* orc_grid_create(int D, const double* x, int nx1, const double* y, int ny1, const double* z, int nz1) {
    return grid_create(D, x, nx1, y, ny1, z, nz1);
}
void orc_grid_free(void* g) { delete (Grid*)g; }
void orc_grid_info(void* gp, int* J, int* E) { Grid* g = (Grid*)gp; *J = g->J; *E = g->E; }
void orc_grid_edges(void* gp, int* in, int* out, int* dir, int* side) {
    Grid* g = (Grid*)gp;
    std::copy(g->e_in.begin(), g->e_in.end(), in);
    std::copy(g->e_out.begin(), g->e_out.end(), out);
    std::copy(g->e_dir.begin(), g->e_dir.end(), dir);
    std::copy(g->e_side.begin(), g->e_side.end(), side);
}
void orc_grid_cj(void* gp, double* cj) { Grid* g = (Grid*)gp; std::copy(g->cj.begin(), g->cj.end(), cj); }
void orc_grid_cells(void* gp, double* lo, double* h) {
    Grid* g = (Grid*)gp;
    std::copy(g->lo.begin(), g->lo.end(), lo);
    std::copy(g->h.begin(), g->h.end(), h);
}

void* orc_sip_create(void* g, int k, double penalty_base, const double* nu, const int* bc, int nq1) {
    return sip_create((Grid*)g, k, penalty_base, nu, bc, nq1);
}
void orc_sip_free(void* s) { delete (Sip*)s; }
void orc_sip_set_chunk(void* s, int chunk) { ((Sip*)s)->chunk = chunk; }
int orc_sip_num_bnd_nodes(void* sp) { return (int)((Sip*)sp)->gbnd.size(); }
// physical coordinates of the boundary-edge nodes, in the layout of gbnd
void orc_sip_bnd_nodes(void* sp, double* X) {
    Sip& s = *(Sip*)sp;
    const Grid& g = *s.g;
    int D = s.D;
    std::vector<double> FX, FW;
    tensor_rule(D - 1, s.nq1, FX, FW);
    for (int e = 0; e < g.E; ++e) {
        int slot = s.bnd_edge_slot[e];
        if (slot < 0) continue;
        int c = g.e_in[e], d = g.e_dir[e], side = g.e_side[e] & 1;
        for (int q = 0; q < s.nqf; ++q) {
            double xi[3];
            face_node_xi(D, d, side, &FX[q * (D - 1 > 0 ? D - 1 : 1)], xi);
            for (int dd = 0; dd < D; ++dd) X[(size_t)(slot + q) * D + dd] = g.lo[D * c + dd] + 0.5 * (xi[dd] + 1.0) * g.h[D * c + dd];
        }
    }
}
void orc_sip_set_bnd_values(void* sp, const double* gv) {
    Sip& s = *(Sip*)sp;
    std::copy(gv, gv + s.gbnd.size(), s.gbnd.begin());
}
void orc_sip_set_rhs(void* sp, const double* f) {
    Sip& s = *(Sip*)sp;
    if (!f) { s.f_rhs.clear(); return; }
    s.f_rhs.assign(f, f + (size_t)s.g->J * s.Np);
}
// volume quadrature nodes (physical) and weights*detJ: J*nqv
int orc_sip_num_vol_nodes(void* sp) { return ((Sip*)sp)->nqv; }
void orc_sip_vol_nodes(void* sp, double* X, double* W) {
    Sip& s = *(Sip*)sp;
    const Grid& g = *s.g;
    int D = s.D;
    std::vector<double> VX, VW;
    tensor_rule(D, s.nq1, VX, VW);
    for (int j = 0; j < g.J; ++j) {
        double detJ = 1;
        for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * j + d];
        for (int q = 0; q < s.nqv; ++q) {
            for (int d = 0; d < D; ++d)
                X[((size_t)j * s.nqv + q) * D + d] = g.lo[D * j + d] + 0.5 * (VX[q * D + d] + 1) * g.h[D * j + d];
            W[(size_t)j * s.nqv + q] = VW[q] * detJ;
        }
    }
}
// L2 projection onto the orthonormal basis from values at the volume nodes
void orc_sip_project(void* sp, const double* fvals, double* coef) {
    Sip& s = *(Sip*)sp;
    const Grid& g = *s.g;
    int D = s.D;
    std::vector<double> VX, VW;
    tensor_rule(D, s.nq1, VX, VW);
    NodeData nd;
    for (int j = 0; j < g.J; ++j) {
        double detJ = 1;
        for (int d = 0; d < D; ++d) detJ *= 0.5 * g.h[D * j + d];
        for (int m = 0; m < s.Np; ++m) coef[(size_t)j * s.Np + m] = 0;
        for (int q = 0; q < s.nqv; ++q) {
            phys_vol(s, j, q, nd);
            for (int m = 0; m < s.Np; ++m) coef[(size_t)j * s.Np + m] += VW[q] * detJ * fvals[(size_t)j * s.nqv + q] * nd.phi[m];
        }
    }
}
void orc_sip_evaluate(void* sp, double alpha, const double* u, double beta, double* y) {
    sip_evaluate(*(Sip*)sp, alpha, u, beta, y);
}
void* orc_sip_assemble(void* sp, double* affine) { return sip_assemble(*(Sip*)sp, affine); }

void* orc_bsr_create(int nrows, int bs, const int64_t* rowptr, const int* col, const double* val) {
    Bsr* M = new Bsr();
    M->nrows = nrows; M->bs = bs;
    M->rowptr.assign(rowptr, rowptr + nrows + 1);
    M->col.assign(col, col + rowptr[nrows]);
    M->val.assign(val, val + (size_t)rowptr[nrows] * bs * bs);
    return M;
}
void orc_bsr_free(void* m) { delete (Bsr*)m; }
void orc_bsr_info(void* mp, int* nrows, int64_t* nnzb, int* bs) {
    Bsr* M = (Bsr*)mp;
    *nrows = M->nrows; *nnzb = M->rowptr[M->nrows]; *bs = M->bs;
}
void orc_bsr_export(void* mp, int64_t* rowptr, int* col, double* val) {
    Bsr* M = (Bsr*)mp;
    std::copy(M->rowptr.begin(), M->rowptr.end(), rowptr);
    std::copy(M->col.begin(), M->col.end(), col);
    std::copy(M->val.begin(), M->val.end(), val);
}
void orc_bsr_spmv(void* mp, double alpha, const double* x, double beta, double* y) {
    bsr_spmv(*(Bsr*)mp, alpha, x, beta, y);
}
int orc_bjacobi_solve(void* mp, double omega, int iters, double* x, const double* b) {
    try {
        BlockJacobi bj;
        bj.omega = omega; bj.NoOfIterations = iters;
        bj.init((Bsr*)mp);
        bj.solve(x, b);
        return 0;
    } catch (...) { return -1; }
}
// GMRES(m) with optional block-Jacobi preconditioner (precond: 0 none, 1 BJ)
int orc_soft_gmres(void* mp, int precond, double omega, int bj_iters, int m, double tol, int maxit,
                   double* x, const double* b, int* iters, double* res0, double* res) {
    Bsr* A = (Bsr*)mp;
    BlockJacobi bj;
    Precond P;
    if (precond == 1) {
        bj.omega = omega; bj.NoOfIterations = bj_iters;
        bj.init(A);
        P = [&](double* xx, const double* bb) { bj.solve(xx, bb); };
    }
    Termination T;
    T.tol = tol; T.MaxIter = maxit;
    GmresStats st = soft_gmres(*A, P, x, b, m, T);
    *iters = st.iters; *res0 = st.res0; *res = st.res;
    return st.converged ? 0 : 1;
}
// persistent hierarchy: create / V-cycle from a level / level SpMV / transfers
void* orc_pmg_create(void* mp, void* gp, const int* deg, int ndeg, int min_coarse, double omega, int pre, int post) {
    try { return pmg_build(*(Bsr*)mp, *(Grid*)gp, deg, ndeg, min_coarse, omega, pre, post); } catch (...) { return nullptr; }
}
void orc_pmg_free(void* p) { delete (Pmg*)p; }
int orc_pmg_levels(void* p, int* sizes) {
    Pmg* P = (Pmg*)p;
    for (size_t l = 0; l < P->lv.size(); ++l) if (sizes) sizes[l] = P->lv[l].A.nrows * P->lv[l].A.bs;
    return (int)P->lv.size();
}
void orc_pmg_vcycle_from(void* p, int l, const double* b, double* x) {
    Pmg* P = (Pmg*)p;
    size_t N = (size_t)P->lv[l].A.nrows * P->lv[l].A.bs;
    std::fill(x, x + N, 0.0);
    pmg_vcycle(*P, l, b, x);
}
void orc_pmg_level_spmv(void* p, int l, const double* x, double* y) {
    Pmg* P = (Pmg*)p;
    bsr_spmv(P->lv[l].A, 1.0, x, 0.0, y);
}
void orc_pmg_set(void* p, double omega, int pre, int post) {
    Pmg* P = (Pmg*)p;
    P->omega = omega; P->pre = pre; P->post = post;
}
// one smoother application on level l: x = omega M_l^-1 r
void orc_pmg_smooth(void* p, int l, const double* r, double* x) {
    Pmg* P = (Pmg*)p;
    bgemv(P->lv[l], r, x, P->omega, false);
}
int orc_pmg_set_level_pmg(void* p, int on) {
    try { return pmg_set_level_pmg(*(Pmg*)p, on != 0); } catch (...) { return -1; }
}
int orc_pmg_set_schwarz(void* p, int on, int overlap, int scale) {
    try { pmg_set_schwarz(*(Pmg*)p, on != 0, overlap, scale != 0); return 0; } catch (...) { return -1; }
}
// level l: info = {block rows, block size, transfer kind to l+1, degree, n0, n1, n2}
void orc_pmg_level_info(void* p, int l, int* info) {
    const MgLev& L = ((Pmg*)p)->lv[l];
    info[0] = L.A.nrows; info[1] = L.A.bs; info[2] = L.kind; info[3] = L.deg;
    info[4] = L.n[0]; info[5] = L.n[1]; info[6] = L.n[2];
}
// aggregation transfer of level l (kind 1): parent aggregate per cell and the
// blocks P_i[fine mode][coarse mode]; returns 0, or -1 for a truncation level
int orc_pmg_level_transfer(void* p, int l, int* parent, double* Pb) {
    const MgLev& L = ((Pmg*)p)->lv[l];
    if (L.kind != 1 || L.parent.empty()) return -1;
    std::copy(L.parent.begin(), L.parent.end(), parent);
    std::copy(L.P.begin(), L.P.end(), Pb);
    return 0;
}
// cell boxes of level l (J x D lower corners and widths); -1 where unknown
int orc_pmg_level_boxes(void* p, int l, double* lo, double* h) {
    const MgLev& L = ((Pmg*)p)->lv[l];
    if (L.lo.empty()) return -1;
    std::copy(L.lo.begin(), L.lo.end(), lo);
    std::copy(L.h.begin(), L.h.end(), h);
    return 0;
}
// a copy of the level matrix as a free-standing BSR handle
void* orc_pmg_level_matrix(void* p, int l) { return new Bsr(((Pmg*)p)->lv[l].A); }
// one V-cycle from a zero guess: x = V(b)
int orc_pmg_vcycle(void* mp, void* gp, const int* deg, int ndeg, int min_coarse, double omega, int pre, int post,
                   const double* b, double* x, int* nlevels) {
    try {
        Pmg* P = pmg_build(*(Bsr*)mp, *(Grid*)gp, deg, ndeg, min_coarse, omega, pre, post);
        size_t N = (size_t)P->lv[0].A.nrows * P->lv[0].A.bs;
        std::fill(x, x + N, 0.0);
        pmg_vcycle(*P, 0, b, x);
        if (nlevels) *nlevels = (int)P->lv.size();
        delete P;
        return 0;
    } catch (...) { return -1; }
}
void* orc_euler_create(void* g, int k, int nq1, double gamma, double mach) { return euler_create((Grid*)g, k, nq1, gamma, mach); }
void orc_euler_free(void* e) { delete (Euler*)e; }
void orc_euler_rhs(void* e, const double* U, double* R) { euler_rhs(*(Euler*)e, U, R); }
void orc_euler_tvd3(void* e, double* U, double dt, int steps) { euler_tvd3(*(Euler*)e, U, dt, steps); }
void orc_hllc(int D, double gamma, double mach, const double* Ui, const double* Uo, const double* n, double* F) {
    hllc_flux(D, gamma, gamma * mach * mach, Ui, Uo, n, F);
}
void orc_euler_flux(int D, double gamma, double mach, const double* U, double* F) { euler_flux(D, gamma, gamma * mach * mach, U, F); }
int orc_num_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"

// XDG system on the device: upload of the element lists, assembly of the
// full (cell, species) block matrix, agglomeration into the reduced system,
// right-hand side, and the two-level preconditioner.
//
// Agglomeration (X/CellAgglomerator.cs:1157 GetRowManipulationMatrix,
// :1755 ManipulateMatrixAndRHS): the DG polynomial of species s in a source
// cell is the extension of the target cell's polynomial, u_src = E u_tgt with
// E_mn = int_{src} phi^src_m phi^tgt_n dx (exact: a polynomial of degree k
// restricted to a box).  With the injection Q (identity on kept blocks, E on
// source blocks) the agglomerated system is Q^T A Q u~ = Q^T b, and the
// full-space solution is Q u~.  Source blocks are removed from the reduced
// DOF map (a cell keeps the species blocks that are not sources).
#include <algorithm>
#include <cmath>
#include <map>

#include "bh_solvers.hpp"
#include "bh_xdgdev.hpp"

namespace bh {

void XdgDevice::upload() {
    XdgProblem& p = *P;
    bh_grid* g = p.g;
    nstride = g->nmax + 1;
    std::vector<double> n1((size_t)3 * nstride, 0.0);
    for (int d = 0; d < p.D; ++d) std::copy(g->nodes[d].begin(), g->nodes[d].end(), n1.begin() + (size_t)d * nstride);
    nodes1.alloc(n1.size() * sizeof(double));
    ctx->upload(nodes1.p, n1.data(), n1.size() * sizeof(double));
    elems.alloc(std::max<size_t>(p.elems.size(), 1) * sizeof(XElem));
    ctx->upload(elems.p, p.elems.data(), p.elems.size() * sizeof(XElem));
    // colour groups (see XDG_COLOURS), each with its cut items (own rules, more nodes) first
    auto colour = [&](const XElem& E) {
        if (E.type == 0) return 0;
        if (E.type == 3) return 1;
        if (E.type == 2) return 2 + 2 * E.dir + (E.side ? 1 : 0);
        int iP[3], iQ[3];
        g->cell_ijk(E.cellP, iP);
        g->cell_ijk(E.cellQ, iQ);
        return 8 + 2 * E.dir + (std::min(iP[E.dir], iQ[E.dir]) & 1);
    };
    std::vector<int> cnt(2 * XDG_COLOURS + 1, 0);
    const int ne = (int)p.elems.size();
    std::vector<int> key(ne);
    for (int e = 0; e < ne; ++e) {
        const int c = colour(p.elems[e]);
        BH_CHECK(c >= 0 && c < XDG_COLOURS, "XDG element colour out of range");
        key[e] = 2 * c + (p.elems[e].rn > 0 ? 0 : 1);
        cnt[key[e] + 1]++;
    }
    for (int i = 0; i < 2 * XDG_COLOURS; ++i) cnt[i + 1] += cnt[i];
    std::vector<int> pm(std::max(ne, 1));
    {
        std::vector<int> pos(cnt.begin(), cnt.end() - 1);
        for (int e = 0; e < ne; ++e) pm[pos[key[e]]++] = e;   // stable: the original order within a key
    }
    colour_off.assign(XDG_COLOURS + 1, 0);
    for (int c = 0; c <= XDG_COLOURS; ++c) colour_off[c] = cnt[2 * c];
    cut_end.assign(XDG_COLOURS, 0);
    for (int c = 0; c < XDG_COLOURS; ++c) cut_end[c] = cnt[2 * c + 1];
    perm.alloc(pm.size() * sizeof(int));
    ctx->upload(perm.p, pm.data(), pm.size() * sizeof(int));
    auto up = [&](DevBuf& b, const std::vector<double>& v) {
        b.alloc(std::max<size_t>(v.size(), 1) * sizeof(double));
        ctx->upload(b.p, v.data(), v.size() * sizeof(double));
    };
    up(nx, p.nx);
    up(nw, p.nw);
    up(nn, p.nn);
    std::vector<int> e;
    basis_exponents(p.D, p.k, e);
    ex.alloc(e.size() * sizeof(int));
    ctx->upload(ex.p, e.data(), e.size() * sizeof(int));
    std::vector<double> x(p.k + 1), w(p.k + 1);
    gauss_rule(p.k + 1, x.data(), w.data());
    up(gx, x);
    up(gw, w);
    spmask.alloc(p.spmask.size() * sizeof(int));
    ctx->upload(spmask.p, p.spmask.data(), p.spmask.size() * sizeof(int));
    offf.alloc(p.off_full.size() * sizeof(int64_t));
    ctx->upload(offf.p, p.off_full.data(), p.off_full.size() * sizeof(int64_t));
}

// ------------------------------------------------------------ host helpers
namespace {

// physical basis values / gradients of cell c at physical x
void phys_basis(const XdgProblem& p, int c, const double* x, double* val, double* grad) {
    int idx[3];
    p.g->cell_ijk(c, idx);
    double xi[3], h[3], detJ = 1;
    for (int d = 0; d < p.D; ++d) {
        const double lo = p.g->nodes[d][idx[d]];
        h[d] = p.g->nodes[d][idx[d] + 1] - lo;
        xi[d] = 2.0 * (x[d] - lo) / h[d] - 1.0;
        detJ *= 0.5 * h[d];
    }
    std::vector<double> g((size_t)p.Np * p.D);
    basis_eval(p.D, p.k, xi, val, grad ? g.data() : nullptr);
    const double sc = 1.0 / std::sqrt(detJ);
    for (int m = 0; m < p.Np; ++m) {
        val[m] *= sc;
        if (grad)
            for (int d = 0; d < p.D; ++d) grad[m * p.D + d] = g[(size_t)m * p.D + d] * sc * 2.0 / h[d];
    }
}

// nodes of an element (tensor rule generated like the assembly kernel does)
void elem_nodes(const XdgProblem& p, const XElem& E, std::vector<double>& X, std::vector<double>& W) {
    X.clear();
    W.clear();
    if (E.rn > 0) {
        X.assign(p.nx.begin() + (size_t)E.r0 * 3, p.nx.begin() + (size_t)(E.r0 + E.rn) * 3);
        W.assign(p.nw.begin() + E.r0, p.nw.begin() + E.r0 + E.rn);
        return;
    }
    const int q = p.k + 1;
    std::vector<double> gx(q), gw(q);
    gauss_rule(q, gx.data(), gw.data());
    int idx[3];
    p.g->cell_ijk(E.cellP, idx);
    double lo[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    for (int d = 0; d < p.D; ++d) { lo[d] = p.g->nodes[d][idx[d]]; h[d] = p.g->nodes[d][idx[d] + 1] - lo[d]; }
    const int nd = E.type == 0 ? p.D : p.D - 1;
    int n = 1;
    for (int i = 0; i < nd; ++i) n *= q;
    for (int t = 0; t < n; ++t) {
        double x[3] = {0, 0, 0}, w = 1;
        int r = t;
        for (int d = 0; d < p.D; ++d) {
            if (E.type != 0 && d == E.dir) {
                x[d] = (E.type == 2 && E.side == 0) ? lo[d] : lo[d] + h[d];
                continue;
            }
            const int id = r % q;
            r /= q;
            x[d] = lo[d] + 0.5 * (gx[id] + 1.0) * h[d];
            w *= gw[id] * 0.5 * h[d];
        }
        X.insert(X.end(), x, x + 3);
        W.push_back(w);
    }
}

inline int spos_of(const XdgProblem& p, int c, int sp) { return (sp == 1 && (p.spmask[c] & 1)) ? 1 : 0; }
inline int spos_red(const XdgProblem& p, int c, int sp) {
    // position of species sp among the kept (non-source) species blocks of cell c
    int pos = 0;
    for (int s = 0; s < sp; ++s)
        if (((p.spmask[c] >> s) & 1) && p.agg_target[s][c] < 0) ++pos;
    return pos;
}

}  // namespace

// ------------------------------------------------------------ build
void xdg_build(bh_xdg* X) {
    XdgProblem& p = *X->P;
    bh_ctx* ctx = X->ctx;
    const int J = p.g->J, D = p.D, Np = p.Np;
    X->X.ctx = ctx;
    X->X.P = &p;
    X->X.upload();
    // full matrix structure: every cell couples to itself and its face neighbours
    std::vector<std::vector<int>> cols(J);
    for (int c = 0; c < J; ++c) {
        int idx[3];
        p.g->cell_ijk(c, idx);
        cols[c].push_back(c);
        for (int d = 0; d < D; ++d)
            for (int s = -1; s <= 1; s += 2) {
                int nb[3] = {idx[0], idx[1], idx[2]};
                nb[d] += s;
                if (nb[d] < 0 || nb[d] >= p.g->n[d]) continue;
                cols[c].push_back(p.g->cell_index(nb[0], nb[1], nb[2]));
            }
        std::sort(cols[c].begin(), cols[c].end());
    }
    X->Afull.build(ctx, J, p.bsz_full, cols);
    xdg_assemble(X->X, X->Afull);
    // extension blocks of the agglomeration pairs
    const int npair = (int)p.agg_pairs.size() / 3;
    X->ext.assign((size_t)npair * Np * Np, 0.0);
    {
        const int q = p.k + 1;
        std::vector<double> gx(q), gw(q), vs(Np), vt(Np);
        gauss_rule(q, gx.data(), gw.data());
        int nq = 1;
        for (int d = 0; d < D; ++d) nq *= q;
        for (int i = 0; i < npair; ++i) {
            const int src = p.agg_pairs[3 * i], tgt = p.agg_pairs[3 * i + 1];
            int idx[3];
            p.g->cell_ijk(src, idx);
            double lo[3], h[3];
            for (int d = 0; d < D; ++d) { lo[d] = p.g->nodes[d][idx[d]]; h[d] = p.g->nodes[d][idx[d] + 1] - lo[d]; }
            double* Eb = &X->ext[(size_t)i * Np * Np];
            for (int t = 0; t < nq; ++t) {
                double x[3] = {0, 0, 0}, w = 1;
                int r = t;
                for (int d = 0; d < D; ++d) {
                    const int id = r % q;
                    r /= q;
                    x[d] = lo[d] + 0.5 * (gx[id] + 1.0) * h[d];
                    w *= gw[id] * 0.5 * h[d];
                }
                phys_basis(p, src, x, vs.data(), nullptr);
                phys_basis(p, tgt, x, vt.data(), nullptr);
                for (int m = 0; m < Np; ++m)
                    for (int n = 0; n < Np; ++n) Eb[m * Np + n] += w * vs[m] * vt[n];
            }
        }
    }
    // Reduced (agglomerated) matrix Q^T A Q, as a uniform BSR over species units (a
    // unit = one kept species block of one cell, Np DOFs; cut cells hold two): the DOF
    // layout is the reduced one, the zero species cross blocks between neighbours drop
    // out.  The host lists the contributions (species sub-blocks of the full matrix
    // with their target unit blocks and extension blocks) from the structure alone;
    // the device flags the nonzero ones, the host builds the unit pattern from the
    // flags, and the device sums the values per unit block.  The full matrix never
    // leaves the device.  The cell diagonal blocks (species coupled) for the smoother
    // are gathered from the unit blocks.
    {
        std::vector<int> pair_of[2];
        for (int s = 0; s < 2; ++s) pair_of[s].assign(J, -1);
        for (int i = 0; i < npair; ++i) pair_of[p.agg_pairs[3 * i + 2]][p.agg_pairs[3 * i]] = i;
        std::vector<int> unit0(J + 1, 0);
        for (int r = 0; r < J; ++r) unit0[r + 1] = unit0[r] + p.bsz_red[r] / Np;
        const int U = unit0[J];
        auto target = [&](int c, int s, int& unit, int& pe) {
            const int t = p.agg_target[s][c];
            const int rc = t >= 0 ? t : c;
            pe = t >= 0 ? pair_of[s][c] : -1;
            unit = unit0[rc] + spos_red(p, rc, s);
        };
        // contributions in row order of the full matrix
        std::vector<int64_t> csrc;
        std::vector<int> crf, cu, cv, cpr, cpc;
        for (int r = 0; r < J; ++r) {
            const int Rf = p.bsz_full[r];
            for (int64_t bp = X->Afull.rowptr[r]; bp < X->Afull.rowptr[r + 1]; ++bp) {
                const int c = X->Afull.col[bp];
                for (int sr = 0; sr < 2; ++sr) {
                    if (!((p.spmask[r] >> sr) & 1)) continue;
                    for (int sc = 0; sc < 2; ++sc) {
                        if (!((p.spmask[c] >> sc) & 1)) continue;
                        int u, v, pr, pc;
                        target(r, sr, u, pr);
                        target(c, sc, v, pc);
                        csrc.push_back(X->Afull.voff[bp] + (int64_t)spos_of(p, c, sc) * Np * Rf + spos_of(p, r, sr) * Np);
                        crf.push_back(Rf);
                        cu.push_back(u);
                        cv.push_back(v);
                        cpr.push_back(pr);
                        cpc.push_back(pc);
                    }
                }
            }
        }
        const int64_t nc = (int64_t)csrc.size();
        std::vector<unsigned char> flg(std::max<int64_t>(nc, 1));
        {
            DevBuf dsrc(std::max<int64_t>(nc, 1) * sizeof(int64_t)), drf(std::max<int64_t>(nc, 1) * sizeof(int)), dfl(std::max<int64_t>(nc, 1));
            ctx->upload(dsrc.p, csrc.data(), nc * sizeof(int64_t));
            ctx->upload(drf.p, crf.data(), nc * sizeof(int));
            qaq_flags(ctx, nc, Np, dsrc.as<int64_t>(), drf.as<int>(), X->Afull.vals.as<double>(), (unsigned char*)dfl.p);
            ctx->download(flg.data(), dfl.p, nc);
        }
        // unit pattern: the blocks of nonzero contributions and every diagonal block,
        // columns ascending per row
        std::vector<std::vector<int>> ucols(U);
        for (int u = 0; u < U; ++u) ucols[u].push_back(u);
        for (int64_t i = 0; i < nc; ++i)
            if (flg[i]) ucols[cu[i]].push_back(cv[i]);
        std::vector<int64_t> urp(U + 1, 0);
        std::vector<int> ucol;
        for (int u = 0; u < U; ++u) {
            auto& cl = ucols[u];
            std::sort(cl.begin(), cl.end());
            cl.erase(std::unique(cl.begin(), cl.end()), cl.end());
            ucol.insert(ucol.end(), cl.begin(), cl.end());
            urp[u + 1] = (int64_t)ucol.size();
            std::vector<int>().swap(cl);
        }
        const int64_t nb = urp[U];
        auto block_of = [&](int u, int v) -> int64_t {
            const auto b0 = ucol.begin() + urp[u], b1 = ucol.begin() + urp[u + 1];
            const auto it = std::lower_bound(b0, b1, v);
            return (it != b1 && *it == v) ? (int64_t)(it - ucol.begin()) : -1;
        };
        // contributions grouped by target block, row order kept within a block
        std::vector<int64_t> cblk(nc, -1), cptr(nb + 1, 0);
        for (int64_t i = 0; i < nc; ++i)
            if (flg[i]) { cblk[i] = block_of(cu[i], cv[i]); ++cptr[cblk[i] + 1]; }
        for (int64_t b = 0; b < nb; ++b) cptr[b + 1] += cptr[b];
        const int64_t nk = cptr[nb];
        std::vector<int64_t> ssrc(std::max<int64_t>(nk, 1));
        std::vector<int> srf(std::max<int64_t>(nk, 1)), spr(std::max<int64_t>(nk, 1)), spc(std::max<int64_t>(nk, 1));
        {
            std::vector<int64_t> fillp(cptr.begin(), cptr.end() - 1);
            for (int64_t i = 0; i < nc; ++i) {
                if (cblk[i] < 0) continue;
                const int64_t k = fillp[cblk[i]]++;
                ssrc[k] = csrc[i];
                srf[k] = crf[i];
                spr[k] = cpr[i];
                spc[k] = cpc[i];
            }
        }
        bh_bsr* M = new bh_bsr();
        X->Ared.reset(M);
        M->ctx = ctx;
        M->nrows = U;
        M->bs = Np;
        M->h_rowptr = urp;
        M->h_col = ucol;
        M->nnzb = nb;
        M->rowptr.alloc((U + 1) * sizeof(int64_t));
        M->col.alloc(std::max<int64_t>(nb, 1) * sizeof(int));
        ctx->upload(M->rowptr.p, urp.data(), (U + 1) * sizeof(int64_t));
        ctx->upload(M->col.p, ucol.data(), nb * sizeof(int));
        M->vals.alloc(std::max<int64_t>(nb, 1) * Np * Np * sizeof(double));
        {
            DevBuf dcp((nb + 1) * sizeof(int64_t)), dsrc(std::max<int64_t>(nk, 1) * sizeof(int64_t));
            DevBuf drf(std::max<int64_t>(nk, 1) * sizeof(int)), dpr(std::max<int64_t>(nk, 1) * sizeof(int)),
                dpc(std::max<int64_t>(nk, 1) * sizeof(int)), dext(std::max<size_t>(X->ext.size(), 1) * sizeof(double));
            ctx->upload(dcp.p, cptr.data(), (nb + 1) * sizeof(int64_t));
            ctx->upload(dsrc.p, ssrc.data(), nk * sizeof(int64_t));
            ctx->upload(drf.p, srf.data(), nk * sizeof(int));
            ctx->upload(dpr.p, spr.data(), nk * sizeof(int));
            ctx->upload(dpc.p, spc.data(), nk * sizeof(int));
            ctx->upload(dext.p, X->ext.data(), X->ext.size() * sizeof(double));
            qaq_gather(ctx, nb, Np, dcp.as<int64_t>(), dsrc.as<int64_t>(), drf.as<int>(), dpr.as<int>(), dpc.as<int>(),
                       X->Afull.vals.as<double>(), dext.as<double>(), M->vals.as<double>());
            ctx->sync();
        }
        // cell diagonal blocks
        std::vector<std::vector<int>> dcols(J);
        std::vector<int64_t> dblk((size_t)4 * J, -1);
        for (int r = 0; r < J; ++r) {
            if (p.bsz_red[r] == 0) continue;
            dcols[r].push_back(r);
            const int nu = p.bsz_red[r] / Np;
            for (int a = 0; a < nu; ++a)
                for (int c = 0; c < nu; ++c) dblk[(size_t)4 * r + a * 2 + c] = block_of(unit0[r] + a, unit0[r] + c);
        }
        X->Dcell.build(ctx, J, p.bsz_red, dcols);
        {
            std::vector<int64_t> doff(J, 0);
            for (int r = 0; r < J; ++r)
                if (X->Dcell.rowptr[r + 1] > X->Dcell.rowptr[r]) doff[r] = X->Dcell.voff[X->Dcell.rowptr[r]];
            DevBuf ddo(std::max(J, 1) * sizeof(int64_t)), ddb(std::max<size_t>(dblk.size(), 1) * sizeof(int64_t));
            ctx->upload(ddo.p, doff.data(), J * sizeof(int64_t));
            ctx->upload(ddb.p, dblk.data(), dblk.size() * sizeof(int64_t));
            qaq_dcell(ctx, J, Np, X->Dcell.rsz_d.as<int>(), ddo.as<int64_t>(), ddb.as<int64_t>(), M->vals.as<double>(),
                      X->Dcell.vals.as<double>());
            ctx->sync();
        }
    }
    // boundary nodes (Dirichlet data)
    X->bnd_x.clear(); X->bnd_w.clear(); X->bnd_nu.clear(); X->bnd_mu.clear();
    X->bnd_cell.clear(); X->bnd_sp.clear(); X->bnd_dir.clear(); X->bnd_side.clear();
    std::vector<double> NX, NW;
    for (const XElem& E : p.elems) {
        if (E.type != 2) continue;
        elem_nodes(p, E, NX, NW);
        for (size_t t = 0; t < NW.size(); ++t) {
            X->bnd_x.insert(X->bnd_x.end(), &NX[3 * t], &NX[3 * t] + 3);
            X->bnd_w.push_back(NW[t]);
            X->bnd_nu.push_back(E.nuP);
            X->bnd_mu.push_back(E.mu);
            X->bnd_cell.push_back(E.cellP);
            X->bnd_sp.push_back(E.spP);
            X->bnd_dir.push_back(E.dir);
            X->bnd_side.push_back(E.side);
        }
    }
    X->rhs_red.alloc(std::max<int64_t>(p.N_red, 1) * sizeof(double));
    ctx->sync();
}

// right-hand side: F - affine, F_m = int f_s phi_m (species-constant source f_s),
// affine = boundary terms of the Dirichlet data g (SIPLaplace.BoundaryEdgeForm):
// -nu g dphi_m/dn + nu mu g phi_m
void xdg_set_rhs(bh_xdg* X, const double* f, const double* g) {
    XdgProblem& p = *X->P;
    const int Np = p.Np, D = p.D;
    X->fsrc = {f[0], f[1]};
    X->rhs_full.assign((size_t)p.N_full, 0.0);
    std::vector<double> NX, NW, v(Np), gr((size_t)Np * D);
    for (const XElem& E : p.elems) {
        if (E.type != 0) continue;
        const double fs = f[E.spP];
        if (fs == 0.0) continue;
        elem_nodes(p, E, NX, NW);
        double* o = &X->rhs_full[p.off_full[E.cellP] + spos_of(p, E.cellP, E.spP) * Np];
        for (size_t t = 0; t < NW.size(); ++t) {
            phys_basis(p, E.cellP, &NX[3 * t], v.data(), nullptr);
            for (int m = 0; m < Np; ++m) o[m] += NW[t] * fs * v[m];
        }
    }
    if (g)
        for (size_t t = 0; t < X->bnd_w.size(); ++t) {
            const int c = X->bnd_cell[t];
            phys_basis(p, c, &X->bnd_x[3 * t], v.data(), gr.data());
            const double nsign = X->bnd_side[t] ? 1.0 : -1.0;
            const double nu = X->bnd_nu[t], mu = X->bnd_mu[t], w = X->bnd_w[t];
            double* o = &X->rhs_full[p.off_full[c] + spos_of(p, c, X->bnd_sp[t]) * Np];
            for (int m = 0; m < Np; ++m) {
                const double dn = nsign * gr[(size_t)m * D + X->bnd_dir[t]];
                o[m] -= w * (-nu * g[t] * dn + nu * mu * g[t] * v[m]);
            }
        }
    // reduced: Q^T b
    std::vector<double> red((size_t)p.N_red, 0.0);
    const int J = p.g->J;
    std::vector<int> pair_of[2];
    for (int s = 0; s < 2; ++s) pair_of[s].assign(J, -1);
    const int npair = (int)p.agg_pairs.size() / 3;
    for (int i = 0; i < npair; ++i) pair_of[p.agg_pairs[3 * i + 2]][p.agg_pairs[3 * i]] = i;
    for (int c = 0; c < J; ++c)
        for (int s = 0; s < 2; ++s) {
            if (!((p.spmask[c] >> s) & 1)) continue;
            const double* b = &X->rhs_full[p.off_full[c] + spos_of(p, c, s) * Np];
            const int t = p.agg_target[s][c];
            if (t < 0) {
                double* o = &red[p.off_red[c] + spos_red(p, c, s) * Np];
                for (int m = 0; m < Np; ++m) o[m] += b[m];
            } else {
                const double* E = &X->ext[(size_t)pair_of[s][c] * Np * Np];
                double* o = &red[p.off_red[t] + spos_red(p, t, s) * Np];
                for (int n = 0; n < Np; ++n) {
                    double a = 0;
                    for (int m = 0; m < Np; ++m) a += E[m * Np + n] * b[m];
                    o[n] += a;
                }
            }
        }
    X->ctx->upload(X->rhs_red.p, red.data(), red.size() * sizeof(double));
}

// u_full = Q u_red (host)
void xdg_extend(const bh_xdg* X, const double* ured, double* ufull) {
    const XdgProblem& p = *X->P;
    const int J = p.g->J, Np = p.Np;
    std::vector<int> pair_of[2];
    for (int s = 0; s < 2; ++s) pair_of[s].assign(J, -1);
    const int npair = (int)p.agg_pairs.size() / 3;
    for (int i = 0; i < npair; ++i) pair_of[p.agg_pairs[3 * i + 2]][p.agg_pairs[3 * i]] = i;
    for (int c = 0; c < J; ++c)
        for (int s = 0; s < 2; ++s) {
            if (!((p.spmask[c] >> s) & 1)) continue;
            double* o = &ufull[p.off_full[c] + spos_of(p, c, s) * Np];
            const int t = p.agg_target[s][c];
            if (t < 0) {
                const double* u = &ured[p.off_red[c] + spos_red(p, c, s) * Np];
                for (int m = 0; m < Np; ++m) o[m] = u[m];
            } else {
                const double* E = &X->ext[(size_t)pair_of[s][c] * Np * Np];
                const double* u = &ured[p.off_red[t] + spos_red(p, t, s) * Np];
                for (int m = 0; m < Np; ++m) {
                    double a = 0;
                    for (int n = 0; n < Np; ++n) a += E[m * Np + n] * u[n];
                    o[m] = a;
                }
            }
        }
}

// ------------------------------------------------------------ preconditioner
XdgPrecond::~XdgPrecond() { delete Ac; }

void XdgPrecond::build(bh_grid* g, int degc, int min_coarse, int kcycle_inner) {
    ctx = Av ? Av->ctx : A->ctx;
    N = Av ? Av->N : (int64_t)A->nrows * A->bs;
    const int D = g->D, J = g->J;
    Npc = basis_count(D, degc);
    BH_CHECK(Npc <= Np && (Av || A->bs == Np), "XDG preconditioner: coarse degree above the DG degree");
    BH_CHECK(Dc->N == N && Dc->nrows == J, "XDG preconditioner: cell blocks do not match the matrix");
    for (int c = 0; c < J; ++c)
        BH_CHECK(Dc->rsz[c] % Np == 0 && Dc->rsz[c] <= 2 * Np, "XDG preconditioner: cell blocks must hold 0, Np or 2 Np values");
    // block inverses of the smoother (cell blocks, species coupled)
    Dc->invert_diag(prec == BH_FP32);
    if (prec == BH_FP32) {
        if (Av) Av->ensure_fp32();
        else A->ensure_fp32();
    }
    // coarse matrix Ac = R^T A R (species merged, first Npc modes of every species block):
    // the cell pattern on the host from the block structure, the values summed on the
    // device (the matrix values never leave it)
    std::vector<int64_t> rp(J + 1, 0);
    std::vector<int> col;
    DevBuf cvals;
    if (Av) {
        rp = Av->rowptr;
        col = Av->col;
        std::vector<int> prow(std::max<size_t>(col.size(), 1));
        for (int r = 0; r < J; ++r)
            for (int64_t p = Av->rowptr[r]; p < Av->rowptr[r + 1]; ++p) prow[p] = r;
        DevBuf dprow(prow.size() * sizeof(int));
        ctx->upload(dprow.p, prow.data(), prow.size() * sizeof(int));
        cvals.alloc(std::max<size_t>(col.size(), 1) * Npc * Npc * sizeof(double));
        if (!col.empty()) xdg_coarse_vblocks(ctx, *Av, Np, Npc, dprow.as<int>(), cvals.as<double>());
        ctx->sync();
    } else {
        std::vector<int> cell_of(A->nrows);
        for (int c = 0, u = 0; c < J; ++c)
            for (int a = 0; a < Dc->rsz[c] / Np; ++a) cell_of[u++] = c;
        // per cell row: (cell column, unit block) in unit-row then unit-column order; a
        // stable sort by cell column keeps that order within each coarse block
        std::vector<std::vector<std::pair<int, int64_t>>> ent(J);
        for (int u = 0; u < A->nrows; ++u)
            for (int64_t bp = A->h_rowptr[u]; bp < A->h_rowptr[u + 1]; ++bp) ent[cell_of[u]].push_back({cell_of[A->h_col[bp]], bp});
        std::vector<int64_t> qptr{0}, plist;
        plist.reserve((size_t)A->nnzb);
        for (int r = 0; r < J; ++r) {
            auto& e = ent[r];
            std::stable_sort(e.begin(), e.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (size_t i = 0; i < e.size(); ++i) {
                if (i == 0 || e[i].first != e[i - 1].first) {
                    if (i > 0) qptr.push_back((int64_t)plist.size());
                    col.push_back(e[i].first);
                }
                plist.push_back(e[i].second);
            }
            if (!e.empty()) qptr.push_back((int64_t)plist.size());
            rp[r + 1] = (int64_t)col.size();
            std::vector<std::pair<int, int64_t>>().swap(e);
        }
        DevBuf dq(qptr.size() * sizeof(int64_t)), dl(std::max<size_t>(plist.size(), 1) * sizeof(int64_t));
        ctx->upload(dq.p, qptr.data(), qptr.size() * sizeof(int64_t));
        ctx->upload(dl.p, plist.data(), plist.size() * sizeof(int64_t));
        cvals.alloc(std::max<size_t>(col.size(), 1) * Npc * Npc * sizeof(double));
        if (!col.empty())
            xdg_coarse_gather(ctx, (int64_t)col.size(), Np, Npc, dq.as<int64_t>(), dl.as<int64_t>(), A->vals.as<double>(),
                              cvals.as<double>());
        ctx->sync();
    }
    Ac = new bh_bsr();
    Ac->ctx = ctx;
    Ac->nrows = J;
    Ac->bs = Npc;
    Ac->h_rowptr = rp;
    Ac->h_col = col;
    Ac->nnzb = rp[J];
    Ac->rowptr.alloc((J + 1) * sizeof(int64_t));
    Ac->col.alloc(std::max<int64_t>(Ac->nnzb, 1) * sizeof(int));
    ctx->upload(Ac->rowptr.p, rp.data(), (J + 1) * sizeof(int64_t));
    ctx->upload(Ac->col.p, col.data(), col.size() * sizeof(int));
    Ac->vals = std::move(cvals);
    cmg.reset(new PMultigrid());
    cmg->prec = prec;
    cmg->omega = omega;
    cmg->pre = pre;
    cmg->post = post;
    cmg->kcycle_level = kcycle_inner > 0 ? 0 : -1;
    cmg->kin = kcycle_inner;
    cmg->build_grid(ctx, g, Ac, {degc}, min_coarse, nullptr);
    const size_t es = prec == BH_FP32 ? sizeof(float) : sizeof(double);
    b.alloc(N * es);
    x.alloc(N * es);
    r.alloc(N * es);
    ctx->sync();
}

template <typename T>
void XdgPrecond::residual(const T* bb, const T* xx, T* rr) {
    const int p = sizeof(T) == 4 ? BH_FP32 : BH_FP64;
    if (Av) Av->spmv(-1.0, xx, 0.0, rr, p, bb);
    else bsr_residual(A, bb, xx, rr, p);
}

template <typename T>
void XdgPrecond::cycle() {
    const int p = sizeof(T) == 4 ? BH_FP32 : BH_FP64;
    T* bb = (T*)b.p;
    T* xx = (T*)x.p;
    T* rr = (T*)r.p;
    Dc->block_gemv(bb, xx, omega, false, p);
    for (int s = 1; s < pre; ++s) {
        residual<T>(bb, xx, rr);
        Dc->block_gemv(rr, xx, omega, true, p);
    }
    residual<T>(bb, xx, rr);
    MgLevel& C = cmg->lv[0];
    xdg_restrict<T>(ctx, *Dc, Np, Npc, rr, (T*)C.b.p);
    cmg->solve_level<T>(0);
    xdg_prolong_add<T>(ctx, *Dc, Np, Npc, (const T*)C.x.p, xx);
    for (int s = 0; s < post; ++s) {
        residual<T>(bb, xx, rr);
        Dc->block_gemv(rr, xx, omega, true, p);
    }
}

void XdgPrecond::precondition(double* xo, const double* bi) {
    if (prec == BH_FP32) {
        convert_f64_f32(ctx, bi, (float*)b.p, N);
        cycle<float>();
        convert_f32_f64(ctx, (const float*)x.p, xo, N);
    } else {
        copy(ctx, (double*)b.p, bi, N);
        cycle<double>();
        copy(ctx, xo, (const double*)x.p, N);
    }
    iterations++;
}

void XdgPrecond::solve_dev(double* xo, const double* bi) {
    DevBuf rr(N * sizeof(double)), c(N * sizeof(double));
    residual<double>(bi, xo, rr.as<double>());
    precondition(c.as<double>(), rr.as<double>());
    axpby(ctx, 1.0, c.as<double>(), 1.0, xo, N);
    converged = 1;
}

}  // namespace bh

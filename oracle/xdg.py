"""XDG oracle (TEST INFRASTRUCTURE): two-phase SIP Poisson with a spherical
interface, restated in numpy/scipy.  Only the tests use it, as the checker.

Three independent parts:

* geometry of a ball / disk against axis-aligned boxes, in closed form
  (area of disk ∩ rectangle, arc length of circle ∩ rectangle) plus
  adaptive 1D quadrature (scipy.integrate.quad) for the 3D volume and surface
  pieces -- this pins the cut-cell volumes, face areas and interface areas of
  the product's dimension-reduction rules without sharing any code with them;
* the agglomeration of X/AgglomerationAlgorithm.cs (sources :393-460, direct
  targets :2064-2200, chains :1815-2055, groups :2232-2330) and the length
  scales of X/MultiphaseAgglomerator.cs:870-1030, from those volumes/areas;
* the bilinear forms XLaplace_Bulk (SIPLaplace with species diffusion and
  species length scales) and XLaplace_Interface in SIP mode
  (L3-solution/BoSSS.Solution.XNSECommon/XLaplace.cs:50-346), integrated
  with a given set of element rules (tensor Gauss or cut-cell nodes), the
  agglomeration Q^T A Q (X/CellAgglomerator.cs:1157, :1755) and the
  right-hand side.

Parity notes: the reference's per-cell face order (Cells2Edges, built in
F/GridData_EdgeData.cs:2283-2320 from its edge enumeration) decides ties of
the strict ">" in the target search; the order used here is, per direction,
the lower face then the upper face.  Ties between equal volume fractions
(e.g. two uncut neighbours, fraction 1) are therefore parity unpinned with
respect to the reference, though pinned between this oracle and the product.
"""
import math

import numpy as np
from scipy import integrate
import scipy.sparse as sp

from . import basis_eval, basis_exponents, basis_count, gauss

EMPTY_EDGE = 1e-10   # EmptyEdgeTreshold of AgglomerationAlgorithm


# ---------------------------------------------------------------- geometry
def _sqrt_int(R, a, b):
    """int_a^b sqrt(R^2 - t^2) dt, |a|, |b| <= R."""
    def F(t):
        t = min(max(t, -R), R)
        return 0.5 * (t * math.sqrt(max(R * R - t * t, 0.0)) + R * R * math.asin(t / R))
    return F(b) - F(a)


def disk_rect_area(R, cx, cy, x0, x1, y0, y1):
    """Area of {(x, y) in [x0,x1]x[y0,y1] : (x-cx)^2 + (y-cy)^2 < R^2} (closed form)."""
    if R <= 0:
        return 0.0
    u0, u1 = max(x0 - cx, -R), min(x1 - cx, R)
    if u0 >= u1:
        return 0.0
    v0, v1 = y0 - cy, y1 - cy
    bps = {u0, u1}
    for v in (v0, v1):
        if abs(v) < R:
            s = math.sqrt(R * R - v * v)
            for u in (-s, s):
                if u0 < u < u1:
                    bps.add(u)
    bps = sorted(bps)
    tot = 0.0
    for a, b in zip(bps[:-1], bps[1:]):
        m = 0.5 * (a + b)
        s = math.sqrt(max(R * R - m * m, 0.0))
        hi, lo = min(v1, s), max(v0, -s)
        if hi <= lo:
            continue
        up = v1 * (b - a) if v1 < s else _sqrt_int(R, a, b)
        dn = v0 * (b - a) if v0 > -s else -_sqrt_int(R, a, b)
        tot += up - dn
    return tot


def circle_rect_angle(R, cx, cy, x0, x1, y0, y1):
    """Total angle of the circle of radius R about (cx, cy) inside the rectangle."""
    if R <= 0:
        return 0.0
    cuts = [0.0, 2 * math.pi]
    for (c, lo, hi, cosine) in ((cx, x0, x1, True), (cy, y0, y1, False)):
        for line in (lo, hi):
            t = (line - c) / R
            if -1 < t < 1:
                base = math.acos(t) if cosine else math.asin(t)
                cand = (base, -base) if cosine else (base, math.pi - base)
                for a in cand:
                    cuts.append(a % (2 * math.pi))
    cuts = sorted(cuts)
    tot = 0.0
    for a, b in zip(cuts[:-1], cuts[1:]):
        if b - a <= 0:
            continue
        m = 0.5 * (a + b)
        x, y = cx + R * math.cos(m), cy + R * math.sin(m)
        if x0 <= x <= x1 and y0 <= y <= y1:
            tot += b - a
    return tot


def _quad(f, a, b, pts):
    pts = sorted(p for p in pts if a < p < b)
    v, _ = integrate.quad(f, a, b, points=pts or None, epsabs=1e-15, epsrel=1e-13, limit=400)
    return v


def box_ball_volume(lo, hi, c, r):
    """Volume (3D) / area (2D) of box ∩ ball."""
    D = len(lo)
    if D == 2:
        return disk_rect_area(r, c[0], c[1], lo[0], hi[0], lo[1], hi[1])
    z0, z1 = max(lo[2], c[2] - r), min(hi[2], c[2] + r)
    if z0 >= z1:
        return 0.0
    f = lambda z: disk_rect_area(math.sqrt(max(r * r - (z - c[2]) ** 2, 0.0)), c[0], c[1], lo[0], hi[0], lo[1], hi[1])
    pts = _kinks(lo, hi, c, r)
    return _quad(f, z0, z1, pts)


def _kinks(lo, hi, c, r):
    """z where the slice radius equals the distance to an edge or corner of the xy box."""
    out = [c[2]]
    ds = []
    for x in (lo[0], hi[0]):
        ds.append(abs(x - c[0]))
        for y in (lo[1], hi[1]):
            ds.append(math.hypot(x - c[0], y - c[1]))
    for y in (lo[1], hi[1]):
        ds.append(abs(y - c[1]))
    for d in ds:
        if d < r:
            s = math.sqrt(r * r - d * d)
            out += [c[2] - s, c[2] + s]
    return out


def box_sphere_area(lo, hi, c, r):
    """Area (3D) / length (2D) of the sphere |x - c| = r inside the box."""
    D = len(lo)
    if D == 2:
        return r * circle_rect_angle(r, c[0], c[1], lo[0], hi[0], lo[1], hi[1])
    z0, z1 = max(lo[2], c[2] - r), min(hi[2], c[2] + r)
    if z0 >= z1:
        return 0.0
    # dS = r dphi dz on the slices (slice radius rho, arc length rho*angle, dS = r/rho * rho*angle dz)
    f = lambda z: r * circle_rect_angle(math.sqrt(max(r * r - (z - c[2]) ** 2, 0.0)), c[0], c[1], lo[0], hi[0], lo[1],
                                        hi[1])
    return _quad(f, z0, z1, _kinks(lo, hi, c, r))


def face_ball_area(D, lo, hi, d, c, r):
    """Measure of (face in the plane x_d = lo[d]) ∩ ball."""
    t = lo[d]
    R2 = r * r - (t - c[d]) ** 2
    if R2 <= 0:
        return 0.0
    R = math.sqrt(R2)
    o = [e for e in range(D) if e != d]
    if D == 2:
        e = o[0]
        a, b = max(lo[e], c[e] - R), min(hi[e], c[e] + R)
        return max(b - a, 0.0)
    return disk_rect_area(R, c[o[0]], c[o[1]], lo[o[0]], hi[o[0]], lo[o[1]], hi[o[1]])


# ---------------------------------------------------------------- cells, faces, agglomeration
class Geometry:
    """Species volumes, face areas and interface areas of every cell of a
    Cartesian grid (reference numbering), computed by the closed forms above."""

    def __init__(self, nodes, c, r):
        self.D = D = len(nodes)
        self.nodes = [np.asarray(a, float) for a in nodes]
        self.n = [len(a) - 1 for a in nodes] + [1] * (3 - D)
        self.c = np.asarray(c, float)[:D]
        self.r = float(r)
        self.J = J = int(np.prod(self.n[:D]))
        self.spmask = np.zeros(J, np.int32)
        self.vol = np.zeros((2, J))
        self.iface = np.zeros(J)
        self.cut = np.zeros(J, bool)
        self.cellvol = np.zeros(J)
        # faces: (a, b, dir, side, area[2])
        self.faces = []
        self.cell_face = -np.ones((J, 6), np.int64)
        for j in range(J):
            lo, hi = self.box(j)
            V = float(np.prod(hi - lo))
            self.cellvol[j] = V
            dmin2 = sum(0.0 if lo[d] <= self.c[d] <= hi[d] else min((lo[d] - self.c[d]) ** 2, (hi[d] - self.c[d]) ** 2)
                        for d in range(D))
            dmax2 = sum(max((lo[d] - self.c[d]) ** 2, (hi[d] - self.c[d]) ** 2) for d in range(D))
            if dmax2 <= r * r:
                va = V
            elif dmin2 >= r * r:
                va = 0.0
            else:
                va = box_ball_volume(lo, hi, self.c, r)
                self.iface[j] = box_sphere_area(lo, hi, self.c, r)
            self.vol[0, j] = va
            self.vol[1, j] = max(V - va, 0.0)
            m = (1 if self.vol[0, j] > 0 else 0) | (2 if self.vol[1, j] > 0 else 0)
            self.spmask[j] = m
            self.cut[j] = m == 3
        for j in range(J):
            idx = self.ijk(j)
            lo, hi = self.box(j)
            for d in range(D):
                for s in (0, 1):
                    bnd = idx[d] == 0 if s == 0 else idx[d] == self.n[d] - 1
                    if not bnd and s == 0:
                        continue
                    if bnd:
                        b = -1
                    else:
                        nb = list(idx)
                        nb[d] += 1
                        b = self.index(*nb)
                    fl, fh = lo.copy(), hi.copy()
                    fl[d] = fh[d] = hi[d] if s else lo[d]
                    A = float(np.prod([fh[e] - fl[e] for e in range(D) if e != d]))
                    aa = face_ball_area(D, fl, fh, d, self.c, r)
                    aa = min(aa, A)
                    fi = len(self.faces)
                    self.faces.append((j, b, d, s, np.array([aa, max(A - aa, 0.0)])))
                    self.cell_face[j, 2 * d + s] = fi
                    if b >= 0:
                        self.cell_face[b, 2 * d] = fi

    def ijk(self, j):
        n = self.n
        if self.D == 2:
            return [j // n[1], j % n[1], 0]
        return [j % n[0], (j // n[0]) % n[1], j // (n[0] * n[1])]

    def index(self, i, jj, k=0):
        n = self.n
        return i * n[1] + jj if self.D == 2 else i + jj * n[0] + k * n[0] * n[1]

    def box(self, j):
        idx = self.ijk(j)
        lo = np.array([self.nodes[d][idx[d]] for d in range(self.D)])
        hi = np.array([self.nodes[d][idx[d] + 1] for d in range(self.D)])
        return lo, hi

    def center(self, j):
        lo, hi = self.box(j)
        return 0.5 * (lo + hi)

    def _neighbours(self, j, s):
        """(neighbour or -1, cut-edge area of species s) per face of cell j, in the
        order per direction: lower face, then upper face (no periodic faces here)."""
        idx = self.ijk(j)
        out = []
        for d in range(self.D):
            for side in (0, 1):
                nb = list(idx)
                nb[d] += 1 if side else -1
                jn = self.index(*nb) if 0 <= nb[d] < self.n[d] else -1
                fi = self.cell_face[j, 2 * d + side]
                out.append((jn, self.faces[fi][4][s]))
        return out

    def agglomerate(self, threshold):
        """Agglomeration map per species, X/AgglomerationAlgorithm.cs restated for one
        MPI rank and a static interface:

        * sources: cut cells of the species with clipped volume fraction <= threshold
          (FindAgglomerationSources, :393-460); candidates: cells where the species is
          present and which are no source ("make sure source cells cannot be target
          cells", FindAgglomerationTargets_Mk3, :1721-1726);
        * direct pass (SearchDirectAgglomeration_Mk3, :2064-2200): the face neighbour
          among the candidates with the largest raw volume fraction (strict >, face
          order), skipping empty cut edges when a non-empty one exists; a source
          without a target goes to the chain list when a neighbour is a source, else
          it fails;
        * chain pass (SearchChainAgglomeration_Mk3, :1815-2055): repeatedly connect a
          waiting source to the final target of a paired neighbour source, directly
          when all its neighbour sources are paired (nearest target, then largest
          target fraction), else one connection per sweep from the globally best
          edge (distance ascending, fraction descending, stable order);
        * group pass (SearchGroupAgglomeration_Mk3, :2232-2330, global mapping): the
          waiting source with the largest species volume (then the lowest index)
          becomes a target and takes every waiting source connected to it through
          non-empty edges (AgglomerationGroup.AddNeighbors, :72-111); then the chain
          pass runs again; the self pairs are removed at the end.

        Returns the pairs (source, target, species) in the order the passes find
        them, the per-species target arrays and the (source, species) that found no
        target."""
        D, J = self.D, self.J
        pairs_out, failed_out = [], []
        target = -np.ones((2, J), np.int64)
        for s in range(2):
            vol = self.vol[s]
            present = ((self.spmask >> s) & 1) == 1
            fsrc = np.clip(np.maximum(vol, 0.0) / self.cellvol, 0.0, 1.0)
            is_src = self.cut & present & (fsrc <= threshold)
            cand = present & ~is_src
            sources = [j for j in range(J) if is_src[j]]
            pairs = []        # [source, target, fracTarget, posTarget]
            need, failed = [], []
            # direct pass
            for j in sources:
                nbs = self._neighbours(j, s)
                non_empty = any(jn >= 0 and a > EMPTY_EDGE for jn, a in nbs)
                if vol[j] <= 0:
                    non_empty = False
                best, tgt = -1.0, -1
                for jn, a in nbs:
                    if jn < 0 or (a <= EMPTY_EDGE and non_empty):
                        continue
                    if not cand[jn]:
                        continue
                    f = vol[jn] / self.cellvol[jn]
                    if f > best:
                        best, tgt = f, jn
                if tgt < 0:
                    if any(jn >= 0 and is_src[jn] for jn, _ in nbs):
                        need.append(j)
                    else:
                        failed.append(j)
                else:
                    pairs.append([j, tgt, best, self.center(tgt)])

            def pair_of(src):
                hits = [p for p in pairs if p[0] == src]
                assert len(hits) <= 1, "a cell is mapped to two different target cells"
                return hits[0] if hits else None

            def chain():
                update = True
                while need and update:
                    update = False
                    loop_pairs, weighted, immediate = [], [], []
                    for j in need:
                        nsrc = npaired = 0
                        wj = []
                        cj = self.center(j)
                        for jn, a in self._neighbours(j, s):
                            if jn < 0:
                                continue
                            if is_src[jn]:
                                nsrc += 1
                            tp = pair_of(jn)
                            if tp is not None:
                                dist = math.sqrt(float(((tp[3] - cj) ** 2).sum()))
                                wj.append((dist, tp[2], j, jn, tp))
                                npaired += 1
                        if npaired == nsrc and npaired > 0:
                            e = sorted(wj, key=lambda t: (t[0], -t[1]))[0]
                            loop_pairs.append([j, e[4][1], e[4][2], e[4][3]])
                            immediate.append(j)
                        else:
                            weighted.extend(wj)
                    for j in immediate:
                        need.remove(j)
                        update = True
                    weighted.sort(key=lambda t: (t[0], -t[1]))
                    if weighted and need:
                        e = weighted[0]
                        loop_pairs.append([e[2], e[4][1], e[4][2], e[4][3]])
                        need.remove(e[2])
                        update = True
                    pairs.extend(loop_pairs)

            def group_neighbours(j):
                nbs = self._neighbours(j, s)
                non_empty = any(jn >= 0 and a > EMPTY_EDGE for jn, a in nbs)
                return [jn for jn, a in nbs if jn >= 0 and (a > EMPTY_EDGE or not non_empty)]

            chain()
            while need:
                gt = min(need, key=lambda p: (-vol[p], p))
                members = [gt]
                nbl = group_neighbours(gt)
                need.remove(gt)
                while True:
                    seen = set()
                    inter = []
                    for c in nbl:
                        if c in seen:
                            continue
                        seen.add(c)
                        if c in need:
                            inter.append(c)
                    if not inter:
                        break
                    for c in inter:
                        if c not in members:
                            members.append(c)
                            nbl = nbl + group_neighbours(c)
                        need.remove(c)
                pos = self.center(gt)
                pairs.extend([[c, gt, vol[gt], pos] for c in members])
                chain()
            failed.extend(need)
            pairs = [p for p in pairs if p[0] != p[1]]
            for p in pairs:
                target[s, p[0]] = p[1]
                pairs_out.append((p[0], p[1], s))
            failed_out.extend((j, s) for j in failed)
        return np.array(pairs_out, np.int64).reshape(-1, 3), target, failed_out

    def agg_groups(self, target, s):
        """cells of each agglomerated group of species s: target -> [target, sources...]."""
        groups = {}
        for j in range(self.J):
            t = target[s, j]
            if t >= 0:
                groups.setdefault(int(t), [int(t)]).append(j)
        return groups

    def length_scales(self, target):
        """Agglomerated species volume / (interface area + cut-edge areas without the
        agglomeration edges) (MultiphaseAgglomerator.cs:870-1030).  Agglomeration
        edges are the faces between two cells of one agglomerated group, target and
        sources (CellAgglomerator.cs:418-446); sources add their volume and surface
        to their target, which hands the sums back to them."""
        J = self.J
        out = np.zeros((2, J))
        for s in range(2):
            present = ((self.spmask >> s) & 1) == 1
            gid = -np.ones(J, np.int64)
            for j in range(J):
                t = target[s, j]
                if t >= 0:
                    gid[j] = t
                    gid[t] = t
            surf = np.where(present, self.iface, 0.0)
            vol = self.vol[s].copy()
            for (a, b, d, side, area) in self.faces:
                if b >= 0 and gid[a] >= 0 and gid[a] == gid[b]:
                    continue
                surf[a] += area[s]
                if b >= 0:
                    surf[b] += area[s]
            for j in range(J):
                t = target[s, j]
                if t >= 0:
                    surf[t] += surf[j]
                    vol[t] += vol[j]
            for j in range(J):
                t = target[s, j]
                if t >= 0:
                    surf[j] = surf[t]
                    vol[j] = vol[t]
            out[s] = np.where(surf > 1e-25, vol / np.where(surf > 1e-25, surf, 1.0), 0.0)
        return out


# ---------------------------------------------------------------- assembly
def _phys_basis(D, k, lo, h, x):
    """orthonormal basis values [n, Np] and gradients [n, Np, D] of the cell (lo, h) at physical x [n, D]."""
    xi = 2.0 * (x - lo) / h - 1.0
    v, g = basis_eval(D, k, xi)
    detJ = np.prod(0.5 * h)
    sc = 1.0 / math.sqrt(detJ)
    return v * sc, g * sc * (2.0 / h)[None, None, :]


def _tensor_nodes(D, k, lo, h, kind, d=None, upper=False):
    """Gauss k+1 tensor nodes/weights of the cell volume (kind 0) or of a face x_d = lo_d (+h_d if upper)."""
    gx, gw = gauss(k + 1)
    dims = [e for e in range(D) if kind == 0 or e != d]
    grids = np.meshgrid(*[gx] * len(dims), indexing="ij")
    wgrid = np.ones_like(grids[0])
    for gi in range(len(dims)):
        wgrid = wgrid * np.meshgrid(*[gw] * len(dims), indexing="ij")[gi]
    X = np.zeros((grids[0].size, D))
    W = wgrid.ravel().copy()
    for gi, e in enumerate(dims):
        X[:, e] = lo[e] + 0.5 * (grids[gi].ravel() + 1) * h[e]
        W = W * 0.5 * h[e]
    if kind != 0:
        X[:, d] = lo[d] + (h[d] if upper else 0.0)
    return X, W


class XdgAssembly:
    """Full (cell, species) block matrix of the two-phase SIP operator, from the
    element list (elements: int rows [type, P, Q, spP, spQ, dir, side, r0, rn]) and
    rule nodes; the penalties are recomputed from the length scales given."""

    def __init__(self, geom, k, nu, penalty_base, lengths, elems, nx, nw, nn):
        self.g = geom
        self.k = k
        D = geom.D
        self.Np = Np = basis_count(D, k)
        p = float(k)
        self.penalty_deg = max((p + 1) * (p + D) / D, (p + 1) ** 2)
        pb = penalty_base * self.penalty_deg
        J = geom.J
        nsp = np.array([bin(int(m)).count("1") for m in geom.spmask])
        self.off = np.concatenate([[0], np.cumsum(nsp * Np)])
        self.N = int(self.off[-1])
        rows, cols, vals = [], [], []

        def spos(c, s):
            return 1 if (s == 1 and (geom.spmask[c] & 1)) else 0

        def dofs(c, s):
            return self.off[c] + spos(c, s) * Np + np.arange(Np)

        for (t, P, Q, sP, sQ, d, side, r0, rn) in elems:
            loP, hiP = geom.box(P)
            hP = hiP - loP
            if rn > 0:
                X, W = nx[r0:r0 + rn, :D], nw[r0:r0 + rn]
            elif t == 0:
                X, W = _tensor_nodes(D, k, loP, hP, 0)
            else:
                X, W = _tensor_nodes(D, k, loP, hP, 1, d, upper=(t == 1 or side == 1))
            vP, gP = _phys_basis(D, k, loP, hP, X)
            if t == 0:
                M = -nu[sP] * np.einsum("q,qmd,qnd->mn", W, gP, gP)
                self._add(rows, cols, vals, dofs(P, sP), dofs(P, sP), M)
                continue
            if t == 3:
                nrm = nn[r0:r0 + rn, :D]
            else:
                nrm = np.zeros((len(W), D))
                nrm[:, d] = 1.0 if (t == 1 or side == 1) else -1.0
            dP = np.einsum("qmd,qd->qm", gP, nrm)
            if t == 2:
                # SIPLaplace boundary: nu (dn u v + dn v u) - nu * 2 pen(l) u v
                mu = 2.0 * pb / lengths[sP, P]
                nuP = nu[sP]
                M = nuP * (np.einsum("q,qn,qm->mn", W, dP, vP) + np.einsum("q,qm,qn->mn", W, dP, vP)) \
                    - nuP * mu * np.einsum("q,qn,qm->mn", W, vP, vP)
                self._add(rows, cols, vals, dofs(P, sP), dofs(P, sP), M)
                continue
            loQ, hiQ = geom.box(Q)
            vQ, gQ = _phys_basis(D, k, loQ, hiQ - loQ, X)
            dQ = np.einsum("qmd,qd->qm", gQ, nrm)
            if t == 1:
                nuP = nuQ = nu[sP]
                mu = pb * max(1.0 / lengths[sP, P], 1.0 / lengths[sP, Q]) * nu[sP]
            else:
                nuP, nuQ = nu[0], nu[1]
                mu = pb * max(1.0 / lengths[0, P], 1.0 / lengths[1, P]) * max(abs(nu[0]), abs(nu[1])) * \
                    (1.0 if nu[0] >= 0 else -1.0)
            # inner-edge SIP form: {nu grad u}.n [v] + {nu grad v}.n [u] - mu [u][v], [w] = w_P - w_Q
            for (a, va, da, nua, sa, ca, sga) in ((0, vP, dP, nuP, sP, P, 1.0), (1, vQ, dQ, nuQ, sQ, Q, -1.0)):      # test side
                for (b, vb, db, nub, sb, cb, sgb) in ((0, vP, dP, nuP, sP, P, 1.0), (1, vQ, dQ, nuQ, sQ, Q, -1.0)):  # trial
                    M = 0.5 * nub * sga * np.einsum("q,qn,qm->mn", W, db, va) \
                        + 0.5 * nua * sgb * np.einsum("q,qm,qn->mn", W, da, vb) \
                        - mu * sga * sgb * np.einsum("q,qn,qm->mn", W, vb, va)
                    self._add(rows, cols, vals, dofs(ca, sa), dofs(cb, sb), M)
        self.A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                               shape=(self.N, self.N))
        self.spos = spos
        self.dofs = dofs

    @staticmethod
    def _add(rows, cols, vals, r, c, M):
        ii, jj = np.meshgrid(r, c, indexing="ij")
        rows.append(ii.ravel())
        cols.append(jj.ravel())
        vals.append(M.ravel())

    # ------------------------------------------------------------ mass matrix
    def mass_blocks(self, elems, nx, nw):
        """XDG mass matrix (X/MassMatrixFactory.cs:35-120: per cell and species
        int_{K_j cap A_s} phi_m phi_n on the species' volume rule), as an (N, Np) array:
        row r of the full layout holds its block row.  Uncut cells use the Gauss k+1
        tensor rule (the identity up to rounding, the basis being orthonormal)."""
        D, k, Np = self.g.D, self.k, self.Np
        out = np.zeros((self.N, Np))
        for (t, P, Q, sP, sQ, d, side, r0, rn) in elems:
            if t != 0:
                continue
            lo, hi = self.g.box(P)
            h = hi - lo
            if rn > 0:
                X, W = nx[r0:r0 + rn, :D], nw[r0:r0 + rn]
            else:
                X, W = _tensor_nodes(D, k, lo, h, 0)
            v, _ = _phys_basis(D, k, lo, h, X)
            out[self.dofs(P, sP)] = np.einsum("q,qm,qn->mn", W, v, v)
        return out

    # ------------------------------------------------------------ agglomeration
    def extension(self, src, tgt):
        """E[m, n] = int_src phi^src_m phi^tgt_n (Gauss k+1, exact for the polynomial product)."""
        D, k = self.g.D, self.k
        lo, hi = self.g.box(src)
        X, W = _tensor_nodes(D, k, lo, hi - lo, 0)
        vs, _ = _phys_basis(D, k, lo, hi - lo, X)
        lt, ht = self.g.box(tgt)
        vt, _ = _phys_basis(D, k, lt, ht - lt, X)
        return np.einsum("q,qm,qn->mn", W, vs, vt)

    def injection(self, target):
        """Q: reduced DOFs (sources dropped) -> full DOFs; returns (Q, reduced offsets)."""
        g, Np = self.g, self.Np
        J = g.J
        keep = np.zeros((2, J), bool)
        for s in range(2):
            keep[s] = (((g.spmask >> s) & 1) == 1) & (target[s] < 0)
        nred = keep.sum(0)
        roff = np.concatenate([[0], np.cumsum(nred * Np)])

        def rpos(c, s):
            return int(keep[0, c]) if s == 1 else 0

        rows, cols, vals = [], [], []
        for c in range(J):
            for s in range(2):
                if not (g.spmask[c] >> s) & 1:
                    continue
                fr = self.dofs(c, s)
                t = target[s, c]
                if t < 0:
                    rc = roff[c] + rpos(c, s) * Np + np.arange(Np)
                    self._add(rows, cols, vals, fr, rc, np.eye(Np))
                else:
                    rc = roff[t] + rpos(t, s) * Np + np.arange(Np)
                    self._add(rows, cols, vals, fr, rc, self.extension(c, t))
        Q = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                          shape=(self.N, int(roff[-1])))
        return Q, roff

    # ------------------------------------------------------------ right-hand side
    def rhs(self, elems, nx, nw, f, nu, penalty_base, lengths, gfun):
        """F - affine: F_m = int f_s phi_m, boundary affine -nu g dphi/dn + nu mu g phi
        with mu = 2 pen(l) (SIPLaplace.BoundaryEdgeForm with u - g)."""
        g, D, k, Np = self.g, self.g.D, self.k, self.Np
        pb = penalty_base * self.penalty_deg
        b = np.zeros(self.N)
        for (t, P, Q, sP, sQ, d, side, r0, rn) in elems:
            if t not in (0, 2):
                continue
            lo, hi = g.box(P)
            h = hi - lo
            if rn > 0:
                X, W = nx[r0:r0 + rn, :D], nw[r0:r0 + rn]
            elif t == 0:
                X, W = _tensor_nodes(D, k, lo, h, 0)
            else:
                X, W = _tensor_nodes(D, k, lo, h, 1, d, upper=(side == 1))
            v, gr = _phys_basis(D, k, lo, h, X)
            if t == 0:
                b[self.dofs(P, sP)] += np.einsum("q,qm->m", W * f[sP], v)
            else:
                gv = gfun(X)
                dn = gr[:, :, d] * (1.0 if side == 1 else -1.0)
                mu = 2.0 * pb / lengths[sP, P]
                b[self.dofs(P, sP)] -= np.einsum("q,qm->m", W * gv, -nu[sP] * dn + nu[sP] * mu * v)
        return b

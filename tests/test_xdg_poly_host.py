"""Cut-cell rules of a polynomial level set (bh_xdg_create_poly, Saye's dimension
reduction with Bernstein bounds) on the host, pinned by closed forms:

* the sphere written as the polynomial |x - c|^2 - r^2: every cell's species
  volumes and interface area equal the oracle's closed forms (the same quantities
  the sphere path is held to), with the error falling as the rule order grows;
* a torus (degree 4 per variable, genus 1, its equator on a grid plane, where the
  surface is perpendicular to the cell faces): total inside volume 2 pi^2 R r^2 and
  total interface area 4 pi^2 R r to 1e-12 / 1e-10;
* a tilted plane: total volume below and total interface area exact (the rules
  integrate the linear height exactly);
* 2D: an ellipse (area pi a b) and the superellipse x^4 + y^4 = s^4
  (area 4 s^2 Gamma(5/4)^2 / Gamma(3/2));
* interface normals of the sphere equal (x - c) / r at every node.
"""
import math

import numpy as np
import pytest

import bosss_amd as B
from oracle import xdg as OX


def poly_from_terms(D, deg, terms):
    """coefficient array [i0, i1(, i2)] from {(i0, i1, i2): value}"""
    C = np.zeros((deg + 1,) * D)
    for idx, v in terms.items():
        C[idx[:D]] += v
    return C


def shifted_square_sum(D, c):
    """|x - c|^2 as a coefficient array of degree 2"""
    C = np.zeros((3,) * D)
    for d in range(D):
        e = [0] * D
        e[d] = 2
        C[tuple(e)] += 1.0
        e[d] = 1
        C[tuple(e)] -= 2 * c[d]
        C[(0,) * D] += c[d] ** 2
    return C


def polymul(A, Bc):
    """product of two tensor coefficient arrays (degrees add per variable)"""
    D = A.ndim
    out = np.zeros(tuple(A.shape[d] + Bc.shape[d] - 1 for d in range(D)))
    for ia in np.ndindex(A.shape):
        if A[ia] == 0:
            continue
        for ib in np.ndindex(Bc.shape):
            if Bc[ib] == 0:
                continue
            out[tuple(ia[d] + ib[d] for d in range(D))] += A[ia] * Bc[ib]
    return out


def pad(A, deg):
    out = np.zeros((deg + 1,) * A.ndim)
    out[tuple(slice(0, s) for s in A.shape)] = A
    return out


def build(D, n, C, q, k=2):
    x = np.linspace(0.0, 1.0, n + 1)
    g = B.Grid(None, *([x] * D))
    return x, B.XdgPoisson.poly(g, k, C, 1.0, 1000.0, quad_order=q)


@pytest.mark.parametrize("D,n,q,tol_v,tol_a", [(2, 10, 10, 1e-13, 1e-13), (3, 12, 10, 2e-9, 5e-7), (3, 12, 16, 1e-12, 5e-10)])
def test_sphere_polynomial_matches_closed_forms(D, n, q, tol_v, tol_a):
    r = 0.27 if D == 3 else 0.31
    c = [0.5, 0.47, 0.52][:D]
    C = shifted_square_sum(D, c)
    C[(0,) * D] -= r * r
    x, X = build(D, n, C, q)
    G = OX.Geometry([x] * D, c, r)
    cp = X.cells()
    h = 1.0 / n
    assert np.array_equal(cp["spmask"], G.spmask)
    assert np.abs(cp["vol"] - G.vol).max() <= tol_v * h ** D
    assert np.abs(cp["iface_area"] - G.iface).max() <= tol_a * h ** (D - 1)


def test_sphere_polynomial_normals():
    c, r = np.array([0.5, 0.47, 0.52]), 0.27
    C = shifted_square_sum(3, c)
    C[0, 0, 0] -= r * r
    _, X = build(3, 8, C, 6)
    e = X.elements()
    ints, nx, nn = e["ints"], e["nx"], e["nn"]
    sel = []
    for (t, P, Q, sP, sQ, d, side, r0, rn) in ints:
        if t == 3 and rn > 0:
            sel.extend(range(r0, r0 + rn))
    sel = np.array(sel)
    assert len(sel) > 100
    expect = (nx[sel] - c) / np.linalg.norm(nx[sel] - c, axis=1)[:, None]
    assert np.abs(nn[sel] - expect).max() <= 1e-12
    assert np.abs(np.linalg.norm(nx[sel] - c, axis=1) - r).max() <= 1e-12


def test_torus_volume_and_area():
    R, rr, c = 0.3, 0.1, [0.5, 0.5, 0.5]
    S = shifted_square_sum(3, c)                      # |x - c|^2
    S[0, 0, 0] += R * R - rr * rr                     # |x - c|^2 + R^2 - r^2
    P = shifted_square_sum(3, [c[0], c[1], 0.0])       # (x - cx)^2 + (y - cy)^2 ... plus z^2
    P[0, 0, 2] -= 1.0                                 # remove z^2
    C = polymul(S, S) - 4 * R * R * pad(P, 4)
    x, X = build(3, 16, C, 10)
    cp = X.cells()
    V, A = 2 * math.pi ** 2 * R * rr * rr, 4 * math.pi ** 2 * R * rr
    assert cp["vol"][0].sum() == pytest.approx(V, rel=1e-12)
    assert cp["iface_area"].sum() == pytest.approx(A, rel=1e-10)
    assert (cp["vol"].sum(0) == pytest.approx(np.full(X.J, 1.0 / 16 ** 3), rel=1e-12))


def test_tilted_plane_exact():
    a, b, cz = 0.3, 0.2, 0.1         # phi = z - (a + b x + cz y): A below the plane
    C = poly_from_terms(3, 1, {(0, 0, 1): 1.0, (0, 0, 0): -a, (1, 0, 0): -b, (0, 1, 0): -cz})
    x, X = build(3, 10, C, 4)
    cp = X.cells()
    assert cp["vol"][0].sum() == pytest.approx(a + b / 2 + cz / 2, rel=1e-13)
    assert cp["iface_area"].sum() == pytest.approx(math.sqrt(1 + b * b + cz * cz), rel=1e-13)


def test_2d_ellipse_and_superellipse():
    ea, eb, c = 0.3, 0.2, [0.5, 0.45]
    C = poly_from_terms(2, 2, {(2, 0): 1 / ea ** 2, (1, 0): -2 * c[0] / ea ** 2, (0, 2): 1 / eb ** 2,
                               (0, 1): -2 * c[1] / eb ** 2, (0, 0): c[0] ** 2 / ea ** 2 + c[1] ** 2 / eb ** 2 - 1.0})
    _, X = build(2, 20, C, 12)
    assert X.cells()["vol"][0].sum() == pytest.approx(math.pi * ea * eb, rel=1e-12)
    s = 0.3
    xq = np.zeros((5, 5))
    # (x - 1/2)^4 + (y - 1/2)^4 - s^4
    for i, v in enumerate([1 / 16, -1 / 2, 3 / 2, -2, 1]):
        xq[i, 0] += v
        xq[0, i] += v
    xq[0, 0] -= s ** 4
    _, X = build(2, 20, xq, 12)
    area = 4 * s * s * math.gamma(1.25) ** 2 / math.gamma(1.5)
    assert X.cells()["vol"][0].sum() == pytest.approx(area, rel=1e-10)

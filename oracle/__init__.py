"""CPU oracle for the DG operator / block-solver path (TEST INFRASTRUCTURE).

A plain FP64 restatement of the BoSSS algorithms, compiled from
``oracle.cpp`` into ``liboracle.so`` and loaded here with ctypes.  Only the
tests, the smoke check and the CPU-baseline leg of the benchmark use it, and
only as the checker; the product library never calls it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_d = ctypes.c_double
_i = ctypes.c_int
_i64 = ctypes.c_int64
_p = ctypes.c_void_p
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "orc_basis_count": (_i, [_i, _i]),
            "orc_basis_exponents": (None, [_i, _i, _ip]),
            "orc_basis_eval": (None, [_i, _i, _i, _dp, _dp, _p]),
            "orc_gauss": (None, [_i, _dp, _dp]),
            "orc_grid_create": (_p, [_i, _dp, _i, _dp, _i, _p, _i]),
            "orc_grid_free": (None, [_p]),
            "orc_grid_info": (None, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
            "orc_grid_edges": (None, [_p, _ip, _ip, _ip, _ip]),
            "orc_grid_cj": (None, [_p, _dp]),
            "orc_grid_cells": (None, [_p, _dp, _dp]),
            "orc_sip_create": (_p, [_p, _i, _d, _dp, _ip, _i]),
            "orc_sip_free": (None, [_p]),
            "orc_sip_set_chunk": (None, [_p, _i]),
            "orc_sip_num_bnd_nodes": (_i, [_p]),
            "orc_sip_bnd_nodes": (None, [_p, _dp]),
            "orc_sip_set_bnd_values": (None, [_p, _dp]),
            "orc_sip_set_rhs": (None, [_p, _p]),
            "orc_sip_num_vol_nodes": (_i, [_p]),
            "orc_sip_vol_nodes": (None, [_p, _dp, _dp]),
            "orc_sip_project": (None, [_p, _dp, _dp]),
            "orc_sip_evaluate": (None, [_p, _d, _dp, _d, _dp]),
            "orc_sip_assemble": (_p, [_p, _dp]),
            "orc_bsr_create": (_p, [_i, _i, _i64p, _ip, _dp]),
            "orc_bsr_free": (None, [_p]),
            "orc_bsr_info": (None, [_p, ctypes.POINTER(_i), ctypes.POINTER(_i64), ctypes.POINTER(_i)]),
            "orc_bsr_export": (None, [_p, _i64p, _ip, _dp]),
            "orc_bsr_spmv": (None, [_p, _d, _dp, _d, _dp]),
            "orc_bjacobi_solve": (_i, [_p, _d, _i, _dp, _dp]),
            "orc_soft_gmres": (_i, [_p, _i, _d, _i, _i, _d, _i, _dp, _dp,
                                    ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_d)]),
            "orc_pmg_vcycle": (_i, [_p, _p, _ip, _i, _i, _d, _i, _i, _dp, _dp, ctypes.POINTER(_i)]),
            "orc_pmg_create": (_p, [_p, _p, _ip, _i, _i, _d, _i, _i]),
            "orc_pmg_free": (None, [_p]),
            "orc_pmg_levels": (_i, [_p, _p]),
            "orc_pmg_vcycle_from": (None, [_p, _i, _dp, _dp]),
            "orc_pmg_level_spmv": (None, [_p, _i, _dp, _dp]),
            "orc_pmg_set": (None, [_p, _d, _i, _i]),
            "orc_pmg_set_schwarz": (_i, [_p, _i, _i, _i]),
            "orc_pmg_set_level_pmg": (_i, [_p, _i]),
            "orc_pmg_smooth": (None, [_p, _i, _p, _p]),
            "orc_pmg_level_info": (None, [_p, _i, _ip]),
            "orc_pmg_level_transfer": (_i, [_p, _i, _ip, _dp]),
            "orc_pmg_level_boxes": (_i, [_p, _i, _dp, _dp]),
            "orc_pmg_level_matrix": (_p, [_p, _i]),
            "orc_euler_create": (_p, [_p, _i, _i, _d, _d]),
            "orc_euler_free": (None, [_p]),
            "orc_euler_rhs": (None, [_p, _dp, _dp]),
            "orc_euler_tvd3": (None, [_p, _dp, _d, _i]),
            "orc_hllc": (None, [_i, _d, _d, _dp, _dp, _dp, _dp]),
            "orc_euler_flux": (None, [_i, _d, _d, _dp, _dp]),
            "orc_num_threads": (_i, []),
            "orc_set_threads": (None, [_i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def basis_count(D, k):
    return lib().orc_basis_count(D, k)


def basis_exponents(D, k):
    n = basis_count(D, k)
    out = np.zeros(n * D, np.int32)
    lib().orc_basis_exponents(D, k, out)
    return out.reshape(n, D)


def basis_eval(D, k, xi):
    xi = np.ascontiguousarray(xi, np.float64).reshape(-1, D)
    n = basis_count(D, k)
    val = np.zeros((xi.shape[0], n))
    grad = np.zeros((xi.shape[0], n, D))
    lib().orc_basis_eval(D, k, xi.shape[0], xi, val, grad.ctypes.data)
    return val, grad


def gauss(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().orc_gauss(n, x, w)
    return x, w


class Grid:
    """Cartesian grid with the reference's cell numbering."""

    def __init__(self, *nodes):
        self.D = len(nodes)
        ns = [np.ascontiguousarray(a, np.float64) for a in nodes]
        z = ns[2] if self.D == 3 else None
        self.h = lib().orc_grid_create(self.D, ns[0], len(ns[0]), ns[1], len(ns[1]),
                                       z.ctypes.data if z is not None else None,
                                       len(z) if z is not None else 0)
        self._keep = ns
        J, E = _i(), _i()
        lib().orc_grid_info(self.h, ctypes.byref(J), ctypes.byref(E))
        self.J, self.E = J.value, E.value

    def edges(self):
        a = [np.zeros(self.E, np.int32) for _ in range(4)]
        lib().orc_grid_edges(self.h, *a)
        return a

    def cj(self):
        c = np.zeros(self.J)
        lib().orc_grid_cj(self.h, c)
        return c

    def cells(self):
        lo = np.zeros(self.J * self.D)
        hh = np.zeros(self.J * self.D)
        lib().orc_grid_cells(self.h, lo, hh)
        return lo.reshape(self.J, self.D), hh.reshape(self.J, self.D)

    def __del__(self):
        try:
            lib().orc_grid_free(self.h)
        except Exception:
            pass


class Bsr:
    def __init__(self, handle=None, nrows=None, bs=None, rowptr=None, col=None, val=None):
        if handle is None:
            handle = lib().orc_bsr_create(nrows, bs, np.ascontiguousarray(rowptr, np.int64),
                                          np.ascontiguousarray(col, np.int32),
                                          np.ascontiguousarray(val, np.float64))
        self.h = handle
        n, nnzb, bs_ = _i(), _i64(), _i()
        lib().orc_bsr_info(self.h, ctypes.byref(n), ctypes.byref(nnzb), ctypes.byref(bs_))
        self.nrows, self.nnzb, self.bs = n.value, nnzb.value, bs_.value

    def export(self):
        rp = np.zeros(self.nrows + 1, np.int64)
        col = np.zeros(self.nnzb, np.int32)
        val = np.zeros(self.nnzb * self.bs * self.bs)
        lib().orc_bsr_export(self.h, rp, col, val)
        return rp, col, val.reshape(self.nnzb, self.bs, self.bs)

    def dense(self):
        rp, col, val = self.export()
        n = self.nrows * self.bs
        A = np.zeros((n, n))
        bs = self.bs
        for r in range(self.nrows):
            for p in range(rp[r], rp[r + 1]):
                c = col[p]
                A[r * bs:(r + 1) * bs, c * bs:(c + 1) * bs] = val[p]
        return A

    def spmv(self, alpha, x, beta, y):
        lib().orc_bsr_spmv(self.h, alpha, np.ascontiguousarray(x, np.float64), beta, y)
        return y

    def bjacobi(self, x, b, omega=1.0, iters=1):
        rc = lib().orc_bjacobi_solve(self.h, omega, iters, x, np.ascontiguousarray(b, np.float64))
        if rc != 0:
            raise ArithmeticException_("block-Jacobi init failed")
        return x

    def pmg_vcycle(self, grid, b, degrees, min_coarse=64, omega=1.0, pre=1, post=1):
        """one p/h multigrid V-cycle from a zero guess (see oracle.cpp pmg_build)."""
        x = np.zeros(self.nrows * self.bs)
        nl = _i()
        rc = lib().orc_pmg_vcycle(self.h, grid.h, np.ascontiguousarray(degrees, np.int32), len(degrees), min_coarse,
                                  omega, pre, post, np.ascontiguousarray(b, np.float64), x, ctypes.byref(nl))
        if rc != 0:
            raise ArithmeticException_("multigrid setup failed")
        return x, nl.value

    def gmres(self, b, x=None, precond=1, omega=1.0, bj_iters=1, restart=30, tol=1e-10, maxit=10000):
        if x is None:
            x = np.zeros(self.nrows * self.bs)
        it, r0, r = _i(), _d(), _d()
        rc = lib().orc_soft_gmres(self.h, precond, omega, bj_iters, restart, tol, maxit, x,
                                  np.ascontiguousarray(b, np.float64),
                                  ctypes.byref(it), ctypes.byref(r0), ctypes.byref(r))
        return x, dict(converged=(rc == 0), iters=it.value, res0=r0.value, res=r.value)

    def __del__(self):
        try:
            lib().orc_bsr_free(self.h)
        except Exception:
            pass


class ArithmeticException_(ArithmeticError):
    pass


class Sip:
    """SIP Laplace operator (div(nu grad u)) on a Cartesian grid."""

    def __init__(self, grid, k, nu, penalty_base=1.3, bc=None, nq1=0):
        self.grid = grid
        self.k = k
        self.D = grid.D
        self.Np = basis_count(grid.D, k)
        self.N = grid.J * self.Np
        nu = np.ascontiguousarray(np.broadcast_to(nu, (grid.J,)), np.float64)
        bc = np.ascontiguousarray(bc if bc is not None else np.zeros(2 * grid.D), np.int32)
        self.h = lib().orc_sip_create(grid.h, k, penalty_base, nu, bc, nq1)
        self._rhs = None

    def set_chunk(self, c):
        lib().orc_sip_set_chunk(self.h, c)

    def bnd_nodes(self):
        n = lib().orc_sip_num_bnd_nodes(self.h)
        X = np.zeros(n * self.D)
        lib().orc_sip_bnd_nodes(self.h, X)
        return X.reshape(n, self.D)

    def set_bnd_values(self, g):
        lib().orc_sip_set_bnd_values(self.h, np.ascontiguousarray(g, np.float64))

    def set_rhs(self, f):
        if f is None:
            lib().orc_sip_set_rhs(self.h, None)
            self._rhs = None
        else:
            self._rhs = np.ascontiguousarray(f, np.float64)
            lib().orc_sip_set_rhs(self.h, self._rhs.ctypes.data)

    def vol_nodes(self):
        nq = lib().orc_sip_num_vol_nodes(self.h)
        X = np.zeros(self.grid.J * nq * self.D)
        W = np.zeros(self.grid.J * nq)
        lib().orc_sip_vol_nodes(self.h, X, W)
        return X.reshape(self.grid.J, nq, self.D), W.reshape(self.grid.J, nq)

    def project(self, fvals):
        c = np.zeros(self.N)
        lib().orc_sip_project(self.h, np.ascontiguousarray(fvals, np.float64).ravel(), c)
        return c

    def evaluate(self, u, alpha=1.0, beta=0.0, y=None):
        if y is None:
            y = np.zeros(self.N)
        lib().orc_sip_evaluate(self.h, alpha, np.ascontiguousarray(u, np.float64), beta, y)
        return y

    def assemble(self):
        aff = np.zeros(self.N)
        h = lib().orc_sip_assemble(self.h, aff)
        return Bsr(handle=h), aff

    def __del__(self):
        try:
            lib().orc_sip_free(self.h)
        except Exception:
            pass


class Pmg:
    """Persistent oracle p/h multigrid hierarchy (see oracle.cpp pmg_build)."""

    def __init__(self, A, grid, degrees, min_coarse=64, omega=1.0, pre=1, post=1):
        self.h = lib().orc_pmg_create(A.h, grid.h, np.ascontiguousarray(degrees, np.int32), len(degrees), min_coarse,
                                      omega, pre, post)
        if not self.h:
            raise ArithmeticException_("multigrid setup failed")
        n = lib().orc_pmg_levels(self.h, None)
        sz = np.zeros(n, np.int32)
        lib().orc_pmg_levels(self.h, sz.ctypes.data)
        self.sizes = sz

    def vcycle(self, b, level=0):
        x = np.zeros(self.sizes[level])
        lib().orc_pmg_vcycle_from(self.h, level, np.ascontiguousarray(b, np.float64), x)
        return x

    def spmv(self, level, x):
        y = np.zeros(self.sizes[level])
        lib().orc_pmg_level_spmv(self.h, level, np.ascontiguousarray(x, np.float64), y)
        return y

    def set(self, omega, pre, post):
        lib().orc_pmg_set(self.h, omega, pre, post)

    def smooth(self, level, r):
        """x = omega M^-1 r with the level's smoother (block-Jacobi or aggregate Schwarz)."""
        x = np.zeros(self.sizes[level])
        lib().orc_pmg_smooth(self.h, level, np.ascontiguousarray(r, np.float64).ctypes.data, x.ctypes.data)
        return x

    def set_level_pmg(self, on=True):
        """LevelPmg on the finest level: low-order coarse correction, then the high-order
        modes of every cell with their diagonal sub-block (AS/LevelPmg.cs, UseDiagonalPmg)."""
        if lib().orc_pmg_set_level_pmg(self.h, 1 if on else 0) != 0:
            raise ArithmeticException_("LevelPmg set-up failed (needs a p-level above the next level)")

    def set_schwarz(self, on=True, overlap=0, scaling=False):
        """Additive Schwarz over the aggregates, grown by `overlap` (0-2) face-neighbour
        layers (with `scaling`, a DOF's block sum divided by its multiplicity), instead
        of block-Jacobi on aggregated levels."""
        if lib().orc_pmg_set_schwarz(self.h, 1 if on else 0, int(overlap), 1 if scaling else 0) != 0:
            raise ArithmeticException_("Schwarz block inversion failed")

    def level_info(self, level):
        """dict(nrows, bs, kind (0 truncation / 1 aggregation to the next level), deg, n)."""
        info = np.zeros(7, np.int32)
        lib().orc_pmg_level_info(self.h, level, info)
        return dict(nrows=int(info[0]), bs=int(info[1]), kind=int(info[2]), deg=int(info[3]),
                    n=[int(v) for v in info[4:7]])

    def level_transfer(self, level):
        """(parent aggregate per cell, blocks P[cell][fine mode][coarse mode]) of an aggregation level."""
        inf = self.level_info(level)
        par = np.zeros(inf["nrows"], np.int32)
        P = np.zeros((inf["nrows"], inf["bs"], inf["bs"]))
        if lib().orc_pmg_level_transfer(self.h, level, par, P.reshape(-1)) != 0:
            raise ValueError("level %d is not an aggregation level" % level)
        return par, P

    def level_boxes(self, level, D):
        """(lower corners, widths) of the level's cells, J x D each."""
        inf = self.level_info(level)
        lo = np.zeros((inf["nrows"], D))
        h = np.zeros((inf["nrows"], D))
        if lib().orc_pmg_level_boxes(self.h, level, lo.reshape(-1), h.reshape(-1)) != 0:
            raise ValueError("level %d carries no cell boxes" % level)
        return lo, h

    def level_matrix(self, level):
        """the level matrix as a free-standing oracle Bsr."""
        return Bsr(handle=lib().orc_pmg_level_matrix(self.h, level))

    def __del__(self):
        try:
            lib().orc_pmg_free(self.h)
        except Exception:
            pass


class Euler:
    """Explicit DG for the compressible Euler equations on a periodic Cartesian
    grid with the HLLC flux (see oracle.cpp, Euler section)."""

    def __init__(self, grid, k=2, nq1=4, gamma=1.4, mach=None):
        self.grid = grid
        self.k = k
        self.D = grid.D
        self.Np = basis_count(grid.D, k)
        self.nv = grid.D + 2
        self.N = grid.J * self.nv * self.Np
        self.gamma = gamma
        self.mach = mach if mach is not None else 1.0 / np.sqrt(gamma)
        self.h = lib().orc_euler_create(grid.h, k, nq1, gamma, self.mach)

    def rhs(self, U):
        R = np.zeros(self.N)
        lib().orc_euler_rhs(self.h, np.ascontiguousarray(U, np.float64), R)
        return R

    def tvd3(self, U, dt, steps):
        U = np.ascontiguousarray(U, np.float64).copy()
        lib().orc_euler_tvd3(self.h, U, dt, steps)
        return U

    def __del__(self):
        try:
            lib().orc_euler_free(self.h)
        except Exception:
            pass


def hllc(D, gamma, mach, Ui, Uo, n):
    F = np.zeros(D + 2)
    lib().orc_hllc(D, gamma, mach, np.ascontiguousarray(Ui, np.float64), np.ascontiguousarray(Uo, np.float64),
                   np.ascontiguousarray(n, np.float64), F)
    return F


def euler_flux(D, gamma, mach, U):
    F = np.zeros((D + 2) * D)
    lib().orc_euler_flux(D, gamma, mach, np.ascontiguousarray(U, np.float64), F)
    return F.reshape(D + 2, D)

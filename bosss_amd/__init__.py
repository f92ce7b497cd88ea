"""bosss_amd -- MI355X backend for the BoSSS DG-operator / block-solver path.

Python mirror of the reference's host interfaces over the C ABI of
``libbosss_hip.so`` (declared in ``include/bosss_hip.h``).  Names follow the
reference so parity tests read like its own tests:

* ``Grid2D.Cartesian2DGrid`` / ``Grid3D.Cartesian3DGrid``
  (BoSSS.Foundation.Grid/Grid2D.cs:396, Grid3D.cs:88)
* ``SipOperator.Evaluate(alpha, beta, output)``  (IEvaluatorNonLin.Evaluate,
  F/DifferentialOperator.cs:1601) and ``ComputeMatrix``
  (F/DifferentialOperator.cs:1787)
* ``BlockMsrMatrix.SpMV(alpha, x, beta, y)``  (P/BlockMsrMatrix.cs:2970)
* ``BlockJacobi``, ``SoftGMRES``, ``FlexGMRES`` with ``Solve(X, B)``
  (AS/BlockJacobi.cs, AS/SoftGMRES.cs, AS/FlexGMRES.cs)

There is no CPU fallback: every call goes to the HIP library, and importing
this package raises if the library has not been built.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbosss_hip.so")

BH_DEVICE_PTR = 1
BH_FP64, BH_FP32, BH_FP16_MFMA = 0, 1, 2

_c_d = ctypes.c_double
_c_i = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_p = ctypes.c_void_p
_c_f = ctypes.c_float
_PP = ctypes.POINTER(ctypes.c_void_p)

# exported symbol -> (restype, argtypes); every declaration of include/bosss_hip.h
SIGNATURES = {
    "bh_init": (_c_i, [_c_i, _PP]),
    "bh_finalize": (_c_i, [_c_p]),
    "bh_last_error": (_c_i, [ctypes.c_char_p, _c_i]),
    "bh_sync": (_c_i, [_c_p]),
    "bh_malloc": (_c_i, [_c_p, ctypes.c_size_t, _PP]),
    "bh_free": (_c_i, [_c_p, _c_p]),
    "bh_memcpy_h2d": (_c_i, [_c_p, _c_p, _c_p, ctypes.c_size_t]),
    "bh_memcpy_d2h": (_c_i, [_c_p, _c_p, _c_p, ctypes.c_size_t]),
    "bh_dot": (_c_i, [_c_p, _c_i64, _c_p, _c_p, ctypes.POINTER(_c_d)]),
    "bh_memcpy_d2d": (_c_i, [_c_p, _c_p, _c_p, ctypes.c_size_t]),
    "bh_event_record": (_c_i, [_c_p, _c_i]),
    "bh_event_elapsed": (_c_i, [_c_p, _c_i, _c_i, ctypes.POINTER(_c_f)]),
    "bh_grid_cartesian": (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_p, _c_i, _c_p, _c_i, _PP]),
    "bh_grid_destroy": (_c_i, [_c_p]),
    "bh_grid_info": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_i)]),
    "bh_grid_edges": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    "bh_grid_cj": (_c_i, [_c_p, _c_p]),
    "bh_sip_create": (_c_i, [_c_p, _c_i, _c_d, _c_p, _c_p, _PP]),
    "bh_op_destroy": (_c_i, [_c_p]),
    "bh_op_info": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i64)]),
    "bh_sip_num_bnd_nodes": (_c_i, [_c_p, ctypes.POINTER(_c_i)]),
    "bh_sip_bnd_nodes": (_c_i, [_c_p, _c_p]),
    "bh_sip_set_bnd_values": (_c_i, [_c_p, _c_p]),
    "bh_sip_set_rhs": (_c_i, [_c_p, _c_p, _c_i]),
    "bh_op_evaluate": (_c_i, [_c_p, _c_d, _c_p, _c_d, _c_p, _c_i, _c_i, _c_i]),
    "bh_op_assemble": (_c_i, [_c_p, _PP, _c_p, _c_i]),
    "bh_op_tuning": (_c_i, [_c_p, _c_p, _c_p]),
    "bh_op_tuning_fp32": (_c_i, [_c_p, _c_p, _c_p, _c_p]),
    "bh_op_set_fp64_path": (_c_i, [_c_p, _c_i, _c_i, _c_i]),
    "bh_op_set_fp32_path": (_c_i, [_c_p, _c_i, _c_i]),
    "bh_bsr_create": (_c_i, [_c_p, _c_i, _c_i, _c_p, _c_p, _c_p, _PP]),
    "bh_bsr_create_v": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p, _PP]),
    "bh_bsr_destroy": (_c_i, [_c_p]),
    "bh_bsr_info": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_i64)]),
    "bh_bsr_export": (_c_i, [_c_p, _c_p, _c_p, _c_p]),
    "bh_bsr_spmv": (_c_i, [_c_p, _c_d, _c_p, _c_d, _c_p, _c_i, _c_i]),
    "bh_bjacobi_create": (_c_i, [_c_p, _c_d, _c_i, _c_i, _PP]),
    "bh_gmres_create": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_d, _c_i, _PP]),
    "bh_fgmres_create": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_d, _c_i, _PP]),
    "bh_pmg_create": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i, _c_d, _c_i, _c_i, _c_i, _c_i, _c_i, _PP]),
    "bh_mlp_create": (_c_i, [_c_p, _c_i, _c_i, _c_i, ctypes.c_uint64, _PP]),
    "bh_mlp_destroy": (_c_i, [_c_p]),
    "bh_mlp_info": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_i)]),
    "bh_mlp_get_params": (_c_i, [_c_p, _c_p]),
    "bh_mlp_set_params": (_c_i, [_c_p, _c_p]),
    "bh_mlp_forward": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_i]),
    "bh_mlp_gradient": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, ctypes.POINTER(_c_d), _c_i]),
    "bh_mlp_gradient_noalloc": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_p]),
    "bh_mlp_adam_step": (_c_i, [_c_p, _c_f]),
    "bh_mlp_features": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i, _c_i]),
    "bh_mlp_backward": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_i]),
    "bh_mlp_correction": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i, _c_i]),
    "bh_mlp_upstream": (_c_i, [_c_p, _c_p, _c_p, _c_d, _c_p, _c_i, _c_i]),
    "bh_mlp_targets": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i, _c_i]),
    "bh_pmg_set_learned": (_c_i, [_c_p, _c_p]),
    "bh_orthomg_create": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i, _c_i, _c_i, _c_d, _c_i, _PP]),
    "bh_orthomg_set_learned": (_c_i, [_c_p, _c_p]),
    "bh_orthomg_info": (_c_i, [_c_p, _c_p]),
    "bh_pmg_learned_stats": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i)]),
    "bh_pmg_set_graphs": (_c_i, [_c_p, _c_i]),
    "bh_pmg_set_schwarz": (_c_i, [_c_p, _c_i]),
    "bh_pmg_set_level_pmg": (_c_i, [_c_p, _c_i]),
    "bh_pmg_set_cell_ilu": (_c_i, [_c_p, _c_i]),
    "bh_pmg_smooth": (_c_i, [_c_p, _c_i, _c_p, _c_p]),
    "bh_pmg_level_info": (_c_i, [_c_p, _c_i, _c_p, _c_p]),
    "bh_pmg_level_transfer": (_c_i, [_c_p, _c_i, _c_p, _c_p]),
    "bh_pmg_level_matrix": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p]),
    "bh_pmg_set_schwarz_overlap": (_c_i, [_c_p, _c_i, _c_i, _c_i]),
    "bh_pmg_set_diag_equilib": (_c_i, [_c_p]),
    "bh_xdg_precond_set_fp64_blocks": (_c_i, [_c_p, _c_i]),
    "bh_xdg_precond_coarse_export": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "bh_pmg_set_learned_fallback": (_c_i, [_c_p, _c_i, _c_d]),
    "bh_pmg_learned_info": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_d)]),
    "bh_pmg_set_learned_cost_rule": (_c_i, [_c_p, _c_i]),
    "bh_pmg_learned_costs": (_c_i, [_c_p, _c_p]),
    "bh_euler_create": (_c_i, [_c_p, _c_i, _c_i, _c_d, _c_d, _PP]),
    "bh_euler_destroy": (_c_i, [_c_p]),
    "bh_euler_info": (_c_i, [_c_p, ctypes.POINTER(_c_i64), ctypes.POINTER(_c_i)]),
    "bh_euler_rhs": (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_i]),
    "bh_euler_implicit_step": (_c_i, [_c_p, _c_p, _c_p, _c_d, _c_d, _c_i, _c_i, _c_p, _c_p, _c_i]),
    "bh_euler_tvd3": (_c_i, [_c_p, _c_p, _c_d, _c_i, _c_i, _c_i]),
    "bh_xdg_create": (_c_i, [_c_p, _c_i, _c_p, _c_d, _c_d, _c_d, _c_d, _c_d, _c_i, _PP]),
    "bh_xdg_destroy": (_c_i, [_c_p]),
    "bh_xdg_info": (_c_i, [_c_p, _c_p]),
    "bh_xdg_cells": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    "bh_xdg_agglomeration": (_c_i, [_c_p, _c_p]),
    "bh_xdg_elements": (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "bh_xdg_matrix": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_p, _c_p]),
    "bh_xdg_bnd_nodes": (_c_i, [_c_p, _c_p]),
    "bh_xdg_set_rhs": (_c_i, [_c_p, _c_p, _c_p]),
    "bh_xdg_rhs": (_c_i, [_c_p, _c_i, _c_p, _c_i]),
    "bh_xdg_spmv": (_c_i, [_c_p, _c_i, _c_d, _c_p, _c_d, _c_p, _c_i, _c_i]),
    "bh_xdg_extend": (_c_i, [_c_p, _c_p, _c_p]),
    "bh_xdg_evaluate": (_c_i, [_c_p, _c_d, _c_p, _c_d, _c_p, _c_i]),
    "bh_xdg_create_poly": (_c_i, [_c_p, _c_i, _c_i, _c_p, _c_d, _c_d, _c_d, _c_d, _c_i, _PP]),
    "bh_xdg_cut_rule_fallbacks": (_c_i, [_c_p]),
    "bh_xdg_precond_create": (_c_i, [_c_p, _c_i, _c_i, _c_d, _c_i, _c_i, _c_i, _c_i, _PP]),
    "bh_xdg_precond_create_v": (_c_i, [_c_p, _c_p, _c_i, _c_i, _c_i, _c_d, _c_i, _c_i, _c_i, _c_i, _PP]),
    "bh_xdg_dof_frame": (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p]),
    "bh_xdg_mass_matrix": (_c_i, [_c_p, _c_p]),
    "bh_xdg_fgmres_create": (_c_i, [_c_p, _c_p, _c_i, _c_d, _c_i, _PP]),
    "bh_solver_solve": (_c_i, [_c_p, _c_p, _c_p, _c_i]),
    "bh_solver_stats": (_c_i, [_c_p, ctypes.POINTER(_c_i), ctypes.POINTER(_c_i), ctypes.POINTER(_c_d),
                               ctypes.POINTER(_c_d)]),
    "bh_solver_destroy": (_c_i, [_c_p]),
}

_LIB = None


class BoSSSHipError(RuntimeError):
    """A non-zero status of the native library (message from bh_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """Load libbosss_hip.so (raises if it is missing: there is no fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run __graft_entry__.build() or make -C bosss_amd/csrc")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(2048)
        lib().bh_last_error(buf, 2048)
        msg = buf.value.decode(errors="replace")
        if rc == -5:
            raise ArithmeticError(msg)
        raise BoSSSHipError(rc, msg)


def _ptr(a):
    return a.ctypes.data if a is not None else None


class DeviceArray:
    """A device (HBM) buffer owned by a Context: the device-pointer path of the C ABI."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = ctypes.c_void_p()
        _check(lib().bh_malloc(ctx.h, self.n * self.dtype.itemsize, ctypes.byref(p)))
        self.ptr = p.value
        ctx._children.append(weakref.ref(self))

    @classmethod
    def from_host(cls, ctx, a, dtype=None):
        a = np.ascontiguousarray(a, dtype or a.dtype)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size == self.n
        _check(lib().bh_memcpy_h2d(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes))

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().bh_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def copy_from(self, other):
        """device-to-device copy of a same-sized array (bh_memcpy_d2d)."""
        assert other.n == self.n and other.dtype == self.dtype
        _check(lib().bh_memcpy_d2d(self.ctx.h, self.ptr, other.ptr, self.n * self.dtype.itemsize))
        return self

    def view(self, offset, n):
        """non-owning view of elements [offset, offset+n) (device-pointer argument)."""
        assert 0 <= offset and offset + n <= self.n
        return _DeviceView(self, offset, n)

    def free(self):
        if self.ptr and self.ctx.h:
            _check(lib().bh_free(self.ctx.h, self.ptr))
        self.ptr = None

    destroy = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _DeviceView:
    def __init__(self, base, offset, n):
        self.base = base
        self.ctx = base.ctx
        self.dtype = base.dtype
        self.n = n
        self.ptr = base.ptr + offset * base.dtype.itemsize

    def download(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().bh_memcpy_d2h(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes))
        return out


def _vec_args(x):
    """(pointer, flags) of a host numpy array or a DeviceArray."""
    if isinstance(x, (DeviceArray, _DeviceView)):
        return x.ptr, BH_DEVICE_PTR
    return x.ctypes.data, 0


class _Owned:
    """A native handle owned by a Context.  The context destroys every live
    child before it finalises, so no destructor can run against a freed
    context (native objects synchronise on the context's stream)."""

    _destroy_fn = None

    def _register(self, ctx, handle):
        self.h = handle
        self._ctx_ref = ctx
        ctx._children.append(weakref.ref(self))

    def destroy(self):
        h = getattr(self, "h", None)
        if h and (self._ctx_ref.h or self._ctx_ref is HOST):
            getattr(lib(), self._destroy_fn)(h)
        self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Context:
    """One GPU, one stream (bh_init)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        _check(lib().bh_init(device, ctypes.byref(h)))
        self.h = h.value
        self._children = []

    def sync(self):
        _check(lib().bh_sync(self.h))

    def record(self, slot):
        _check(lib().bh_event_record(self.h, slot))

    def elapsed_ms(self, s0, s1):
        ms = ctypes.c_float()
        _check(lib().bh_event_elapsed(self.h, s0, s1, ctypes.byref(ms)))
        return ms.value

    def close(self):
        if self.h:
            for r in reversed(self._children):
                o = r()
                if o is not None:
                    o.destroy()
            self._children = []
            _check(lib().bh_finalize(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _HostOnly:
    """Stand-in context of host-only objects (grid + XDG set-up without a GPU)."""

    h = None

    def __init__(self):
        self._children = []


HOST = _HostOnly()


class Grid(_Owned):
    """Cartesian grid (reference numbering; see bh_grid_cartesian).  With
    ``ctx=None`` the grid is host-only: it serves the XDG set-up (cut-cell
    rules, agglomeration) and nothing that runs on the GPU."""

    _destroy_fn = "bh_grid_destroy"

    def __init__(self, ctx, *nodes):
        ctx = HOST if ctx is None else ctx
        self.ctx = ctx
        self.D = len(nodes)
        ns = [np.ascontiguousarray(a, np.float64) for a in nodes] + [None] * (3 - len(nodes))
        h = ctypes.c_void_p()
        _check(lib().bh_grid_cartesian(ctx.h, self.D, _ptr(ns[0]), len(ns[0]), _ptr(ns[1]), len(ns[1]),
                                       _ptr(ns[2]), len(ns[2]) if ns[2] is not None else 0, ctypes.byref(h)))
        self._register(ctx, h.value)
        D, J, E = _c_i(), _c_i(), _c_i()
        _check(lib().bh_grid_info(self.h, ctypes.byref(D), ctypes.byref(J), ctypes.byref(E)))
        self.J, self.E = J.value, E.value

    def edges(self):
        a = [np.zeros(self.E, np.int32) for _ in range(4)]
        _check(lib().bh_grid_edges(self.h, *[x.ctypes.data for x in a]))
        return a

    def cj(self):
        c = np.zeros(self.J)
        _check(lib().bh_grid_cj(self.h, c.ctypes.data))
        return c


class Grid2D:
    @staticmethod
    def Cartesian2DGrid(ctx, xNodes, yNodes):
        return Grid(ctx, xNodes, yNodes)


class Grid3D:
    @staticmethod
    def Cartesian3DGrid(ctx, xNodes, yNodes, zNodes):
        return Grid(ctx, xNodes, yNodes, zNodes)


class BlockMsrMatrix(_Owned):
    """Device block-sparse matrix (bh_bsr)."""

    _destroy_fn = "bh_bsr_destroy"

    def __init__(self, ctx, handle=None, nrows=None, bs=None, rowptr=None, col=None, vals=None):
        self.ctx = ctx
        if handle is None:
            h = ctypes.c_void_p()
            rp = np.ascontiguousarray(rowptr, np.int64)
            cl = np.ascontiguousarray(col, np.int32)
            vv = np.ascontiguousarray(vals, np.float64)
            _check(lib().bh_bsr_create(ctx.h, nrows, bs, rp.ctypes.data, cl.ctypes.data, vv.ctypes.data,
                                       ctypes.byref(h)))
            handle = h.value
        self._register(ctx, handle)
        n, bs_, nnzb = _c_i(), _c_i(), _c_i64()
        _check(lib().bh_bsr_info(self.h, ctypes.byref(n), ctypes.byref(bs_), ctypes.byref(nnzb)))
        self.nrows, self.bs, self.nnzb = n.value, bs_.value, nnzb.value
        self.N = self.nrows * self.bs

    @classmethod
    def variable(cls, ctx, rowptr, col, blk_i0, blk_len, vals):
        """Variable block sizes (bh_bsr_create_v): block row r of length blk_len[r]
        at offset blk_i0[r] of the vector layout; vals = every stored block,
        row-major, in rowptr order.  Vectors have max(blk_i0 + blk_len) entries."""
        rp = np.ascontiguousarray(rowptr, np.int64)
        cl = np.ascontiguousarray(col, np.int32)
        i0 = np.ascontiguousarray(blk_i0, np.int64)
        ln = np.ascontiguousarray(blk_len, np.int32)
        vv = np.ascontiguousarray(vals, np.float64)
        h = ctypes.c_void_p()
        _check(lib().bh_bsr_create_v(ctx.h, len(ln), rp.ctypes.data, cl.ctypes.data, i0.ctypes.data, ln.ctypes.data,
                                     vv.ctypes.data, ctypes.byref(h)))
        M = cls(ctx, handle=h.value)
        M.N = int((i0 + ln).max()) if len(ln) else 0
        return M

    def export(self):
        rp = np.zeros(self.nrows + 1, np.int64)
        col = np.zeros(self.nnzb, np.int32)
        val = np.zeros(self.nnzb * self.bs * self.bs)
        _check(lib().bh_bsr_export(self.h, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
        return rp, col, val.reshape(self.nnzb, self.bs, self.bs)

    def SpMV(self, alpha, x, beta, y, prec=BH_FP64):
        xp, fx = _vec_args(x)
        yp, fy = _vec_args(y)
        assert fx == fy, "x and y must both be host or both be device arrays"
        _check(lib().bh_bsr_spmv(self.h, alpha, xp, beta, yp, prec, fx))
        return y


class SipOperator(_Owned):
    """DifferentialOperator with one SIPLaplace component + RHS source."""

    _destroy_fn = "bh_op_destroy"

    def __init__(self, grid, k, nu, penalty_base=1.3, bc=None):
        self.grid = grid
        self.ctx = grid.ctx
        nu = np.ascontiguousarray(np.broadcast_to(nu, (grid.J,)), np.float64)
        self.nu = nu
        bc = np.ascontiguousarray(bc if bc is not None else np.zeros(2 * grid.D), np.int32)
        h = ctypes.c_void_p()
        _check(lib().bh_sip_create(grid.h, k, penalty_base, nu.ctypes.data, bc.ctypes.data, ctypes.byref(h)))
        self._register(self.ctx, h.value)
        self.k = k
        Np, N = _c_i(), _c_i64()
        _check(lib().bh_op_info(self.h, ctypes.byref(Np), ctypes.byref(N)))
        self.Np, self.N = Np.value, N.value

    def bnd_nodes(self):
        n = _c_i()
        _check(lib().bh_sip_num_bnd_nodes(self.h, ctypes.byref(n)))
        X = np.zeros(n.value * self.grid.D)
        _check(lib().bh_sip_bnd_nodes(self.h, X.ctypes.data))
        return X.reshape(n.value, self.grid.D)

    def set_bnd_values(self, g):
        g = np.ascontiguousarray(g, np.float64)
        _check(lib().bh_sip_set_bnd_values(self.h, g.ctypes.data))

    def set_rhs(self, f):
        if f is None:
            _check(lib().bh_sip_set_rhs(self.h, None, 0))
        else:
            p, fl = _vec_args(f if isinstance(f, DeviceArray) else np.ascontiguousarray(f, np.float64))
            _check(lib().bh_sip_set_rhs(self.h, p, fl))

    def Evaluate(self, alpha, beta, output, u, affine=True, prec=BH_FP64):
        """output <- beta*output + alpha*Op(u)  (IEvaluatorNonLin.Evaluate)."""
        up, fu = _vec_args(u)
        yp, fy = _vec_args(output)
        assert fu == fy
        _check(lib().bh_op_evaluate(self.h, alpha, up, beta, yp, 1 if affine else 0, prec, fu))
        return output

    def mfma_tuning(self):
        """(waves per workgroup, workgroups per XCD, ms) picked by the autotuner of the MFMA apply."""
        cfg = np.zeros(5, np.int32)
        ms = (_c_f * 3)()
        _check(lib().bh_op_tuning(self.h, cfg.ctypes.data, ms))
        return int(cfg[0]), int(cfg[1]), ms[0]

    def fp64_tuning(self):
        """FP64 apply choice: ('vector' | 'modal' | 'mfma' | None, waves, workgroups per XCD, vector ms, MFMA ms)."""
        cfg = np.zeros(5, np.int32)
        ms = (_c_f * 3)()
        _check(lib().bh_op_tuning(self.h, cfg.ctypes.data, ms))
        path = {0: None, 1: "vector", 2: "mfma", 3: "modal", 4: "tiled"}[int(cfg[2])]
        return path, int(cfg[3]), int(cfg[4]), ms[1], ms[2]

    def fp32_tuning(self):
        """FP32 apply choice: ('vector' | 'modal' | 'tiled' | None, waves per workgroup, ms)."""
        p, w, ms = _c_i(), _c_i(), _c_f()
        _check(lib().bh_op_tuning_fp32(self.h, ctypes.byref(p), ctypes.byref(w), ctypes.byref(ms)))
        return {0: None, 1: "vector", 3: "modal", 4: "tiled"}[p.value], w.value, ms.value

    def force_fp64_path(self, path, waves=8, wg_per_xcd=64):
        """Pin the FP64 apply to 'vector' (face rule node by node), 'modal' (the vector
        kernel with modal face terms), 'tiled' (modal, traces exchanged; waves = 2, 4
        or 8 per workgroup) or 'mfma'; tests compare each against the oracle."""
        code = {"vector": 1, "mfma": 2, "modal": 3, "tiled": 4}[path]
        _check(lib().bh_op_set_fp64_path(self.h, code, int(waves), int(wg_per_xcd)))

    def force_fp32_path(self, path, waves=4):
        """Pin the FP32 apply to 'vector', 'modal' or 'tiled' (waves = 2, 4 or 8)."""
        code = {"vector": 1, "modal": 3, "tiled": 4}[path]
        _check(lib().bh_op_set_fp32_path(self.h, code, int(waves)))

    def ComputeMatrix(self):
        """Assemble on the device -> (BlockMsrMatrix, affine vector)."""
        h = ctypes.c_void_p()
        aff = np.zeros(self.N)
        _check(lib().bh_op_assemble(self.h, ctypes.byref(h), aff.ctypes.data, 0))
        return BlockMsrMatrix(self.ctx, handle=h.value), aff


class _Solver(_Owned):
    _destroy_fn = "bh_solver_destroy"

    def __init__(self, ctx, handle, N, keep=()):
        self._register(ctx, handle)
        self.N = N
        self._keep = keep

    def Solve(self, X, B):
        xp, fx = _vec_args(X)
        bp, fb = _vec_args(B)
        assert fx == fb
        _check(lib().bh_solver_solve(self.h, xp, bp, fx))
        return X

    def stats(self):
        it, cv, r0, r = _c_i(), _c_i(), _c_d(), _c_d()
        _check(lib().bh_solver_stats(self.h, ctypes.byref(it), ctypes.byref(cv), ctypes.byref(r0), ctypes.byref(r)))
        return dict(iterations=it.value, converged=bool(cv.value), res0=r0.value, res=r.value)

    @property
    def ThisLevelIterations(self):
        return self.stats()["iterations"]

    @property
    def Converged(self):
        return self.stats()["converged"]


class BlockJacobi(_Solver):
    def __init__(self, M, omega=1.0, NoOfIterations=1, prec=BH_FP64):
        h = ctypes.c_void_p()
        _check(lib().bh_bjacobi_create(M.h, omega, NoOfIterations, prec, ctypes.byref(h)))
        super().__init__(M.ctx, h.value, M.N, (M,))


class SoftGMRES(_Solver):
    def __init__(self, M=None, op=None, Precond=None, MaxKrylovDim=80, ConvergenceCriterion=1e-10,
                 MaxSolverIterations=10000):
        h = ctypes.c_void_p()
        _check(lib().bh_gmres_create(M.h if M else None, op.h if op else None, Precond.h if Precond else None,
                                     MaxKrylovDim, ConvergenceCriterion, MaxSolverIterations, ctypes.byref(h)))
        super().__init__((M or op).ctx, h.value, (M or op).N, (M, op, Precond))


class FlexGMRES(_Solver):
    def __init__(self, M=None, op=None, Precond=None, MaxKrylovDim=30, ConvergenceCriterion=1e-8,
                 MaxSolverIterations=10000):
        h = ctypes.c_void_p()
        _check(lib().bh_fgmres_create(M.h if M else None, op.h if op else None, Precond.h if Precond else None,
                                      MaxKrylovDim, ConvergenceCriterion, MaxSolverIterations, ctypes.byref(h)))
        super().__init__((M or op).ctx, h.value, (M or op).N, (M, op, Precond))


class PMultigrid(_Solver):
    """p/h multigrid V-cycle preconditioner (bh_pmg_create); the p-multigrid of
    AS/PmgConfig.cs on the MultigridOperator hierarchy."""

    def __init__(self, op, M=None, degrees=None, min_coarse=64, omega=0.8, pre=2, post=2, kcycle_inner=2,
                 prec=BH_FP64, matrix_free=True):
        degrees = np.ascontiguousarray(degrees if degrees is not None else [op.k, 1], np.int32)
        h = ctypes.c_void_p()
        _check(lib().bh_pmg_create(op.h, M.h if M else None, degrees.ctypes.data, len(degrees), min_coarse, omega,
                                   pre, post, kcycle_inner, prec, 1 if matrix_free else 0, ctypes.byref(h)))
        super().__init__(op.ctx, h.value, op.N, (op, M))



class CellMLP(_Owned):
    """Per-cell correction network (bh_mlp_*): in -> 128 -> 128 -> out, ReLU,
    BF16 weights / FP32 accumulation (MFMA) for inference, FP32 training."""

    _destroy_fn = "bh_mlp_destroy"

    def __init__(self, ctx, n_in, n_out, seed=0, n_hidden=128):
        h = ctypes.c_void_p()
        _check(lib().bh_mlp_create(ctx.h, n_in, n_hidden, n_out, seed, ctypes.byref(h)))
        self._register(ctx, h.value)
        self.ctx = ctx
        self.n_in, self.n_out, self.n_hidden = n_in, n_out, n_hidden
        n, ip, op_ = _c_i(), _c_i(), _c_i()
        _check(lib().bh_mlp_info(self.h, ctypes.byref(n), ctypes.byref(ip), ctypes.byref(op_)))
        self.n_params, self.in_p, self.out_p = n.value, ip.value, op_.value

    def get_params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(lib().bh_mlp_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        _check(lib().bh_mlp_set_params(self.h, p.ctypes.data))

    def forward(self, X, Y=None):
        if isinstance(X, (DeviceArray, _DeviceView)):
            n = X.n // self.in_p
            _check(lib().bh_mlp_forward(self.h, n, X.ptr, Y.ptr, BH_DEVICE_PTR))
            return Y
        X = np.ascontiguousarray(X, np.float32)
        n = X.shape[0]
        Y = np.zeros((n, self.out_p), np.float32)
        _check(lib().bh_mlp_forward(self.h, n, X.ctypes.data, Y.ctypes.data, 0))
        return Y

    def gradient(self, X, T, want_grad=True):
        loss = _c_d()
        g = np.zeros(self.n_params, np.float32) if want_grad else None
        if isinstance(X, (DeviceArray, _DeviceView)):
            n = X.n // self.in_p
            _check(lib().bh_mlp_gradient(self.h, n, X.ptr, T.ptr, _ptr(g), ctypes.byref(loss), BH_DEVICE_PTR))
        else:
            X = np.ascontiguousarray(X, np.float32)
            T = np.ascontiguousarray(T, np.float32)
            _check(lib().bh_mlp_gradient(self.h, X.shape[0], X.ctypes.data, T.ctypes.data, _ptr(g), ctypes.byref(loss), 0))
        return loss.value, g

    def gradient_noalloc(self, X, T, grad=None, loss=None):
        """bh_mlp_gradient_noalloc: device arrays only; grad (n_params FP32) and
        loss (one FP64) are optional device outputs.  Stream-ordered, no sync."""
        n = X.n // self.in_p
        _check(lib().bh_mlp_gradient_noalloc(self.h, n, X.ptr, T.ptr, grad.ptr if grad is not None else None,
                                             loss.ptr if loss is not None else None))

    def backward(self, X, dY):
        """bh_mlp_backward: parameter gradient from an upstream gradient dL/dY.
        Device arrays: the gradient stays in the network for adam_step (returns
        None); host arrays: also returned."""
        if isinstance(X, (DeviceArray, _DeviceView)):
            n = X.n // self.in_p
            _check(lib().bh_mlp_backward(self.h, n, X.ptr, dY.ptr, None, BH_DEVICE_PTR))
            return None
        X = np.ascontiguousarray(X, np.float32)
        dY = np.ascontiguousarray(dY, np.float32)
        g = np.zeros(self.n_params, np.float32)
        _check(lib().bh_mlp_backward(self.h, X.shape[0], X.ctypes.data, dY.ctypes.data, g.ctypes.data, 0))
        return g

    def adam_step(self, lr):
        _check(lib().bh_mlp_adam_step(self.h, lr))


def mlp_features(op, r, X, s, in_p):
    """device feature map of a residual (all DeviceArrays)."""
    _check(lib().bh_mlp_features(op.h, r.ptr, X.ptr, s.ptr, in_p, BH_DEVICE_PTR))


def dot(ctx, x, y):
    """x . y of two FP64 device arrays (bh_dot)."""
    out = _c_d()
    _check(lib().bh_dot(ctx.h, x.n, x.ptr, y.ptr, ctypes.byref(out)))
    return out.value


def mlp_correction(op, s, Y, e, out_p):
    """device correction e = s h^2 / nu * Y (all DeviceArrays)."""
    _check(lib().bh_mlp_correction(op.h, s.ptr, Y.ptr, e.ptr, out_p, BH_DEVICE_PTR))


def mlp_upstream(op, s, g, c, dY, out_p):
    """device upstream gradient dY = c s h^2 / nu * g (all DeviceArrays)."""
    _check(lib().bh_mlp_upstream(op.h, s.ptr, g.ptr, float(c), dY.ptr, out_p, BH_DEVICE_PTR))


def mlp_targets(op, s, e, T, out_p):
    _check(lib().bh_mlp_targets(op.h, s.ptr, e.ptr, T.ptr, out_p, BH_DEVICE_PTR))


def pmg_set_learned(pmg, mlp):
    _check(lib().bh_pmg_set_learned(pmg.h, mlp.h if mlp is not None else None))
    pmg._mlp = mlp


def pmg_learned_stats(pmg):
    a, r = _c_i(), _c_i()
    _check(lib().bh_pmg_learned_stats(pmg.h, ctypes.byref(a), ctypes.byref(r)))
    return a.value, r.value


def pmg_set_learned_fallback(pmg, probe=8, min_gain=1e-2):
    """After `probe` calls, a mean removed |r'|^2 fraction below min_gain switches the network off."""
    _check(lib().bh_pmg_set_learned_fallback(pmg.h, int(probe), float(min_gain)))


def pmg_set_learned_cost_rule(pmg, on=True):
    """At the end of the probe window keep the network only if flexible GMRES with it
    reduces the residual at least as fast per millisecond as without it (on by default)."""
    _check(lib().bh_pmg_set_learned_cost_rule(pmg.h, 1 if on else 0))


def pmg_learned_costs(pmg):
    """dict(cycle_ms, step_ms, cycle_log_reduction, step_log_reduction): per iteration of
    flexible GMRES preconditioned without (cycle_) and with (step_) the learned step."""
    out = np.zeros(4)
    _check(lib().bh_pmg_learned_costs(pmg.h, out.ctypes.data))
    return {"cycle_ms": float(out[0]), "step_ms": float(out[1]), "cycle_log_reduction": float(out[2]),
            "step_log_reduction": float(out[3])}


def pmg_learned_info(pmg):
    """(still active, mean removed |r'|^2 fraction per step over the probe window)"""
    a, g = _c_i(), _c_d()
    _check(lib().bh_pmg_learned_info(pmg.h, ctypes.byref(a), ctypes.byref(g)))
    return bool(a.value), g.value


def pmg_set_cell_ilu(pmg, ilu_level=1):
    """CellILU (block ILU on the pattern of A^(ilu_level+1)) as the finest level's smoother;
    ilu_level < 0: block-Jacobi again."""
    _check(lib().bh_pmg_set_cell_ilu(pmg.h, int(ilu_level)))


def pmg_smooth(pmg, level, r):
    """x = S(r): the smoother of `level` from a zero guess (host FP64 vectors)."""
    r = np.ascontiguousarray(r, np.float64)
    x = np.zeros_like(r)
    _check(lib().bh_pmg_smooth(pmg.h, int(level), r.ctypes.data, x.ctypes.data))
    return x


def pmg_levels(pmg):
    """The multigrid mapping: per level dict(nrows, bs, kind (0 truncation, 1 aggregation,
    -1 coarsest), deg, n, nnzb)."""
    info = np.zeros(7, np.int32)
    _check(lib().bh_pmg_level_info(pmg.h, -1, info.ctypes.data, None))
    out = []
    for lvl in range(int(info[0])):
        nz = _c_i64()
        _check(lib().bh_pmg_level_info(pmg.h, lvl, info.ctypes.data, ctypes.byref(nz)))
        out.append(dict(nrows=int(info[0]), bs=int(info[1]), kind=int(info[2]), deg=int(info[3]),
                        n=[int(v) for v in info[4:7]], nnzb=int(nz.value)))
    return out


def pmg_level_transfer(pmg, level):
    """(parent aggregate per cell, P[cell][fine mode][coarse mode]) of an aggregation level."""
    inf = pmg_levels(pmg)[level]
    par = np.zeros(inf["nrows"], np.int32)
    P = np.zeros((inf["nrows"], inf["bs"], inf["bs"]))
    _check(lib().bh_pmg_level_transfer(pmg.h, int(level), par.ctypes.data, P.ctypes.data))
    return par, P


def pmg_level_matrix(pmg, level):
    """(rowptr, col, blocks[nnzb, bs, bs]) of a level's Galerkin matrix."""
    inf = pmg_levels(pmg)[level]
    rp = np.zeros(inf["nrows"] + 1, np.int64)
    col = np.zeros(inf["nnzb"], np.int32)
    val = np.zeros(inf["nnzb"] * inf["bs"] * inf["bs"])
    _check(lib().bh_pmg_level_matrix(pmg.h, int(level), rp.ctypes.data, col.ctypes.data, val.ctypes.data))
    return rp, col, val.reshape(inf["nnzb"], inf["bs"], inf["bs"])


def pmg_set_level_pmg(pmg, enable=True):
    """LevelPmg on the finest level: low-order coarse correction, then each cell's
    high-order modes with the inverse of their diagonal sub-block."""
    _check(lib().bh_pmg_set_level_pmg(pmg.h, 1 if enable else 0))


def pmg_set_schwarz(pmg, enable=True, overlap=0, overlap_scaling=False):
    """Additive Schwarz on the multigrid aggregates, grown by `overlap` (0-2) layers of
    face neighbours, as the smoother of the aggregated levels; with overlap_scaling a
    DOF's sum of block solutions is divided by the number of blocks holding it."""
    _check(lib().bh_pmg_set_schwarz_overlap(pmg.h, 1 if enable else 0, int(overlap), 1 if overlap_scaling else 0))


def pmg_set_diag_equilib(pmg):
    """DiagBlockEquilib: every level but the coarsest in the basis where its diagonal blocks
    are the identity (assembled hierarchies only; irreversible)."""
    _check(lib().bh_pmg_set_diag_equilib(pmg.h))


def pmg_set_graphs(pmg, enable):
    """HIP-graph replay of the synchronisation-free V-cycles on (default) / off."""
    _check(lib().bh_pmg_set_graphs(pmg.h, 1 if enable else 0))


class EulerOperator(_Owned):
    """Explicit DG Euler/HLLC operator on a periodic Cartesian grid (bh_euler_*)."""

    _destroy_fn = "bh_euler_destroy"

    def __init__(self, grid, k=2, nq1=4, gamma=1.4, mach=None):
        self.grid = grid
        self.ctx = grid.ctx
        self.mach = mach if mach is not None else 1.0 / np.sqrt(gamma)
        h = ctypes.c_void_p()
        _check(lib().bh_euler_create(grid.h, k, nq1, gamma, self.mach, ctypes.byref(h)))
        self._register(self.ctx, h.value)
        N, Np = _c_i64(), _c_i()
        _check(lib().bh_euler_info(self.h, ctypes.byref(N), ctypes.byref(Np)))
        self.N, self.Np = N.value, Np.value

    def rhs(self, U, R, prec=BH_FP64):
        up, fu = _vec_args(U)
        rp, fr = _vec_args(R)
        assert fu == fr
        _check(lib().bh_euler_rhs(self.h, up, rp, prec, fu))
        return R

    def tvd3(self, U, dt, steps, prec=BH_FP64):
        up, fu = _vec_args(U)
        _check(lib().bh_euler_tvd3(self.h, up, dt, steps, prec, fu))
        return U

    def implicit_step(self, Un, dt, U=None, ConvCrit=1e-10, MaxIter=50, restart=30):
        """Backward-Euler step U - Un - dt R(U) = 0 by Newton-Krylov (bh_euler_implicit_step);
        returns (U, dict(newton, gmres, cuts, res0, res))."""
        Un = np.ascontiguousarray(Un, np.float64)
        U = Un.copy() if U is None else np.ascontiguousarray(U, np.float64)
        info = np.zeros(3, np.int32)
        res = np.zeros(2)
        _check(lib().bh_euler_implicit_step(self.h, Un.ctypes.data, U.ctypes.data, dt, ConvCrit, MaxIter, restart,
                                            info.ctypes.data, res.ctypes.data, 0))
        return U, dict(newton=int(info[0]), gmres=int(info[1]), cuts=int(info[2]), res0=float(res[0]),
                       res=float(res[1]))


class OrthonormalizationMultigrid(_Solver):
    """OrthonormalizationMultigrid (bh_orthomg_create) on the finest level of a
    PMultigrid: pre-smoother, coarse-grid and post-smoother corrections (and an
    optional learned one), each orthonormalised in the A-image against all earlier
    corrections of the solve and added with the residual-minimising coefficient."""

    def __init__(self, mg, op, M=None, pre=1, NoOfPostSmootherSweeps=1, coarse_reps=1, max_hist=64,
                 ConvergenceCriterion=1e-10, MaxSolverIterations=100):
        h = ctypes.c_void_p()
        _check(lib().bh_orthomg_create(mg.h, M.h if M else None, op.h, pre, NoOfPostSmootherSweeps, coarse_reps,
                                       max_hist, ConvergenceCriterion, MaxSolverIterations, ctypes.byref(h)))
        super().__init__(op.ctx, h.value, op.N, (mg, op, M))
        self._net = None

    def set_learned(self, net):
        _check(lib().bh_orthomg_set_learned(self.h, net.h if net is not None else None))
        self._net = net

    def info(self):
        """dict(added, cancellations, random, restarts) over the solver's life."""
        v = np.zeros(4, np.int32)
        _check(lib().bh_orthomg_info(self.h, v.ctypes.data))
        return dict(added=int(v[0]), cancellations=int(v[1]), random=int(v[2]), restarts=int(v[3]))


class XdgPoisson(_Owned):
    """Two-phase SIP Poisson with a spherical interface (bh_xdg_*): the
    XLaplace_Bulk / XLaplace_Interface operator of XdgPoisson3
    (L3-solution/BoSSS.Solution.XNSECommon/XLaplace.cs:50-346) on cut cells,
    with cell agglomeration (X/AgglomerationAlgorithm.cs)."""

    _destroy_fn = "bh_xdg_destroy"

    def __init__(self, grid, k, center, radius, nu_in, nu_out, penalty_base=1.3, agg_threshold=0.1, quad_order=0,
                 _handle=None):
        self.grid = grid
        self.ctx = grid.ctx
        self.k = k
        if _handle is None:
            c = np.zeros(3)
            c[:grid.D] = center
            h = ctypes.c_void_p()
            _check(lib().bh_xdg_create(grid.h, k, c.ctypes.data, radius, nu_in, nu_out, penalty_base, agg_threshold,
                                       quad_order, ctypes.byref(h)))
            _handle = h.value
        self._register(self.ctx, _handle)
        info = np.zeros(13, np.int64)
        _check(lib().bh_xdg_info(self.h, info.ctypes.data))
        (self.N_full, self.N, self.ncut, self.npairs, self.nfailed, self.nelem, self.nnodes, self.nvals_full,
         self.nvals, self.nbnd, self.Np, self.J, self.nunits) = [int(v) for v in info]

    @classmethod
    def poly(cls, grid, k, coeffs, nu_in, nu_out, penalty_base=1.3, agg_threshold=0.1, quad_order=0):
        """Level set phi(x) = sum coeffs[i0, i1(, i2)] x^i0 y^i1 z^i2 (an array of shape
        (deg+1,)*D indexed [i0, i1, i2]); species A where phi < 0 (bh_xdg_create_poly)."""
        c = np.asarray(coeffs, np.float64)
        assert c.ndim == grid.D and len(set(c.shape)) == 1
        deg = c.shape[0] - 1
        flat = np.ascontiguousarray(c.transpose(tuple(range(grid.D))[::-1]).ravel())   # i0 fastest
        h = ctypes.c_void_p()
        _check(lib().bh_xdg_create_poly(grid.h, k, deg, flat.ctypes.data, nu_in, nu_out, penalty_base, agg_threshold,
                                        quad_order, ctypes.byref(h)))
        return cls(grid, k, None, None, nu_in, nu_out, _handle=h.value)

    def cells(self):
        J = self.J
        m = np.zeros(J, np.int32)
        vol = np.zeros((2, J))
        ia = np.zeros(J)
        ln = np.zeros((2, J))
        _check(lib().bh_xdg_cells(self.h, m.ctypes.data, vol.ctypes.data, ia.ctypes.data, ln.ctypes.data))
        return dict(spmask=m, vol=vol, iface_area=ia, len=ln)

    def agglomeration(self):
        p = np.zeros((self.npairs, 3), np.int32)
        if self.npairs:
            _check(lib().bh_xdg_agglomeration(self.h, p.ctypes.data))
        return p

    def elements(self):
        ints = np.zeros((self.nelem, 9), np.int32)
        dbls = np.zeros((self.nelem, 3))
        nx = np.zeros((self.nnodes, 3))
        nw = np.zeros(self.nnodes)
        nn = np.zeros((self.nnodes, 3))
        _check(lib().bh_xdg_elements(self.h, ints.ctypes.data, dbls.ctypes.data, nx.ctypes.data, nw.ctypes.data,
                                     nn.ctypes.data))
        return dict(ints=ints, dbls=dbls, nx=nx, nw=nw, nn=nn)

    def matrix(self, which=1):
        """(dof_off, rowptr, col, voff, vals) of the full (0, cell rows) or agglomerated
        (1, species-unit rows) matrix."""
        nr = self.nunits if which else self.J
        nv = self.nvals if which else self.nvals_full
        off = np.zeros(nr + 1, np.int64)
        rp = np.zeros(nr + 1, np.int64)
        _check(lib().bh_xdg_matrix(self.h, which, off.ctypes.data, rp.ctypes.data, None, None, None))
        nb = int(rp[-1])
        col = np.zeros(nb, np.int32)
        voff = np.zeros(nb + 1, np.int64)
        vals = np.zeros(nv)
        _check(lib().bh_xdg_matrix(self.h, which, None, None, col.ctypes.data, voff.ctypes.data, vals.ctypes.data))
        return off, rp, col, voff, vals

    def scipy_matrix(self, which=1):
        import scipy.sparse as sp
        off, rp, col, voff, vals = self.matrix(which)
        rows, cols, data = [], [], []
        for r in range(len(off) - 1):
            R = int(off[r + 1] - off[r])
            for p in range(rp[r], rp[r + 1]):
                c = col[p]
                C = int(off[c + 1] - off[c])
                blk = vals[voff[p]:voff[p] + R * C].reshape(C, R).T
                ii, jj = np.meshgrid(np.arange(R) + off[r], np.arange(C) + off[c], indexing="ij")
                rows.append(ii.ravel()); cols.append(jj.ravel()); data.append(blk.ravel())
        n = int(off[-1])
        return sp.csr_matrix((np.concatenate(data), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))

    def bnd_nodes(self):
        x = np.zeros((self.nbnd, 3))
        if self.nbnd:
            _check(lib().bh_xdg_bnd_nodes(self.h, x.ctypes.data))
        return x

    def set_rhs(self, f, g=None):
        f = np.ascontiguousarray(f, np.float64)
        gp = None
        if g is not None:
            g = np.ascontiguousarray(g, np.float64)
            assert g.shape == (self.nbnd,)
            gp = g.ctypes.data
        _check(lib().bh_xdg_set_rhs(self.h, f.ctypes.data, gp))

    def rhs(self, which=1):
        b = np.zeros(self.N if which else self.N_full)
        _check(lib().bh_xdg_rhs(self.h, which, b.ctypes.data, 0))
        return b

    def SpMV(self, alpha, x, beta, y, which=1, prec=BH_FP64):
        xp, fx = _vec_args(x)
        yp, fy = _vec_args(y)
        assert fx == fy
        _check(lib().bh_xdg_spmv(self.h, which, alpha, xp, beta, yp, prec, fx))
        return y

    def Evaluate(self, alpha, beta, output, u):
        """y <- beta y + alpha A u on the full DOF layout (matrix-free, bh_xdg_evaluate)."""
        up, fu = _vec_args(u)
        yp, fy = _vec_args(output)
        assert fu == fy
        _check(lib().bh_xdg_evaluate(self.h, alpha, up, beta, yp, fu))
        return output

    def dof_frame(self, which=1):
        """Per cell: offset in the XDG constant frame (j * 2 Np), length, and offset
        in the compressed layout of the full (0) or agglomerated (1) system."""
        i0 = np.zeros(self.J, np.int64)
        ln = np.zeros(self.J, np.int32)
        c0 = np.zeros(self.J, np.int64)
        _check(lib().bh_xdg_dof_frame(self.h, which, i0.ctypes.data, ln.ctypes.data, c0.ctypes.data))
        return i0, ln, c0

    def mass_matrix(self):
        """Per (cell, species) mass blocks over the full layout: array (N_full, Np),
        row r holds the block row of DOF r (identity blocks in uncut cells)."""
        out = np.zeros((self.N_full, self.Np))
        _check(lib().bh_xdg_mass_matrix(self.h, out.ctypes.data))
        return out

    def extend(self, u):
        u = np.ascontiguousarray(u, np.float64)
        out = np.zeros(self.N_full)
        _check(lib().bh_xdg_extend(self.h, u.ctypes.data, out.ctypes.data))
        return out


class XdgPrecond(_Solver):
    """Two-level preconditioner of the agglomerated XDG system (bh_xdg_precond_create).
    The defaults (one pre- and one post-sweep, a 4-step K-cycle, levels down to 512
    cells) were the fastest of a configuration sweep on a sphere in 48^3 at k = 2."""

    def __init__(self, xdg, coarse_degree=1, min_coarse=512, omega=0.8, pre=1, post=1, kcycle_inner=4, prec=BH_FP32):
        h = ctypes.c_void_p()
        _check(lib().bh_xdg_precond_create(xdg.h, coarse_degree, min_coarse, omega, pre, post, kcycle_inner, prec,
                                           ctypes.byref(h)))
        super().__init__(xdg.ctx, h.value, xdg.N, (xdg,))


class XdgPrecondV(_Solver):
    """The two-level XDG preconditioner for a caller's variable-block matrix of
    cell blocks in the caller's vector layout (bh_xdg_precond_create_v)."""

    def __init__(self, grid, M, Np, coarse_degree=1, min_coarse=512, omega=0.8, pre=1, post=1, kcycle_inner=4,
                 prec=BH_FP32):
        h = ctypes.c_void_p()
        _check(lib().bh_xdg_precond_create_v(grid.h, M.h, Np, coarse_degree, min_coarse, omega, pre, post,
                                             kcycle_inner, prec, ctypes.byref(h)))
        super().__init__(M.ctx, h.value, M.N, (grid, M))


def xdg_cut_rule_fallbacks():
    """boxes (process-wide) where the polynomial cut rules used the low-order fallback"""
    v = ctypes.c_long()
    _check(lib().bh_xdg_cut_rule_fallbacks(ctypes.byref(v)))
    return v.value


def xdg_precond_coarse_matrix(precond):
    """(rowptr, col, blocks[nnzb, bs, bs]) of an XDG preconditioner's coarse matrix."""
    n, bs, nz = _c_i(), _c_i(), _c_i64()
    _check(lib().bh_xdg_precond_coarse_export(precond.h, ctypes.byref(n), ctypes.byref(bs), ctypes.byref(nz), None, None, None))
    rp = np.zeros(n.value + 1, np.int64)
    col = np.zeros(nz.value, np.int32)
    val = np.zeros(nz.value * bs.value * bs.value)
    _check(lib().bh_xdg_precond_coarse_export(precond.h, None, None, None, rp.ctypes.data, col.ctypes.data, val.ctypes.data))
    return rp, col, val.reshape(nz.value, bs.value, bs.value)


def xdg_precond_set_fp64_blocks(precond, on=True):
    """FP32 XDG preconditioner: cell blocks applied with their FP64 inverse."""
    _check(lib().bh_xdg_precond_set_fp64_blocks(precond.h, 1 if on else 0))


class XdgFlexGMRES(_Solver):
    """FGMRES on the agglomerated XDG system (bh_xdg_fgmres_create)."""

    def __init__(self, xdg, Precond=None, MaxKrylovDim=30, ConvergenceCriterion=1e-8, MaxSolverIterations=10000):
        h = ctypes.c_void_p()
        _check(lib().bh_xdg_fgmres_create(xdg.h, Precond.h if Precond else None, MaxKrylovDim, ConvergenceCriterion,
                                          MaxSolverIterations, ctypes.byref(h)))
        super().__init__(xdg.ctx, h.value, xdg.N, (xdg, Precond))


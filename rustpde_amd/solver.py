"""SolverField on the GPU: Hholtz / HholtzAdi / Poisson of a field's bases (src/solver.rs),
solving batches of right-hand sides (the reference's Solve::solve, one field at a time)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr
from .funspace import _Dev

HHOLTZ_ADI, HHOLTZ, POISSON = 0, 1, 2


class Solver:
    """Solver(kind, base0, base1, c): kind in {HHOLTZ_ADI, HHOLTZ, POISSON}; bases (kind, n)."""

    def __init__(self, kind, base0, base1, c):
        h = ctypes.c_void_p()
        check(_lib.load().rbc_solver_create(int(kind), base0[0], base0[1], base1[0], base1[1], float(c[0]),
                                            float(c[1]), ctypes.byref(h)))
        self._h = h
        v = [ctypes.c_int() for _ in range(5)]
        check(_lib.load().rbc_solver_shape(h, *[ctypes.byref(x) for x in v]))
        o0, o1, s0, s1, cp = [x.value for x in v]
        self.shape_in, self.shape_out, self.is_complex = (o0, o1), (s0, s1), bool(cp)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().rbc_solver_destroy(self._h)
            self._h = None

    def solve_device(self, d_rhs, d_out, batch):
        """rhs/out are device pointers (ints or c_void_p) holding `batch` contiguous fields."""
        check(_lib.load().rbc_solver_solve(self._h, ctypes.c_void_p(int(getattr(d_rhs, "value", d_rhs) or 0)),
                                           ctypes.c_void_p(int(getattr(d_out, "value", d_out) or 0)), int(batch)))

    def solve_device_noalloc(self, d_rhs, d_out, batch):
        """solve_device without any device allocation: the work arrays live in the library's
        fixed per-device scratch (32 MiB); raises RbcError when they do not fit.  Calls that
        overlap on one GPU share that scratch and race."""
        check(_lib.load().rbc_solver_solve_noalloc(self._h, ctypes.c_void_p(int(getattr(d_rhs, "value", d_rhs) or 0)),
                                                   ctypes.c_void_p(int(getattr(d_out, "value", d_out) or 0)),
                                                   int(batch)))

    def set_mixed(self, iters):
        """FP32 eigen-transforms with `iters` chained solves of FP64-residual refinement (0: FP64)"""
        check(_lib.load().rbc_solver_set_mixed(self._h, int(iters)))

    def residual(self, rhs, sol):
        """relative max-norm of P_x f P_y^T - L(u) for host arrays rhs [.., o0, o1], sol [.., s0, s1]"""
        a = np.ascontiguousarray(rhs, dtype=float)
        u = np.ascontiguousarray(sol, dtype=float)
        batch = 1 if a.ndim == 2 else a.shape[0]
        din, dsol = _Dev(a.nbytes), _Dev(u.nbytes)
        L = _lib.load()
        check(L.rbc_memcpy_h2d(din.p, a.ctypes.data, a.nbytes))
        check(L.rbc_memcpy_h2d(dsol.p, u.ctypes.data, u.nbytes))
        out = ctypes.c_double()
        check(L.rbc_solver_residual(self._h, din.p, dsol.p, batch, ctypes.byref(out)))
        return out.value

    def solve(self, rhs):
        """Host arrays [o0, o1] or [batch, o0, o1] -> [.., s0, s1]."""
        a = np.ascontiguousarray(rhs, dtype=complex if self.is_complex else float)
        batch = 1 if a.ndim == 2 else a.shape[0]
        out = np.zeros(((batch,) if a.ndim == 3 else ()) + self.shape_out, dtype=a.dtype)
        din, dout = _Dev(a.nbytes), _Dev(out.nbytes)
        L = _lib.load()
        check(L.rbc_memcpy_h2d(din.p, a.ctypes.data, a.nbytes))
        self.solve_device(din.p, dout.p, batch)
        check(L.rbc_memcpy_d2h(out.ctypes.data, dout.p, out.nbytes))
        return out

    def tensor_setup(self):
        n, fb = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().rbc_solver_tensor_info(self._h, ctypes.byref(n), ctypes.byref(fb)))
        if n.value == 0:
            return None
        m = n.value
        lam, fwd, bwd = np.zeros(m), np.zeros((m, m)), np.zeros((m, m))
        check(_lib.load().rbc_solver_get_tensor(self._h, dptr(lam), dptr(fwd), dptr(bwd)))
        return (lam, fwd, bwd) if fb.value else (lam, None, None)

    def set_tensor(self, lam, fwd=None, bwd=None):
        lam = np.ascontiguousarray(lam, dtype=float)
        f = None if fwd is None else dptr(np.ascontiguousarray(fwd, dtype=float))
        b = None if bwd is None else dptr(np.ascontiguousarray(bwd, dtype=float))
        check(_lib.load().rbc_solver_set_tensor(self._h, dptr(lam), f, b))

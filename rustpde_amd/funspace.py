"""Space2 of funspace on the GPU (funspace/src/space2.rs) over the fs_* C entry points.

Host numpy arrays in the reference layouts go in and out; device staging uses the
library's own allocation helpers, so no other GPU runtime binding is needed.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check

CHEBYSHEV, CHEB_DIRICHLET, CHEB_NEUMANN, CHEB_DIRICHLET_BC, FOURIER_R2C = 0, 1, 2, 3, 4


class _Dev:
    def __init__(self, nbytes):
        self.p = ctypes.c_void_p()
        check(_lib.load().rbc_dev_malloc(ctypes.byref(self.p), max(int(nbytes), 8)))

    def __del__(self):
        if self.p:
            _lib.load().rbc_dev_free(self.p)


class Space2:
    """Space2::new(base0, base1) with bases given as (kind, n)."""

    def __init__(self, base0, base1):
        h = ctypes.c_void_p()
        check(_lib.load().fs_space2_create(base0[0], base0[1], base1[0], base1[1], ctypes.byref(h)))
        self._h = h
        v = [ctypes.c_int() for _ in range(7)]
        check(_lib.load().fs_space2_shape(h, *[ctypes.byref(x) for x in v]))
        p0, p1, s0, s1, o0, o1, c = [x.value for x in v]
        self.shape_physical, self.shape_spectral, self.shape_ortho = (p0, p1), (s0, s1), (o0, o1)
        self.is_complex = bool(c)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.load().fs_space2_destroy(self._h)
            self._h = None

    def _run(self, fn, arr, in_shape, in_cplx, out_shape, out_cplx, *extra):
        a = np.ascontiguousarray(arr, dtype=complex if in_cplx else float)
        batch = 1 if a.ndim == 2 else a.shape[0]
        if a.shape[-2:] != tuple(in_shape):
            raise ValueError("shape %s expected %s" % (a.shape, in_shape))
        out = np.zeros(((batch,) if a.ndim == 3 else ()) + tuple(out_shape), dtype=complex if out_cplx else float)
        din, dout = _Dev(a.nbytes), _Dev(out.nbytes)
        L = _lib.load()
        check(L.rbc_memcpy_h2d(din.p, a.ctypes.data, a.nbytes))
        check(fn(self._h, din.p, dout.p, *extra, batch))
        check(L.rbc_memcpy_d2h(out.ctypes.data, dout.p, out.nbytes))
        return out

    def forward(self, v):
        return self._run(_lib.load().fs_forward, v, self.shape_physical, False, self.shape_spectral, self.is_complex)

    def backward(self, vhat):
        return self._run(_lib.load().fs_backward, vhat, self.shape_spectral, self.is_complex, self.shape_physical, False)

    def to_ortho(self, vhat):
        return self._run(_lib.load().fs_to_ortho, vhat, self.shape_spectral, self.is_complex, self.shape_ortho,
                         self.is_complex)

    def from_ortho(self, p):
        return self._run(_lib.load().fs_from_ortho, p, self.shape_ortho, self.is_complex, self.shape_spectral,
                         self.is_complex)

    def gradient(self, vhat, deriv, scale=None):
        s0, s1 = (1.0, 1.0) if scale is None else scale
        return self._run(_lib.load().fs_gradient, vhat, self.shape_spectral, self.is_complex, self.shape_ortho,
                         self.is_complex, int(deriv[0]), int(deriv[1]), float(s0), float(s1))


def chebyshev(n):
    return (CHEBYSHEV, n)


def cheb_dirichlet(n):
    return (CHEB_DIRICHLET, n)


def cheb_neumann(n):
    return (CHEB_NEUMANN, n)


def cheb_dirichlet_bc(n):
    return (CHEB_DIRICHLET_BC, n)


def fourier_r2c(n):
    return (FOURIER_R2C, n)

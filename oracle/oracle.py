"""ctypes front end of the CPU FP64 oracle (oracle/rbc_oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker, never as the product.
Spectral arrays of periodic (Fourier x Chebyshev) fields are complex128; everything
else is float64; all arrays are C-ordered [n0, n1] like the reference's ndarray.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "librbc_oracle.so")
_LD_PATH = os.path.join(_HERE, "build", "librbc_oracle_ld.so")
_lib = None
_ldlib = None

# base kinds (funspace constructors): chebyshev, cheb_dirichlet, cheb_neumann,
# cheb_dirichlet_bc, fourier_r2c
CHEB, CD, CN, CDBC, FO = 0, 1, 2, 3, 4
FIELDS = {"temp": 0, "ux": 1, "uy": 2, "pres": 3, "pseu": 4, "field": 5, "fieldbc": 6}
_dp = ctypes.POINTER(ctypes.c_double)


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle`")
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_last_error.restype = ctypes.c_char_p
        L.orc_navier_create.restype = ctypes.c_void_p
        L.orc_navier_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] * 2
        for name in ("orc_navier_destroy",):
            getattr(L, name).argtypes = [ctypes.c_void_p]
        L.orc_navier_time.restype = ctypes.c_double
        L.orc_navier_time.argtypes = [ctypes.c_void_p]
        L.orc_navier_shape.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_navier_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_navier_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_navier_forward.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_navier_backward.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_navier_set_velocity.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        L.orc_navier_set_temperature.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        L.orc_navier_update.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_navier_eval.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_navier_tensor_info.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.orc_navier_tensor_get.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_navier_tensor_set.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
        L.orc_space_apply.argtypes = [ctypes.c_int] * 7 + [_dp, _dp, _dp]
        L.orc_solver_apply.argtypes = [ctypes.c_int] * 5 + [_dp, ctypes.c_int, _dp, _dp]
        L.orc_solver_create.restype = ctypes.c_void_p
        L.orc_solver_create.argtypes = [ctypes.c_int] * 5 + [_dp]
        L.orc_solver_destroy.argtypes = [ctypes.c_void_p]
        L.orc_solver_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_solver_tensor.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int),
                                        ctypes.POINTER(ctypes.c_int)]
        L.orc_eig.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_fdma_solve.argtypes = [ctypes.c_int, _dp, _dp, _dp]
        L.orc_dct1.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_cheb_diff.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_stats_enable.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double]
        L.orc_stats_update.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_stats_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_stats_info.argtypes = [ctypes.c_void_p, _dp]
        L.orc_navier_set_time.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_navier_conv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_navier_conv_term.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_navier_finish_conv.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_navier_set_solid.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_max_threads.restype = ctypes.c_int
        L.orc_adjoint_create.restype = ctypes.c_void_p
        L.orc_adjoint_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] * 2
        L.orc_adjoint_destroy.argtypes = [ctypes.c_void_p]
        L.orc_adjoint_part.restype = ctypes.c_void_p
        L.orc_adjoint_part.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_adjoint_update.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_adjoint_time.restype = ctypes.c_double
        L.orc_adjoint_time.argtypes = [ctypes.c_void_p]
        L.orc_adjoint_residual.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_adjoint_unsmoothed.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_adjoint_norms.argtypes = [ctypes.c_void_p, _dp]
        L.orc_adjoint_exit.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.orc_adjoint_smoother.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
        L.orc_adjoint_set_residual.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_adjoint_advance.argtypes = [ctypes.c_void_p]
        L.orc_adjoint_conv.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + lib().orc_last_error().decode())


def set_threads(n):
    lib().orc_set_threads(int(n))


def max_threads():
    return lib().orc_max_threads()


def _base_len(kind, n, spectral):
    if not spectral:
        return n
    return {CHEB: n, CD: n - 2, CN: n - 2, CDBC: 2, FO: n // 2 + 1}[kind]


def _ortho_len(kind, n):
    return n // 2 + 1 if kind == FO else n


def space_apply(k0, n0, k1, n1, op, arr, deriv=(0, 0), scale=None):
    """op: forward | backward | to_ortho | from_ortho | gradient (funspace Space2)."""
    cplx = k0 == FO
    ops = {"forward": 0, "backward": 1, "to_ortho": 2, "from_ortho": 3, "gradient": 4}
    o = ops[op]
    if o == 0:
        shp, out_cplx = (_base_len(k0, n0, True), _base_len(k1, n1, True)), cplx
    elif o == 1:
        shp, out_cplx = (n0, n1), False
    elif o in (2, 4):
        shp, out_cplx = (_ortho_len(k0, n0), _ortho_len(k1, n1)), cplx
    else:
        shp, out_cplx = (_base_len(k0, n0, True), _base_len(k1, n1, True)), cplx
    inp = np.ascontiguousarray(arr, dtype=complex if (cplx and o != 0) else float)
    out = np.zeros(shp, dtype=complex if out_cplx else float)
    sc = None if scale is None else np.ascontiguousarray(scale, dtype=float)
    _check(lib().orc_space_apply(k0, n0, k1, n1, o, deriv[0], deriv[1], _p(sc) if sc is not None else None,
                                 _p(inp.view(float)) if inp.dtype == complex else _p(inp), _p(out.view(float))))
    return out


def solver_apply(kind, k0, n0, k1, n1, c, rhs):
    """kind: 'hholtz_adi' | 'hholtz' | 'poisson'; rhs in ortho space."""
    sk = {"hholtz_adi": 0, "hholtz": 1, "poisson": 2}[kind]
    cplx = np.iscomplexobj(rhs)
    inp = np.ascontiguousarray(rhs, dtype=complex if cplx else float)
    out = np.zeros((_base_len(k0, n0, True), _base_len(k1, n1, True)), dtype=inp.dtype)
    cc = np.ascontiguousarray(c, dtype=float)
    _check(lib().orc_solver_apply(sk, k0, n0, k1, n1, _p(cc), int(cplx), _p(inp.view(float)), _p(out.view(float))))
    return out


class Solver:
    """Build-once Hholtz / HholtzAdi / Poisson of the field bases (k0, n0) x (k1, n1)."""

    def __init__(self, kind, k0, n0, k1, n1, c):
        sk = {"hholtz_adi": 0, "hholtz": 1, "poisson": 2}[kind]
        cc = np.ascontiguousarray(c, dtype=float)
        self._h = lib().orc_solver_create(sk, k0, n0, k1, n1, _p(cc))
        if not self._h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())
        self.shape_in = (_ortho_len(k0, n0), _ortho_len(k1, n1))
        self.shape_out = (_base_len(k0, n0, True), _base_len(k1, n1, True))

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_solver_destroy(self._h)
            self._h = None

    def solve(self, rhs):
        cplx = np.iscomplexobj(rhs)
        a = np.ascontiguousarray(rhs, dtype=complex if cplx else float)
        out = np.zeros(self.shape_out, dtype=a.dtype)
        _check(lib().orc_solver_solve(self._h, a.shape[0], a.shape[1], int(cplx), _p(a.view(float)),
                                      _p(out.view(float))))
        return out

    def tensor_setup(self):
        n, fb = ctypes.c_int(), ctypes.c_int()
        _check(lib().orc_solver_tensor(self._h, None, None, None, ctypes.byref(n), ctypes.byref(fb)))
        if n.value == 0:
            return None
        m = n.value
        lam, fwd, bwd = np.zeros(m), np.zeros((m, m)), np.zeros((m, m))
        _check(lib().orc_solver_tensor(self._h, _p(lam), _p(fwd), _p(bwd), ctypes.byref(n), ctypes.byref(fb)))
        return (lam, fwd, bwd) if fb.value else (lam, None, None)


def eig(a):
    a = np.ascontiguousarray(a, dtype=float)
    n = a.shape[0]
    lam, q, qi = np.zeros(n), np.zeros((n, n)), np.zeros((n, n))
    _check(lib().orc_eig(n, _p(a), _p(lam), _p(q), _p(qi)))
    return lam, q, qi


def fdma_solve(a, b):
    a = np.ascontiguousarray(a, dtype=float)
    x = np.array(b, dtype=float)
    _check(lib().orc_fdma_solve(a.shape[0], _p(a), _p(x), _p(x)))
    return x


def dct1(x):
    x = np.ascontiguousarray(x, dtype=float)
    y = np.zeros_like(x)
    _check(lib().orc_dct1(len(x), _p(x), _p(y)))
    return y


def cheb_diff(c, times):
    d = np.array(c, dtype=float)
    _check(lib().orc_cheb_diff(len(d), times, _p(d)))
    return d


class Navier2D:
    """Oracle Navier2D (navier.rs).  new(...) / new_periodic(...) mirror the reference."""

    def __init__(self, nx, ny, ra, pr, dt, aspect, adiabatic=True, periodic=False):
        self.periodic = periodic
        self._h = lib().orc_navier_create(nx, ny, ra, pr, dt, aspect, int(adiabatic), int(periodic))
        if not self._h:
            raise RuntimeError("oracle: " + lib().orc_last_error().decode())

    @classmethod
    def new(cls, nx, ny, ra, pr, dt, aspect, adiabatic):
        return cls(nx, ny, ra, pr, dt, aspect, adiabatic, False)

    @classmethod
    def _from_handle(cls, h, periodic, owner):
        self = cls.__new__(cls)
        self.periodic = periodic
        self._h = h
        self._owner = owner  # keeps the object the handle shares alive
        return self

    @classmethod
    def new_periodic(cls, nx, ny, ra, pr, dt, aspect):
        return cls(nx, ny, ra, pr, dt, aspect, True, True)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_navier_destroy(self._h)
            self._h = None

    def _shape(self, which, spectral):
        a, b = ctypes.c_int(), ctypes.c_int()
        _check(lib().orc_navier_shape(self._h, FIELDS[which], int(spectral), ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def get(self, which, spectral=True):
        shp = self._shape(which, spectral)
        out = np.zeros(shp, dtype=complex if (spectral and self.periodic) else float)
        _check(lib().orc_navier_get(self._h, FIELDS[which], int(spectral), _p(out.view(float))))
        return out

    def set(self, which, arr, spectral=True):
        shp = self._shape(which, spectral)
        a = np.ascontiguousarray(arr, dtype=complex if (spectral and self.periodic) else float)
        assert a.shape == shp, (a.shape, shp)
        _check(lib().orc_navier_set(self._h, FIELDS[which], int(spectral), _p(a.view(float))))

    def set_phys_forward(self, which, v):
        self.set(which, v, spectral=False)
        _check(lib().orc_navier_forward(self._h, FIELDS[which]))

    def backward(self, which):
        _check(lib().orc_navier_backward(self._h, FIELDS[which]))
        return self.get(which, spectral=False)

    def set_velocity(self, amp, m, n):
        _check(lib().orc_navier_set_velocity(self._h, amp, m, n))

    def set_temperature(self, amp, m, n):
        _check(lib().orc_navier_set_temperature(self._h, amp, m, n))

    def update(self, nsteps=1):
        _check(lib().orc_navier_update(self._h, int(nsteps)))

    def get_time(self):
        return lib().orc_navier_time(self._h)

    def conv(self, which):
        """Dealiased convective term of temp / ux / uy from the current state (navier.rs:538-616),
        spectral array of the `field` space."""
        shp = self._shape("field", True)
        out = np.zeros(shp, dtype=complex if self.periodic else float)
        _check(lib().orc_navier_conv(self._h, FIELDS[which], _p(out.view(float))))
        return out

    def conv_term(self, which, deriv, u):
        """conv_term(field, u, deriv) (conv_term.rs:22-42): u * d^deriv field, physical [nx, ny]."""
        shp = self._shape("field", False)
        ua = np.ascontiguousarray(u, dtype=float)
        assert ua.shape == shp, (ua.shape, shp)
        out = np.zeros(shp)
        _check(lib().orc_navier_conv_term(self._h, FIELDS[which], int(deriv[0]), int(deriv[1]), _p(ua), _p(out)))
        return out

    def finish_conv(self, phys):
        """forward transform in the `field` space, then dealias (navier.rs:562-568, 1022-1032)."""
        a = np.ascontiguousarray(phys, dtype=float)
        assert a.shape == self._shape("field", False)
        out = np.zeros(self._shape("field", True), dtype=complex if self.periodic else float)
        _check(lib().orc_navier_finish_conv(self._h, _p(a), _p(out.view(float))))
        return out

    def set_time(self, t):
        _check(lib().orc_navier_set_time(self._h, float(t)))

    def set_solid(self, solid):
        """navier.solid = Some([mask, value]) or None (physical [nx, ny] arrays)."""
        if solid is None:
            _check(lib().orc_navier_set_solid(self._h, None, None))
            return
        m = np.ascontiguousarray(solid[0], dtype=float)
        v = np.ascontiguousarray(solid[1], dtype=float)
        _check(lib().orc_navier_set_solid(self._h, _p(m), _p(v)))

    # Statistics (src/navier/statistics.rs): 0 t_avg, 1 ux_avg, 2 uy_avg, 3 nusselt
    def stats_enable(self, save_stat, write_stat):
        _check(lib().orc_stats_enable(self._h, save_stat, write_stat))

    def stats_update(self):
        u = ctypes.c_int()
        _check(lib().orc_stats_update(self._h, ctypes.byref(u)))
        return bool(u.value)

    def stats_get(self, which, spectral=True):
        shp = self._shape("field", spectral)
        out = np.zeros(shp, dtype=complex if (spectral and self.periodic) else float)
        _check(lib().orc_stats_get(self._h, which, int(spectral), _p(out.view(float))))
        return out

    def stats_info(self):
        """(avg_time, tot_time, num_save)"""
        o = np.zeros(3)
        _check(lib().orc_stats_info(self._h, _p(o)))
        return o[0], o[1], int(o[2])

    def _eval(self, k):
        v = ctypes.c_double()
        _check(lib().orc_navier_eval(self._h, k, ctypes.byref(v)))
        return v.value

    def eval_nu(self):
        return self._eval(0)

    def eval_nuvol(self):
        return self._eval(1)

    def eval_re(self):
        return self._eval(2)

    def div_norm(self):
        return self._eval(3)

    def tensor_setup(self, s):
        """(lam, fwd, bwd) of tensor solver s (0 ux, 1 uy, 2 temp, 3 pres); fwd/bwd None if diagonal."""
        n, fb = ctypes.c_int(), ctypes.c_int()
        _check(lib().orc_navier_tensor_info(self._h, s, ctypes.byref(n), ctypes.byref(fb)))
        if n.value == 0:
            return None
        m = n.value
        lam = np.zeros(m)
        fwd = np.zeros((m, m))
        bwd = np.zeros((m, m))
        _check(lib().orc_navier_tensor_get(self._h, s, _p(lam), _p(fwd), _p(bwd)))
        return (lam, fwd, bwd) if fb.value else (lam, None, None)

    def set_tensor(self, s, lam, fwd=None, bwd=None):
        lam = np.ascontiguousarray(lam, dtype=float)
        f = None if fwd is None else _p(np.ascontiguousarray(fwd, dtype=float))
        b = None if bwd is None else _p(np.ascontiguousarray(bwd, dtype=float))
        _check(lib().orc_navier_tensor_set(self._h, s, _p(lam), f, b))


def ldlib():
    """The oracle evaluated in long double (oracle/rbc_oracle_ld.cpp)."""
    global _ldlib
    if _ldlib is None:
        if not os.path.exists(_LD_PATH):
            raise RuntimeError("long double oracle missing: run `make -C oracle`")
        L = ctypes.CDLL(_LD_PATH)
        L.orcld_last_error.restype = ctypes.c_char_p
        L.orcld_navier_create.restype = ctypes.c_void_p
        L.orcld_navier_create.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] * 2
        L.orcld_navier_destroy.argtypes = [ctypes.c_void_p]
        L.orcld_navier_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orcld_navier_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orcld_navier_forward.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orcld_navier_set_velocity.argtypes = [ctypes.c_void_p] + [ctypes.c_double] * 3
        L.orcld_navier_update.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        _ldlib = L
    return _ldlib


class NavierLD:
    """The oracle's Navier2D evaluated in long double; fields read and written as float64 /
    complex128 arrays (rounded on the way out).  Shapes as the FP64 oracle's."""

    def __init__(self, nx, ny, ra, pr, dt, aspect, adiabatic=True, periodic=False, threads=None):
        L = ldlib()
        if threads:
            L.orc_set_threads(int(threads))
        self.periodic = periodic
        self._shapes = Navier2D(nx, ny, ra, pr, dt, aspect, adiabatic, periodic)  # shapes only
        self._h = L.orcld_navier_create(nx, ny, ra, pr, dt, aspect, int(adiabatic), int(periodic))
        if not self._h:
            raise RuntimeError("oracle (long double): " + L.orcld_last_error().decode())

    @classmethod
    def new(cls, nx, ny, ra, pr, dt, aspect, adiabatic, threads=None):
        return cls(nx, ny, ra, pr, dt, aspect, adiabatic, False, threads)

    @classmethod
    def new_periodic(cls, nx, ny, ra, pr, dt, aspect, threads=None):
        return cls(nx, ny, ra, pr, dt, aspect, True, True, threads)

    def __del__(self):
        if getattr(self, "_h", None) and _ldlib is not None:
            _ldlib.orcld_navier_destroy(self._h)
            self._h = None

    def _chk(self, rc):
        if rc != 0:
            raise RuntimeError("oracle (long double): " + ldlib().orcld_last_error().decode())

    def get(self, which, spectral=True):
        shp = self._shapes._shape(which, spectral)
        out = np.zeros(shp, dtype=complex if (spectral and self.periodic) else float)
        self._chk(ldlib().orcld_navier_get(self._h, FIELDS[which], int(spectral), _p(out.view(float))))
        return out

    def set(self, which, arr, spectral=True):
        shp = self._shapes._shape(which, spectral)
        a = np.ascontiguousarray(arr, dtype=complex if (spectral and self.periodic) else float)
        assert a.shape == shp, (a.shape, shp)
        self._chk(ldlib().orcld_navier_set(self._h, FIELDS[which], int(spectral), _p(a.view(float))))

    def set_phys_forward(self, which, v):
        self.set(which, v, spectral=False)
        self._chk(ldlib().orcld_navier_forward(self._h, FIELDS[which]))

    def set_velocity(self, amp, m, n):
        self._chk(ldlib().orcld_navier_set_velocity(self._h, amp, m, n))

    def update(self, nsteps=1):
        self._chk(ldlib().orcld_navier_update(self._h, int(nsteps)))


class NavierAdjoint:
    """Oracle Navier2DAdjoint (navier_adjoint.rs).  `fields` is a Navier2D view of the adjoint
    fields [0] (temp, ux, uy, pres, pseu, fieldbc; diagnostics and tensors of the adjoint's
    Poisson solver), `navier` the Navier-Stokes step whose residual is descended."""

    def __init__(self, nx, ny, ra, pr, dt, aspect, adiabatic=True, periodic=False):
        L = lib()
        self.periodic = periodic
        self._h = L.orc_adjoint_create(nx, ny, ra, pr, dt, aspect, int(adiabatic), int(periodic))
        if not self._h:
            raise RuntimeError("oracle: " + L.orc_last_error().decode())
        self.fields = Navier2D._from_handle(L.orc_adjoint_part(self._h, 0), periodic, self)
        self.navier = Navier2D._from_handle(L.orc_adjoint_part(self._h, 1), periodic, self)

    @classmethod
    def new(cls, nx, ny, ra, pr, dt, aspect, adiabatic):
        return cls(nx, ny, ra, pr, dt, aspect, adiabatic, False)

    @classmethod
    def new_periodic(cls, nx, ny, ra, pr, dt, aspect):
        return cls(nx, ny, ra, pr, dt, aspect, True, True)

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            for part in ("fields", "navier"):
                obj = self.__dict__.pop(part, None)
                if obj is not None and obj._h:
                    _lib.orc_navier_destroy(obj._h)
                    obj._h = None
            _lib.orc_adjoint_destroy(self._h)
            self._h = None

    def update(self, nsteps=1):
        _check(lib().orc_adjoint_update(self._h, int(nsteps)))

    def get_time(self):
        return lib().orc_adjoint_time(self._h)

    def residual(self, which):
        """smoothed residual [1] of 'temp' | 'ux' | 'uy' (spectral)"""
        shp = self.fields._shape(which, True)
        out = np.zeros(shp, dtype=complex if self.periodic else float)
        _check(lib().orc_adjoint_residual(self._h, {"temp": 0, "ux": 1, "uy": 2}[which], _p(out.view(float))))
        return out

    def set_residual(self, which, arr):
        """replace the smoothed residual [1] of 'temp' | 'ux' | 'uy' (spectral)"""
        shp = self.fields._shape(which, True)
        a = np.ascontiguousarray(arr, dtype=complex if self.periodic else float)
        assert a.shape == shp, (a.shape, shp)
        _check(lib().orc_adjoint_set_residual(self._h, {"temp": 0, "ux": 1, "uy": 2}[which], _p(a.view(float))))

    def advance(self):
        """the update without recomputing the residuals (navier_adjoint.rs:766-788 after :765)"""
        _check(lib().orc_adjoint_advance(self._h))

    def conv(self, which):
        """conv_temp / conv_ux / conv_uy (navier_adjoint.rs:488-578) of the current state"""
        shp = self.fields._shape("field", True)
        out = np.zeros(shp, dtype=complex if self.periodic else float)
        _check(lib().orc_adjoint_conv(self._h, {"temp": 0, "ux": 1, "uy": 2}[which], _p(out.view(float))))
        return out

    def unsmoothed(self, which):
        """fields_unsmoothed of 'ux' | 'uy' | 'temp' (ortho space of the field)"""
        shp = self.fields._shape("field", True)
        out = np.zeros(shp, dtype=complex if self.periodic else float)
        _check(lib().orc_adjoint_unsmoothed(self._h, {"ux": 0, "uy": 1, "temp": 2}[which], _p(out.view(float))))
        return out

    def norms(self):
        """[|ux1|, |uy1|, |temp1|, |unsmoothed ux|, |unsmoothed uy|, |unsmoothed temp|]"""
        o = np.zeros(6)
        _check(lib().orc_adjoint_norms(self._h, _p(o)))
        return o

    def exit(self):
        r = ctypes.c_int()
        _check(lib().orc_adjoint_exit(self._h, ctypes.byref(r)))
        return bool(r.value)

    def smoother_tensor(self, s):
        """(lam, fwd, bwd) of smoother s (0 ux, 1 uy, 2 temp)"""
        hp = ctypes.c_void_p()
        _check(lib().orc_adjoint_smoother(self._h, int(s), ctypes.byref(hp)))
        n, fb = ctypes.c_int(), ctypes.c_int()
        _check(lib().orc_solver_tensor(hp, None, None, None, ctypes.byref(n), ctypes.byref(fb)))
        m = n.value
        lam, fwd, bwd = np.zeros(m), np.zeros((m, m)), np.zeros((m, m))
        _check(lib().orc_solver_tensor(hp, _p(lam), _p(fwd), _p(bwd), ctypes.byref(n), ctypes.byref(fb)))
        return (lam, fwd, bwd) if fb.value else (lam, None, None)

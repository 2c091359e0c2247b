"""Navier2D on the GPU: host mirror of rustpde's Navier2D (src/navier/navier.rs).

Same constructor names and arguments, the Integrate methods (update, get_time, get_dt,
callback, exit) and the diagnostics; fields are reached through `temp`, `ux`, `uy`,
`pres[0]`, `pres[1]` and `fieldbc` whose `.v` / `.vhat` read and write the device state
in the reference's layouts (numpy, C order; complex128 spectral arrays when periodic).
"""
import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import check, dptr

TEMP, UX, UY, PRES, PSEU, FIELDBC = 0, 1, 2, 3, 4, 6
PHYS, SPEC = 0, 1
NU, NUVOL, RE, DIV = 0, 1, 2, 3


def get_nu(ra, pr, height):
    """navier.rs:46-49"""
    return np.sqrt(pr / (ra / height ** 3.0))


def get_ka(ra, pr, height):
    """navier.rs:52-55"""
    return np.sqrt(1.0 / ((ra / height ** 3.0) * pr))


class FieldProxy:
    """Field2 view (src/field.rs:67-82): `.v` physical, `.vhat` spectral, on the device."""

    def __init__(self, navier, which):
        self._n, self._w = navier, which

    def _shape(self, space):
        a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(_lib.load().rbc_field_shape(self._n._h, self._w, space, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return (a.value, b.value), bool(c.value)

    def _get(self, space):
        shp, cplx = self._shape(space)
        out = np.zeros(shp, dtype=complex if cplx else float)
        check(_lib.load().rbc_get_field(self._n._h, self._w, space, dptr(out.view(float))))
        return out

    def _set(self, space, arr):
        shp, cplx = self._shape(space)
        a = np.ascontiguousarray(arr, dtype=complex if cplx else float)
        if a.shape != shp:
            raise ValueError("shape %s expected %s" % (a.shape, shp))
        if self._w == FIELDBC and space == SPEC:
            check(_lib.load().rbc_set_temp_bc(self._n._h, dptr(a.view(float))))
        else:
            check(_lib.load().rbc_set_field(self._n._h, self._w, space, dptr(a.view(float))))

    v = property(lambda self: self._get(PHYS), lambda self, a: self._set(PHYS, a))
    vhat = property(lambda self: self._get(SPEC), lambda self, a: self._set(SPEC, a))

    def forward(self):
        """Field2::forward (src/field.rs:103-105)"""
        check(_lib.load().rbc_forward(self._n._h, self._w))

    def backward(self):
        """Field2::backward (src/field.rs:108-110)"""
        check(_lib.load().rbc_backward(self._n._h, self._w))


class Navier2D:
    """GPU Navier2D.  Use Navier2D.new(...) or Navier2D.new_periodic(...)."""

    def __init__(self, handle, periodic, nx, ny, ra, pr, dt, aspect):
        self._h = handle
        self.periodic = periodic
        self.nx, self.ny, self.ra, self.pr, self.aspect = nx, ny, ra, pr, aspect
        self.scale = [aspect, 1.0]
        self.nu = get_nu(ra, pr, 2.0)
        self.ka = get_ka(ra, pr, 2.0)
        self.temp = FieldProxy(self, TEMP)
        self.ux = FieldProxy(self, UX)
        self.uy = FieldProxy(self, UY)
        self.pres = [FieldProxy(self, PRES), FieldProxy(self, PSEU)]
        self.fieldbc = FieldProxy(self, FIELDBC)
        self.diagnostics = {"time": [], "Nu": [], "Nuvol": [], "Re": []}
        self.write_intervall = None
        self.statistics = None  # Statistics (navier.rs:193-194); see Navier2D.init_statistics
        self.data_dir = None  # where callback writes snapshots and statistics (None: no files)

    @classmethod
    def new(cls, nx, ny, ra, pr, dt, aspect, adiabatic):
        """Navier2D::new (navier.rs:219-307): Chebyshev x Chebyshev, HholtzAdi + Poisson."""
        h = ctypes.c_void_p()
        check(_lib.load().rbc_create_confined(nx, ny, ra, pr, dt, aspect, int(bool(adiabatic)), ctypes.byref(h)))
        return cls(h, False, nx, ny, ra, pr, dt, aspect)

    @classmethod
    def new_periodic(cls, nx, ny, ra, pr, dt, aspect):
        """Navier2D::new_periodic (navier.rs:384-467): Fourier x Chebyshev, Hholtz + Poisson."""
        h = ctypes.c_void_p()
        check(_lib.load().rbc_create_periodic(nx, ny, ra, pr, dt, aspect, ctypes.byref(h)))
        return cls(h, True, nx, ny, ra, pr, dt, aspect)

    def close(self):
        if getattr(self, "_h", None) and not getattr(self, "_borrowed", False):
            _lib.load().rbc_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Integrate (src/lib.rs:135-146, navier.rs:732-863)
    def update(self, nsteps=1):
        """One (or nsteps) semi-implicit steps, queued on the device stream."""
        check(_lib.load().rbc_update(self._h, int(nsteps)))

    def synchronize(self):
        check(_lib.load().rbc_synchronize(self._h))

    def get_time(self):
        return _lib.load().rbc_get_time(self._h)

    def get_dt(self):
        return _lib.load().rbc_get_dt(self._h)

    @property
    def time(self):
        return self.get_time()

    @time.setter
    def time(self, t):
        check(_lib.load().rbc_set_time(self._h, float(t)))

    @property
    def dt(self):
        return self.get_dt()

    def _eval(self, what):
        v = ctypes.c_double()
        check(_lib.load().rbc_eval(self._h, what, ctypes.byref(v)))
        return v.value

    def eval_nu(self):
        return self._eval(NU)

    def eval_nuvol(self):
        return self._eval(NUVOL)

    def eval_re(self):
        return self._eval(RE)

    def div_norm(self):
        return self._eval(DIV)

    @staticmethod
    def _on_cadence(t, every, dt):
        """(time % every) < dt / 2 or > every - dt / 2 (navier.rs:784-786, :798-799)."""
        r = math.fmod(t, every)
        return r < dt / 2.0 or r > every - dt / 2.0

    def callback(self):
        """navier.rs:775-853: the snapshot (at write_intervall, every call when it is None)
        and the statistics update and file (at their save/write cadence), then the printed
        and recorded diagnostics.  Files go to `data_dir` (the reference's data/) and are
        written only when it is set; snapshots are the container of rustpde_amd.checkpoint."""
        from . import checkpoint
        t, dt = self.get_time(), self.get_dt()
        if self.data_dir is not None:
            os.makedirs(self.data_dir, exist_ok=True)
            if self.write_intervall is None or self._on_cadence(t, self.write_intervall, dt):
                self.write(os.path.join(self.data_dir, checkpoint.flow_filename(t)))
        if self.statistics is not None:
            st = self.statistics
            if self._on_cadence(t, st.save_stat, dt):
                st.update()
            if self.data_dir is not None and self._on_cadence(t, st.write_stat, dt):
                st.write(os.path.join(self.data_dir, "statistics" + checkpoint._SUFFIX))
        div, nu, nuvol, re = self.div_norm(), self.eval_nu(), self.eval_nuvol(), self.eval_re()
        print("time = %4.2f      |div| = %4.2e     Nu = %5.3e     Nuv = %5.3e    Re = %5.3e" % (t, div, nu, nuvol, re))
        for k, v in (("time", t), ("Nu", nu), ("Nuvol", nuvol), ("Re", re)):
            self.diagnostics[k].append(v)

    def exit(self):
        """navier.rs:855-862: True if the divergence norm is NaN."""
        r = ctypes.c_int()
        check(_lib.load().rbc_exit(self._h, ctypes.byref(r)))
        return bool(r.value)

    # ---- checkpoint / restart (navier.rs:961-1008) and statistics (navier.rs:193-194)
    def write(self, filename):
        """Navier2D::write: fields (v with the boundary field for temp, vhat), coordinates
        and scalars, in the container of rustpde_amd.checkpoint."""
        from . import checkpoint
        checkpoint.write_navier(self, filename)
        print(" ==> %r" % filename)

    def read(self, filename):
        """Navier2D::read: temp, ux, uy, pres[0] (broadcast across resolutions) and time."""
        from . import checkpoint
        checkpoint.read_navier(self, filename)
        print(" <== %r" % filename)

    @property
    def solid(self):
        """navier.solid (navier.rs:189-190): [mask, value] physical arrays, or None."""
        return getattr(self, "_solid", None)

    @solid.setter
    def solid(self, masks):
        if masks is None:
            check(_lib.load().rbc_set_solid(self._h, None, None))
            self._solid = None
            return
        m = np.ascontiguousarray(masks[0], dtype=float)
        v = np.ascontiguousarray(masks[1], dtype=float)
        if m.shape != (self.nx, self.ny) or v.shape != (self.nx, self.ny):
            raise ValueError("solid masks must be [nx, ny] = %s" % ((self.nx, self.ny),))
        check(_lib.load().rbc_set_solid(self._h, dptr(m), dptr(v)))
        self._solid = [m, v]

    def init_statistics(self, save_stat, write_stat):
        """navier.statistics = Some(Statistics::new(&navier, save_stat, write_stat))."""
        from .statistics import Statistics
        self.statistics = Statistics.new(self, save_stat, write_stat)
        return self.statistics

    # ---- initial conditions (navier.rs:927-948)
    def set_velocity(self, amp, m, n):
        check(_lib.load().rbc_set_velocity(self._h, amp, m, n))

    def set_temperature(self, amp, m, n):
        check(_lib.load().rbc_set_temperature(self._h, amp, m, n))

    def set_temp_bc(self, vhat):
        """Navier2D::set_temp_bc: replace fieldbc.vhat (and every quantity derived from it)."""
        self.fieldbc.vhat = vhat

    # ---- execution options (equivalent implementations; for A/B checks)
    OPT_GRAPH, OPT_MFMA, OPT_FFT, OPT_BANDS, OPT_DCT, OPT_BATCH, OPT_FUSE, OPT_MIXED, OPT_CONV3 = 0, 1, 2, 3, 4, 5, 6, 7, 8

    def set_option(self, option, value):
        check(_lib.load().rbc_set_option(self._h, int(option), int(value)))

    # ---- tensor solver set-up (FdmaTensor pub fields, fdma_tensor.rs:73-91)
    def tensor_setup(self, s):
        n, fb = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().rbc_tensor_info(self._h, s, ctypes.byref(n), ctypes.byref(fb)))
        if n.value == 0:
            return None
        m = n.value
        lam, fwd, bwd = np.zeros(m), np.zeros((m, m)), np.zeros((m, m))
        check(_lib.load().rbc_get_tensor(self._h, s, dptr(lam), dptr(fwd), dptr(bwd)))
        return (lam, fwd, bwd) if fb.value else (lam, None, None)

    def set_tensor(self, s, lam, fwd=None, bwd=None):
        lam = np.ascontiguousarray(lam, dtype=float)
        f = None if fwd is None else dptr(np.ascontiguousarray(fwd, dtype=float))
        b = None if bwd is None else dptr(np.ascontiguousarray(bwd, dtype=float))
        check(_lib.load().rbc_set_tensor(self._h, s, dptr(lam), f, b))

"""Navier2DAdjoint on the GPU: host mirror of rustpde's adjoint descent for steady states
(src/navier/navier_adjoint.rs), over the rbc_adjoint_* entry points.

Same constructor names and arguments; `temp`, `ux`, `uy` are pairs [adjoint field,
smoothed residual] as in the reference (index 0 the field with .v / .vhat / forward /
backward, index 1 the residual, read-only), `pres` the pair [pressure, pseudo pressure],
`navier` the Navier2D whose step defines the residual (dt_navier = 1e-2).  The Integrate
methods (update, get_time, get_dt, callback, exit) let integrate() drive it.
"""
import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import check, dptr
from .navier import (DIV, FIELDBC, NU, NUVOL, PRES, PSEU, RE, TEMP, UX, UY, FieldProxy, Navier2D, get_ka,
                     get_nu)


class _View:
    """Navier2D-like owner of a borrowed handle (the proxies only need ._h)."""

    def __init__(self, h):
        self._h = h


class ResidualProxy:
    """the smoothed residual [1] of a field (read-only): .vhat spectral, .unsmoothed the
    residual before smoothing (fields_unsmoothed, ortho space)"""

    def __init__(self, adjoint, which):
        self._a, self._w = adjoint, which

    def _read(self, kind):
        a = self._a
        if kind == 0:
            (n0, n1), _ = (a.temp[0] if self._w == TEMP else a.ux[0])._shape(1)
        else:
            n0, n1 = (a.nx // 2 + 1 if a.periodic else a.nx), a.ny
        out = np.zeros((n0, n1), dtype=complex if a.periodic else float)
        check(_lib.load().rbc_adjoint_residual(a._h, self._w, kind, dptr(out.view(float))))
        return out

    @property
    def vhat(self):
        return self._read(0)

    @property
    def unsmoothed(self):
        return self._read(1)


class Navier2DAdjoint:
    """Use Navier2DAdjoint.new(...) or Navier2DAdjoint.new_periodic(...)."""

    def __init__(self, handle, periodic, nx, ny, ra, pr, dt, aspect):
        self._h = handle
        self.periodic, self.nx, self.ny, self.ra, self.pr, self.aspect = periodic, nx, ny, ra, pr, aspect
        self.scale = [aspect, 1.0]
        self.nu, self.ka = get_nu(ra, pr, 2.0), get_ka(ra, pr, 2.0)
        L = _lib.load()
        p0, p1 = ctypes.c_void_p(), ctypes.c_void_p()
        check(L.rbc_adjoint_part(handle, 0, ctypes.byref(p0)))
        check(L.rbc_adjoint_part(handle, 1, ctypes.byref(p1)))
        self._fields = _View(p0)
        self.temp = [FieldProxy(self._fields, TEMP), ResidualProxy(self, TEMP)]
        self.ux = [FieldProxy(self._fields, UX), ResidualProxy(self, UX)]
        self.uy = [FieldProxy(self._fields, UY), ResidualProxy(self, UY)]
        self.pres = [FieldProxy(self._fields, PRES), FieldProxy(self._fields, PSEU)]
        self.fieldbc = FieldProxy(self._fields, FIELDBC)
        self.navier = Navier2D(p1, periodic, nx, ny, ra, pr, 1e-2, aspect)
        self.navier._borrowed = True
        self.diagnostics = {"time": [], "Nu": [], "Nuvol": [], "Re": []}
        self.residuals = []  # (time, unsmoothed total, smoothed total) as data/residual.txt
        self.write_intervall = None
        self.data_dir = None

    @classmethod
    def new(cls, nx, ny, ra, pr, dt, aspect, adiabatic):
        """Navier2DAdjoint::new (navier_adjoint.rs:187-289): Chebyshev x Chebyshev."""
        h = ctypes.c_void_p()
        check(_lib.load().rbc_adjoint_create_confined(nx, ny, ra, pr, dt, aspect, int(bool(adiabatic)),
                                                      ctypes.byref(h)))
        return cls(h, False, nx, ny, ra, pr, dt, aspect)

    @classmethod
    def new_periodic(cls, nx, ny, ra, pr, dt, aspect):
        """Navier2DAdjoint::new_periodic (navier_adjoint.rs:305-447): Fourier x Chebyshev."""
        h = ctypes.c_void_p()
        check(_lib.load().rbc_adjoint_create_periodic(nx, ny, ra, pr, dt, aspect, ctypes.byref(h)))
        return cls(h, True, nx, ny, ra, pr, dt, aspect)

    def close(self):
        if getattr(self, "_h", None):
            self.navier._h = None
            self._fields._h = None
            _lib.load().rbc_adjoint_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- Integrate
    def update(self, nsteps=1):
        check(_lib.load().rbc_adjoint_update(self._h, int(nsteps)))

    def get_time(self):
        return _lib.load().rbc_adjoint_get_time(self._h)

    def get_dt(self):
        return _lib.load().rbc_adjoint_get_dt(self._h)

    @property
    def time(self):
        return self.get_time()

    @time.setter
    def time(self, t):
        check(_lib.load().rbc_adjoint_set_time(self._h, float(t)))

    def reset_time(self):
        """navier_adjoint.rs:987-991"""
        self.time = 0.0
        self.navier.time = 0.0

    def norms(self):
        """[|ux1|, |uy1|, |temp1|, |unsmoothed ux|, |unsmoothed uy|, |unsmoothed temp|]"""
        out = np.zeros(6)
        check(_lib.load().rbc_adjoint_norms(self._h, dptr(out)))
        return out

    def exit(self):
        """navier_adjoint.rs:882-897: NaN divergence, or the smoothed residual below 1e-8"""
        r = ctypes.c_int()
        check(_lib.load().rbc_adjoint_exit(self._h, ctypes.byref(r)))
        return bool(r.value)

    def _eval(self, what):
        v = ctypes.c_double()
        check(_lib.load().rbc_eval(self._fields._h, what, ctypes.byref(v)))
        return v.value

    def eval_nu(self):
        return self._eval(NU)

    def eval_nuvol(self):
        return self._eval(NUVOL)

    def eval_re(self):
        return self._eval(RE)

    def div_norm(self):
        return self._eval(DIV)

    def callback(self):
        """navier_adjoint.rs:798-880: the snapshot (data/adjoint{time}.h5 in the reference; this
        library's container, written only when data_dir is set), the printed diagnostics and the
        residual norms."""
        from . import checkpoint
        t, dt = self.get_time(), self.get_dt()
        if self.data_dir is not None:
            os.makedirs(self.data_dir, exist_ok=True)
            r = math.fmod(t, self.write_intervall) if self.write_intervall else 0.0
            if self.write_intervall is None or r < dt / 2.0 or r > self.write_intervall - dt / 2.0:
                self.write(os.path.join(self.data_dir, "adjoint%s%s" % (format(t, "0>8.2f"), checkpoint._SUFFIX)))
        div, nu, nuvol, re = self.div_norm(), self.eval_nu(), self.eval_nuvol(), self.eval_re()
        print("time = %4.2f      |div| = %4.2e     Nu = %5.3e     Nuv = %5.3e    Re = %5.3e" % (t, div, nu, nuvol, re))
        for k, v in (("time", t), ("Nu", nu), ("Nuvol", nuvol), ("Re", re)):
            self.diagnostics[k].append(v)
        n = self.norms()
        print("|U| = %10.2e\n|V| = %10.2e\n|T| = %10.2e" % (n[0], n[1], n[2]))
        self.residuals.append((t, float(n[3:].sum()), float(n[:3].sum())))

    # ---- initial conditions (navier_adjoint.rs:974-985)
    def set_velocity(self, amp, m, n):
        check(_lib.load().rbc_set_velocity(self._fields._h, amp, m, n))

    def set_temperature(self, amp, m, n):
        check(_lib.load().rbc_set_temperature(self._fields._h, amp, m, n))

    def set_temp_bc(self, vhat):
        self.fieldbc.vhat = vhat

    # ---- tensors of the adjoint's solvers: 0, 1, 2 the smoothers of ux, uy, temp; 3 Poisson
    def tensor_setup(self, s):
        return Navier2D.tensor_setup(self._fields, s)

    def set_tensor(self, s, lam, fwd=None, bwd=None):
        Navier2D.set_tensor(self._fields, s, lam, fwd, bwd)

    # ---- checkpoint (navier_adjoint.rs:995-1066)
    def write(self, filename):
        """temp (with the boundary field), ux, uy, pres and the scalars, in the container of
        rustpde_amd.checkpoint"""
        from . import checkpoint
        data = {}
        for g, f in (("temp", self.temp[0]), ("ux", self.ux[0]), ("uy", self.uy[0]), ("pres", self.pres[0])):
            f.backward()
            v = f.v
            if g == "temp":
                v = v + self.fieldbc.v
            data[g + "/v"] = v
            data[g + "/vhat"] = f.vhat
        x0, y0 = checkpoint.coords(self.nx, self.periodic), checkpoint.coords(self.ny, False)
        data.update({"x": x0 * self.scale[0], "dx": checkpoint.get_dx(x0, self.periodic) * self.scale[0],
                     "y": y0 * self.scale[1], "dy": checkpoint.get_dx(y0, False) * self.scale[1]})
        data.update({"time": self.get_time(), "ra": self.ra, "pr": self.pr, "nu": self.nu, "kappa": self.ka})
        checkpoint.write_container(filename, data)
        print(" ==> %r" % filename)

    def read(self, filename):
        """temp, ux, uy (broadcast across resolutions) and time (navier_adjoint.rs:995-1003)"""
        from . import checkpoint
        data = checkpoint.read_container(filename)
        for g, f in (("temp", self.temp[0]), ("ux", self.ux[0]), ("uy", self.uy[0])):
            x = data[g + "/vhat"]
            cur = f.vhat
            if x.shape != cur.shape:
                x = checkpoint.broadcast_2d(x, cur)
            f.vhat = x
        self.time = float(data["time"][0])
        print(" <== %r" % filename)

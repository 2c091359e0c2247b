"""Ensembles of independent confined Navier2D instances advanced together on the GPU.

Each instance is Navier2D::new(nx, ny, ra, pr, dt, aspect, adiabatic) (navier.rs:219-307)
with its own parameters and fields; instances of equal shape, Pr, dt, aspect and sidewall
condition share one set of batched launches per step (rbc_create_ensemble).
"""
import ctypes

from . import _lib
from ._lib import RbcDesc, check
from .navier import Navier2D


class Ensemble:
    """descs: sequence of (nx, ny, ra, pr, dt, aspect, adiabatic)."""

    def __init__(self, descs):
        descs = [tuple(d) for d in descs]
        arr = (RbcDesc * len(descs))(*[RbcDesc(int(d[0]), int(d[1]), float(d[2]), float(d[3]), float(d[4]),
                                               float(d[5]), int(bool(d[6]))) for d in descs])
        L = _lib.load()
        h = ctypes.c_void_p()
        check(L.rbc_create_ensemble(len(descs), ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(h)))
        self._h = h
        self.instances = []
        for i, d in enumerate(descs):
            ih = ctypes.c_void_p()
            check(L.rbc_ensemble_instance(h, i, ctypes.byref(ih)))
            nav = Navier2D(ih, False, d[0], d[1], d[2], d[3], d[4], d[5])
            nav._borrowed = True  # owned by the ensemble
            self.instances.append(nav)

    def __len__(self):
        return len(self.instances)

    def __getitem__(self, i):
        return self.instances[i]

    @property
    def groups(self):
        n, g = ctypes.c_int(), ctypes.c_int()
        check(_lib.load().rbc_ensemble_size(self._h, ctypes.byref(n), ctypes.byref(g)))
        return g.value

    def update(self, nsteps=1):
        """Integrate::update of every instance, nsteps times."""
        check(_lib.load().rbc_ensemble_update(self._h, int(nsteps)))

    def synchronize(self):
        check(_lib.load().rbc_ensemble_synchronize(self._h))

    def set_option(self, option, value):
        """rbc_ensemble_set_option: an execution option (Navier2D.OPT_*) of every instance,
        e.g. Navier2D.OPT_MIXED for the mixed-precision step."""
        check(_lib.load().rbc_ensemble_set_option(self._h, int(option), int(value)))

    def timing(self, on):
        """Per-launch kernel timing of every group (steps run without graph replay)."""
        check(_lib.load().rbc_ensemble_timing_enable(self._h, 1 if on else 0))

    def timing_read(self, kernel_class):
        """(launches, summed ms, summed work) of one kernel class over all groups."""
        n, ms, w = ctypes.c_longlong(), ctypes.c_double(), ctypes.c_double()
        check(_lib.load().rbc_ensemble_timing_read(self._h, int(kernel_class), ctypes.byref(n), ctypes.byref(ms),
                                                   ctypes.byref(w)))
        return n.value, ms.value, w.value

    def close(self):
        if getattr(self, "_h", None):
            for nav in self.instances:
                nav._h = None
            _lib.load().rbc_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

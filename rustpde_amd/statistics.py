"""Statistics of a GPU Navier2D: host mirror of rustpde's Statistics (src/navier/statistics.rs).

`Statistics.new(navier, save_stat, write_stat)` (statistics.rs:54-88) enables the device
arrays of a handle; `update()` is the statistics.update call that Navier2D::callback makes
(navier.rs:796-807); `t_avg`, `ux_avg`, `uy_avg` and `nusselt` expose `.vhat` (the field
space's spectral array) and `.v` (its backward transform, as Statistics::write forms it,
statistics.rs:176-179).  `write`/`read` store and restore the arrays and scalars in the
checkpoint container of `rustpde_amd.checkpoint` under the dataset names of the reference's
statistics file (statistics.rs:169-193).
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, dptr

T_AVG, UX_AVG, UY_AVG, NUSSELT = 0, 1, 2, 3
PHYS, SPEC = 0, 1
NAMES = {"temp": T_AVG, "ux": UX_AVG, "uy": UY_AVG, "nusselt": NUSSELT}


class StatField:
    """One statistics array (a Field2 of the field space): `.vhat` and `.v`."""

    def __init__(self, stats, which):
        self._s, self._w = stats, which

    def _get(self, space):
        nav = self._s.navier
        if space == PHYS:
            out = np.zeros((nav.nx, nav.ny))
        elif nav.periodic:
            out = np.zeros((nav.nx // 2 + 1, nav.ny), dtype=complex)
        else:
            out = np.zeros((nav.nx, nav.ny))
        check(_lib.load().rbc_stats_get(nav._h, self._w, space, dptr(out.view(float))))
        return out

    @property
    def v(self):
        return self._get(PHYS)

    @property
    def vhat(self):
        return self._get(SPEC)

    @vhat.setter
    def vhat(self, arr):
        nav = self._s.navier
        shp = (nav.nx // 2 + 1, nav.ny) if nav.periodic else (nav.nx, nav.ny)
        a = np.ascontiguousarray(arr, dtype=complex if nav.periodic else float)
        if a.shape != shp:
            raise ValueError("shape %s expected %s" % (a.shape, shp))
        check(_lib.load().rbc_stats_set(nav._h, self._w, dptr(a.view(float))))


class Statistics:
    """Statistics<T, S> of one Navier2D handle (statistics.rs:10-41)."""

    def __init__(self, navier):
        self.navier = navier
        self.t_avg = StatField(self, T_AVG)
        self.ux_avg = StatField(self, UX_AVG)
        self.uy_avg = StatField(self, UY_AVG)
        self.nusselt = StatField(self, NUSSELT)

    @classmethod
    def new(cls, navier, save_stat, write_stat):
        """Statistics::new (statistics.rs:54-88): zero arrays, avg_time 0, tot_time = navier.time."""
        check(_lib.load().rbc_stats_enable(navier._h, float(save_stat), float(write_stat)))
        return cls(navier)

    def update(self):
        """statistics.update(that, ux.to_ortho(), uy.to_ortho(), time) (navier.rs:800-806);
        returns False and changes nothing when navier.time < tot_time (statistics.rs:134-140)."""
        u = ctypes.c_int()
        check(_lib.load().rbc_stats_update(self.navier._h, ctypes.byref(u)))
        if not u.value:
            print("Statistics time mismatch (navier < stat): %r < %r" % (self.navier.get_time(), self.tot_time))
        return bool(u.value)

    def _info(self):
        d = [ctypes.c_double() for _ in range(4)]
        n = ctypes.c_longlong()
        check(_lib.load().rbc_stats_info(self.navier._h, *[ctypes.byref(x) for x in d], ctypes.byref(n)))
        return [x.value for x in d] + [n.value]

    save_stat = property(lambda self: self._info()[0])
    write_stat = property(lambda self: self._info()[1])
    avg_time = property(lambda self: self._info()[2])
    tot_time = property(lambda self: self._info()[3])
    num_save = property(lambda self: self._info()[4])

    def set_info(self, avg_time, tot_time, num_save):
        check(_lib.load().rbc_stats_set_info(self.navier._h, float(avg_time), float(tot_time), int(num_save)))

    def write(self, filename):
        """Statistics::write (statistics.rs:163-193): arrays (v, vhat) and scalars."""
        from . import checkpoint
        nav = self.navier
        data = {}
        for name, f in (("temp", self.t_avg), ("ux", self.ux_avg), ("uy", self.uy_avg), ("nusselt", self.nusselt)):
            data[name + "/v"] = f.v
            data[name + "/vhat"] = f.vhat
        info = self._info()
        data.update({"tot_time": info[3], "avg_time": info[2], "num_save": info[4], "ra": nav.ra, "pr": nav.pr,
                     "nu": nav.nu, "ka": nav.ka})
        checkpoint.write_container(filename, data)
        print(" ==> %r" % filename)

    def read(self, filename):
        """Statistics::read (statistics.rs:197-212): arrays (broadcast to this shape when
        the file's differs, read.rs:62-77) and tot_time, avg_time, num_save."""
        from . import checkpoint
        data = checkpoint.read_container(filename)
        for name, f in (("temp", self.t_avg), ("ux", self.ux_avg), ("uy", self.uy_avg), ("nusselt", self.nusselt)):
            f.vhat = checkpoint.broadcast_2d(data[name + "/vhat"], f.vhat)
        self.set_info(float(data["avg_time"][0]), float(data["tot_time"][0]), int(data["num_save"][0]))
        print(" <== %r" % filename)

"""Checkpoint / restart of a GPU Navier2D (Navier2D::write / read, navier.rs:961-1008).

A snapshot holds exactly the datasets the reference writes, under the same paths:

* per field group (``temp``, ``ux``, ``uy``, ``pres``): ``<g>/v`` (physical; ``temp/v``
  includes the boundary field, navier.rs:987-990) and ``<g>/vhat`` (spectral; complex
  arrays as ``<g>/vhat_re`` and ``<g>/vhat_im``, as write_to_hdf5_complex stores them,
  hdf5-interface/src/lib.rs:195-232);
* ``x``, ``dx``, ``y``, ``dy`` (coordinates and cell widths, field/write.rs:88-92);
* scalars as one-element arrays (write_scalar_to_hdf5, hdf5-interface/src/lib.rs:257-268):
  ``time``, ``ra``, ``pr``, ``nu``, ``kappa``;
* ``solid/mask`` when an obstacle is set (navier.rs:995-998).

The container follows the file name: ``.h5`` is an HDF5 file written and read through the
HDF5 C library (rustpde_amd.h5; the reference's files read here and ours there), anything
else a NumPy ``.npz`` archive of the same datasets (loaded with ``allow_pickle=False``),
for hosts without libhdf5.  ``callback`` names its files with ``_SUFFIX``: ``.h5`` where the
library is found.  Reading restores temp, ux, uy and pres[0] from their ``vhat``
(broadcast to the handle's shape when the file holds another resolution,
field/read.rs:62-77 and :112-121) and the time.
"""
import math

import numpy as np

from . import h5

_SUFFIX = ".h5" if h5.available() else ".npz"


def write_container(filename, data):
    """Write {dataset path: array or scalar} (complex arrays split into _re/_im)."""
    flat = {}
    for k, v in data.items():
        a = np.asarray(v)
        if a.ndim == 0:
            a = a.reshape(1)
        if np.iscomplexobj(a):
            flat[k + "_re"] = np.ascontiguousarray(a.real)
            flat[k + "_im"] = np.ascontiguousarray(a.imag)
        else:
            flat[k] = a
    if str(filename).endswith(".h5"):
        h5.write(str(filename), flat)
        return
    with open(filename, "wb") as f:  # a file object: numpy appends no suffix
        np.savez(f, **flat)


def read_container(filename):
    """{dataset path: array} with _re/_im pairs joined into complex arrays."""
    if str(filename).endswith(".h5"):
        z = h5.read(str(filename))
    else:
        with np.load(filename, allow_pickle=False) as f:
            z = {k: f[k] for k in f.keys()}
    out = {}
    for k in z:
        if k.endswith("_im") and k[:-3] + "_re" in z:
            continue
        if k.endswith("_re") and k[:-3] + "_im" in z:
            out[k[:-3]] = z[k] + 1j * z[k[:-3] + "_im"]
        else:
            out[k] = z[k]
    return out


def broadcast_2d(old, new):
    """field/read.rs:112-121: copy the overlap [..min, ..min] of `old` into `new`
    (entries of `new` outside the overlap keep their values)."""
    new = np.array(new, copy=True)
    s0, s1 = min(old.shape[0], new.shape[0]), min(old.shape[1], new.shape[1])
    new[:s0, :s1] = old[:s0, :s1]
    return new


def coords(n, periodic, scale=1.0):
    """Grid points of a base times the axis scale: Chebyshev x_i = -sin(pi (m - 2 i) / (2 m)),
    m = n - 1 (funspace chebyshev/ortho.rs:71-80), Fourier 0 + (2 pi / n) i
    (fourier/c2c.rs:63-66); navier.rs:500-515 scales them."""
    i = np.arange(n, dtype=float)
    if periodic:
        x = 0.0 + (2.0 * math.pi / n) * i
    else:
        m = float(n - 1)
        x = -np.sin(math.pi * (m - 2.0 * i) / (2.0 * m))
    return x * scale


def get_dx(x, periodic):
    """Cell widths (field.rs:134-163)."""
    if periodic:
        return np.full(len(x), x[2] - x[1])
    dx = np.zeros(len(x))
    for i in range(len(x)):
        left = x[0] if i == 0 else (x[i] + x[i - 1]) / 2.0
        right = x[-1] if i == len(x) - 1 else (x[i + 1] + x[i]) / 2.0
        dx[i] = right - left
    return dx


def flow_filename(time):
    """data/flow{:0>8.2}.h5 of navier.rs:782 (the .npz suffix on hosts without libhdf5)."""
    return "flow%s%s" % (format(time, "0>8.2f"), _SUFFIX)


def write_navier(nav, filename):
    """Navier2D::write (navier.rs:974-1008)."""
    groups = (("temp", nav.temp), ("ux", nav.ux), ("uy", nav.uy), ("pres", nav.pres[0]))
    data = {}
    for g, f in groups:
        f.backward()
        v = f.v
        if g == "temp":
            v = v + nav.fieldbc.v
        data[g + "/v"] = v
        data[g + "/vhat"] = f.vhat
    # field.rs:134-163 forms dx from the unscaled points; navier.rs:500-515 scales both
    x0, y0 = coords(nav.nx, nav.periodic), coords(nav.ny, False)
    data.update({"x": x0 * nav.scale[0], "dx": get_dx(x0, nav.periodic) * nav.scale[0],
                 "y": y0 * nav.scale[1], "dy": get_dx(y0, False) * nav.scale[1]})
    if getattr(nav, "solid", None) is not None:  # navier.rs:995-998
        data["solid/mask"] = np.asarray(nav.solid[0], dtype=float)
    data.update({"time": nav.get_time(), "ra": nav.ra, "pr": nav.pr, "nu": nav.nu, "kappa": nav.ka})
    write_container(filename, data)


def read_navier(nav, filename):
    """Navier2D::read (navier.rs:961-971)."""
    data = read_container(filename)
    for g, f in (("temp", nav.temp), ("ux", nav.ux), ("uy", nav.uy), ("pres", nav.pres[0])):
        x = data[g + "/vhat"]
        cur = f.vhat
        if x.shape != cur.shape:
            print("Attention! Broadcast from shape %s to shape %s." % (x.shape, cur.shape))
            x = broadcast_2d(x, cur)
        f.vhat = x
        f.backward()
    nav.time = float(data["time"][0])

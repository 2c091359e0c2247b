"""The oracle's solid-obstacle terms (navier.rs:551-560, :579-584, :603-608) pinned without
the reference, and the mask builders (src/navier/solid_masks.rs:34-139).

* A mask whose value field equals the physical boundary temperature leaves a step from the
  conduction state (u = 0, T = 0) exactly as the step without a mask: the temperature damp is
  mask (T + T_bc - value) = 0 and the velocity damps vanish; a wrong sign of the value or a
  dropped boundary field fails it.
* A uniform mask (= 1) is a linear relaxation of the velocity at rate 1/eta = 100: at
  vanishing amplitude one step from u equals the step without a mask from (1 - 100 dt) u,
  so eta, the sign and the product with the mask are fixed.
"""
import math

import numpy as np
import pytest

from helpers import rel_err
from rustpde_amd.checkpoint import coords, get_dx
from rustpde_amd.solid_masks import solid_cylinder_inner, solid_porosity, solid_roughness_sinusoid


def _new(oracle, periodic, nx, ny, dt=0.01):
    if periodic:
        return oracle.Navier2D.new_periodic(nx, ny, 1e4, 1.0, dt, 2.0)
    return oracle.Navier2D.new(nx, ny, 1e4, 1.0, dt, 1.0, True)


@pytest.mark.parametrize("periodic,nx,ny", [(False, 16, 16), (True, 16, 17)])
def test_solid_at_boundary_temperature_is_inert(oracle, periodic, nx, ny):
    a, b = _new(oracle, periodic, nx, ny), _new(oracle, periodic, nx, ny)
    x = coords(nx, periodic, 2.0 if periodic else 1.0)
    y = coords(ny, False, 1.0)
    mask = solid_cylinder_inner(x, y, x[nx // 2], 0.0, 0.4)[0]
    a.set_solid([mask, a.get("fieldbc", False)])
    a.update(1)  # one step: its damps read the step-start fields (u = 0, T = 0)
    b.update(1)
    for f in ("temp", "ux", "uy", "pres"):
        assert np.array_equal(a.get(f), b.get(f)), f
    # a value field of the opposite sign drives the temperature
    c = _new(oracle, periodic, nx, ny)
    c.set_solid([mask, -c.get("fieldbc", False)])
    c.update(1)
    assert np.abs(c.get("temp")).max() > 1e-3


@pytest.mark.parametrize("periodic,nx,ny", [(False, 20, 20), (True, 16, 17)])
def test_uniform_mask_relaxes_velocity(oracle, periodic, nx, ny):
    dt, eps = 1e-3, 1e-6
    rng = np.random.default_rng(7)
    vel = []
    for f in ("ux", "uy"):  # low modes only: the 2/3 dealiasing leaves them untouched
        shp = _new(oracle, periodic, nx, ny).get(f).shape
        v = np.zeros(shp, dtype=complex if periodic else float)
        v[:4, :4] = rng.standard_normal((4, 4)) + (1j * rng.standard_normal((4, 4)) if periodic else 0.0)
        if periodic:
            v[0] = v[0].real  # a real field: the k = 0 row is real (c2r drops Im X_0)
        vel.append(v)

    def run(scale, solid):
        o = _new(oracle, periodic, nx, ny, dt)
        for f, v in zip(("ux", "uy"), vel):
            o.set(f, eps * scale * v)
        if solid:
            o.set_solid([np.ones((nx, ny)), o.get("fieldbc", False)])
        o.update(1)
        return [o.get(f) for f in ("ux", "uy")]

    a = run(1.0, True)                  # mask = 1: damp -100 u in the convective term
    b = run(1.0 - dt / 1e-2, False)     # no mask, velocity (1 - dt / eta) u
    c = run(1.0, False)                 # no mask, velocity u
    for fa, fb, fc in zip(a, b, c):
        d_ab, d_cb = np.abs(fa - fb).max(), np.abs(fc - fb).max()
        assert d_cb > 1e-9 and d_ab < 1e-4 * d_cb, (d_ab, d_cb)


def test_mask_builders():
    x = coords(33, False, 1.0)
    y = coords(33, False, 1.0)
    m, v = solid_cylinder_inner(x, y, 0.0, 0.0, 0.5)
    assert m[16, 16] == 1.0 and m[0, 0] == 0.0 and not v.any()
    r = np.sqrt(x[:, None] ** 2 + y[None, :] ** 2)
    assert np.all(m[r < 0.45] == 1.0) and np.all(m[r >= 0.55] == 0.0)
    band = (r >= 0.45) & (r < 0.55)
    assert np.allclose(m[band], 0.5 * (1.0 - np.tanh(2.0 * (r[band] - 0.5) / 0.05)))
    xr = coords(32, True, 1.0)
    m, v = solid_roughness_sinusoid(xr, y, 0.1, 10.0)
    up = np.sin(10.0 * xr) + 0.5 >= 0.0  # roughness height >= 0: the plates are solid there
    assert np.all(m[up, 0] == 1.0) and np.all(v[up, 0] == 0.5) and np.all(v[up, -1] == -0.5)
    assert np.all(m[~up, 0] < 1.0) and not m[:, 16].any()
    m, _ = solid_porosity(x, y, 0.3, 0.8)
    n = math.floor(math.sqrt(0.2 * 4.0 * 4.0 / (math.pi * 0.09)) + 0.5)
    w = get_dx(x, False)[:, None] * get_dx(y, False)[None, :] / 4.0
    area = float((m * w).sum())  # 9 disks of diameter 0.3 in the 2 x 2 box: 0.159
    assert n == 3 and m.max() <= 1.0 + 1e-15 and abs(area - 9 * math.pi * 0.15 ** 2 / 4.0) < 0.02, area

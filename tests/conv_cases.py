"""Closed-form cases that exercise the convective terms of Navier2D::update
(navier.rs:538-616, conv_term.rs:22-42, dealias navier.rs:1022-1032).

Each case sets physical fields whose one-step effect is known from the continuous equations
and whose discretisation is exact or spectrally accurate, so that a flipped sign, a dropped
term (the boundary field's), a wrong x scale or a product with the wrong velocity component
fails the check.  The cases run against any object with the adapter below: the oracle's
Navier2D and the GPU Navier2D.
"""
import numpy as np


def cheb_nodes(n):
    """x_j = -cos(pi j / (n - 1)) (funspace/src/chebyshev/ortho.rs:71-80)."""
    return -np.cos(np.pi * np.arange(n) / (n - 1))


def fourier_nodes(n):
    """x_j = 2 pi j / n (funspace/src/fourier/c2c.rs:63-66), before the aspect scaling."""
    return 2.0 * np.pi * np.arange(n) / n


def grids(nx, ny, periodic):
    """(xh, y): unscaled x coordinate (Chebyshev in [-1, 1], or Fourier in [0, 2 pi)) and y,
    as [nx, ny] arrays."""
    xh = fourier_nodes(nx) if periodic else cheb_nodes(nx)
    y = cheb_nodes(ny)
    return np.meshgrid(xh, y, indexing="ij")


class Adapter:
    """set / get physical fields on the oracle (`orc=True`) or on the GPU Navier2D."""

    def __init__(self, nav, orc):
        self.nav, self.orc = nav, orc

    def set_phys(self, name, v):
        if self.orc:
            self.nav.set_phys_forward(name, v)
        else:
            f = getattr(self.nav, name)
            f.v = v
            f.forward()

    def get_phys(self, name):
        if self.orc:
            return self.nav.backward(name)
        f = getattr(self.nav, name)
        f.backward()
        return f.v

    def update(self, n=1):
        self.nav.update(n)


def make(factory, periodic, nx, ny, aspect, adiabatic=False, ra=1e9, dt=1e-3):
    if periodic:
        return factory.new_periodic(nx, ny, ra, 1.0, dt, aspect)
    return factory.new(nx, ny, ra, 1.0, dt, aspect, adiabatic)


def temp_advection(new_nav, orc, periodic):
    """Temperature advected by a velocity over one step.

    T' is advected by (u, v); with the diffusion-only step (u = v = 0, same T') subtracted,
    T_new - T_diff = -dt (u dT'/dx + v dT'/dy + v dT_bc/dy) up to O(dt ka k^2) relative
    (T_bc = -y/2, so v dT_bc/dy = -v/2: the boundary field's convective term).
    Returns (max abs deviation, max abs of the expected change)."""
    a = 1.5
    nx, ny = (16, 33) if periodic else (40, 40)
    xh, y = grids(nx, ny, periodic)
    if periodic:
        t0 = 0.1 * np.cos(xh) * np.cos(0.5 * np.pi * y)
        dtdx = -0.1 / a * np.sin(xh) * np.cos(0.5 * np.pi * y)
        dtdy = -0.05 * np.pi * np.cos(xh) * np.sin(0.5 * np.pi * y)
        u = 0.7 * (1.0 - y ** 2)
        v = 0.3 * np.cos(xh) * (1.0 - y ** 2) + 0.2 * np.sin(2 * xh) * (1 - y ** 2) ** 2
    else:
        t0 = 0.1 * np.sin(np.pi * xh) * np.sin(np.pi * y)
        dtdx = 0.1 * np.pi / a * np.cos(np.pi * xh) * np.sin(np.pi * y)
        dtdy = 0.1 * np.pi * np.sin(np.pi * xh) * np.cos(np.pi * y)
        u = 0.7 * (1.0 - xh ** 2) * (1.0 - y ** 2) * y
        v = 0.3 * (1.0 - xh ** 2) * (1.0 - y ** 2) * (1.0 + xh)
    out = []
    for vel in (True, False):
        nav = Adapter(new_nav(periodic, nx, ny, a), orc)
        nav.set_phys("temp", t0)
        if vel:
            nav.set_phys("ux", u)
            nav.set_phys("uy", v)
        nav.update(1)
        out.append(nav.get_phys("temp"))
    dt = 1e-3
    expect = -dt * (u * dtdx + v * dtdy + v * (-0.5))
    return float(np.abs((out[0] - out[1]) - expect).max()), float(np.abs(expect).max())


def reynolds_stress_mean_flow(new_nav, orc):
    """The x-mean of ux after one step from a divergence-free wave (periodic).

    Stream function psi = eps [cos(x~) f1(y) + sin(x~) f2(y)], x~ = x / aspect,
    f1 = (1 - y^2)^2, f2 = y (1 - y^2)^2:  u = d psi/dy, v = -d psi/dx vanish with the wall.
    The x-mean of u.grad(u) is d/dy <u v> = d/dy [eps^2 (1 - y^2)^4 / (2 a)], the mean pressure
    gradient along x is zero and the projection does not touch k_x = 0 of ux, so the mean
    flow after one step is U(y) = 4 dt eps^2 y (1 - y^2)^3 / a  (to O(dt nu k^2) relative).
    Every product is a polynomial of degree <= 9 in y times modes 0, 2 in x: exact below the
    2/3 cut.  Returns (max abs deviation, max abs of U)."""
    a, eps, dt = 1.5, 0.2, 1e-3
    nx, ny = 16, 33
    xh, y = grids(nx, ny, True)
    f1, f1p = (1 - y ** 2) ** 2, -4 * y * (1 - y ** 2)
    f2, f2p = y * (1 - y ** 2) ** 2, (1 - y ** 2) * (1 - 5 * y ** 2)
    u = eps * (np.cos(xh) * f1p + np.sin(xh) * f2p)
    v = eps / a * (np.sin(xh) * f1 - np.cos(xh) * f2)
    nav = Adapter(new_nav(True, nx, ny, a), orc)
    nav.set_phys("ux", u)
    nav.set_phys("uy", v)
    nav.update(1)
    mean = nav.get_phys("ux").mean(axis=0)
    expect = 4 * dt * eps ** 2 * y[0] * (1 - y[0] ** 2) ** 3 / a
    return float(np.abs(mean - expect).max()), float(np.abs(expect).max())

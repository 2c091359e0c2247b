"""Pins the oracle's convective terms (oracle/rbc_oracle.cpp conv_term, finish_conv,
conv_temp, conv_vel) without a second restatement:

* conv_term against the analytic products of conv_term.rs:64-119, tightened from the
  reference's 1e-3 to spectral accuracy, in the confined and the periodic space, with an
  x scale (aspect 1.5) so that a missing 1/scale fails;
* dealias (navier.rs:1022-1032) against closed-form spectra: products of Chebyshev and
  Fourier modes whose spectra straddle the cut floor(2 n / 3) in both directions, so that
  an off-by-one cut or a cut along the wrong axis fails;
* conv_temp / conv_ux / conv_uy (navier.rs:538-616) against the analytic u d/dx + v d/dy of
  each field (and the boundary field's terms for temp), through the pinned forward + dealias;
* update() invariants that run the convective terms (tests/conv_cases.py): a temperature
  advected over one step, and the mean flow forced by a wave's Reynolds stress.
"""
import numpy as np
import pytest

import conv_cases as C


def _orc(oracle):
    def new(periodic, nx, ny, a, adiabatic=False):
        return C.make(oracle.Navier2D, periodic, nx, ny, a, adiabatic)
    return new


@pytest.mark.parametrize("periodic", [False, True])
def test_conv_term_analytic(oracle, periodic):
    """conv_term.rs:64-119: temp = cos . sin, u = sin . sin; u dT/dx and u dT/dy."""
    a = 1.5
    nx, ny = (32, 33) if periodic else (32, 32)
    o = C.make(oracle.Navier2D, periodic, nx, ny, a, adiabatic=True)
    xh, y = C.grids(nx, ny, periodic)
    if periodic:  # temp: Fourier x cheb_dirichlet
        t = np.cos(2 * xh) * np.sin(np.pi * y)
        dtdx = -2.0 / a * np.sin(2 * xh) * np.sin(np.pi * y)
        dtdy = np.pi * np.cos(2 * xh) * np.cos(np.pi * y)
        u = np.sin(3 * xh) * np.sin(np.pi * y)
    else:  # temp: cheb_neumann x cheb_dirichlet (adiabatic side walls)
        t = np.cos(np.pi * xh) * np.sin(np.pi * y)
        dtdx = -np.pi / a * np.sin(np.pi * xh) * np.sin(np.pi * y)
        dtdy = np.pi * np.cos(np.pi * xh) * np.cos(np.pi * y)
        u = np.sin(np.pi * xh) * np.sin(np.pi * y)
    o.set_phys_forward("temp", t)
    cx = o.conv_term("temp", (1, 0), u)
    cy = o.conv_term("temp", (0, 1), u)
    assert np.abs(cx - u * dtdx).max() < 1e-11
    assert np.abs(cy - u * dtdy).max() < 1e-11
    # the same for a velocity field (cheb_dirichlet x cheb_dirichlet / Fourier x cheb_dirichlet)
    o.set_phys_forward("ux", u)
    du = (3.0 / a * np.cos(3 * xh) if periodic else np.pi / a * np.cos(np.pi * xh)) * np.sin(np.pi * y)
    assert np.abs(o.conv_term("ux", (1, 0), t) - t * du).max() < 1e-11


def _cheb_mul(p, q):
    """T_p T_q = (T_{p+q} + T_{|p-q|}) / 2 as {index: coefficient}."""
    out = {}
    for k in (p + q, abs(p - q)):
        out[k] = out.get(k, 0.0) + 0.5
    return out


def _four_mul(a, b):
    """product of x-functions given as {k (any sign): complex coefficient of e^{ikx}}"""
    out = {}
    for ka, ca in a.items():
        for kb, cb in b.items():
            out[ka + kb] = out.get(ka + kb, 0.0) + ca * cb
    return out


def _cos(k):
    return {k: 0.5, -k: 0.5} if k else {0: 1.0}


def _sin(k):
    return {k: -0.5j, -k: 0.5j}


def test_dealias_closed_form_confined(oracle):
    """Spectrum of (T8(x)[T1 + T2](y)) * ([T7 + T8](x) T14(y)) in chebyshev(24)^2: rows and
    columns 15 are kept, 16 are zeroed (cut floor(2 * 24 / 3) = 16)."""
    n = 24
    o = oracle.Navier2D.new(n, n, 1e5, 1.0, 0.01, 1.0, True)
    xs = C.cheb_nodes(n)
    u_terms = [((8, 1), 1.0), ((8, 2), 1.0)]
    w_terms = [((7, 14), 1.0), ((8, 14), 1.0)]
    coef = np.zeros((n, n))
    for (px, py), cu in u_terms:
        for (qx, qy), cw in w_terms:
            for kx, ax in _cheb_mul(px, qx).items():
                for ky, ay in _cheb_mul(py, qy).items():
                    coef[kx, ky] += cu * cw * ax * ay
    phys = np.polynomial.chebyshev.chebgrid2d(xs, xs, coef)
    got = o.finish_conv(phys)
    want = coef.copy()
    want[16:, :] = 0.0
    want[:, 16:] = 0.0
    assert coef[15, 15] != 0 and coef[16, 15] != 0 and coef[15, 16] != 0  # the case straddles the cut
    assert np.abs(got - want).max() < 1e-13
    assert np.all(got[16:, :] == 0.0) and np.all(got[:, 16:] == 0.0)


def test_dealias_closed_form_periodic(oracle):
    """Spectrum of cos(8x)[T1 + T3](y) * [cos(7x) + sin(8x)][T19 + T20](y) in
    fourier_r2c(48) x chebyshev(33): x rows < 16 and y columns < 22 kept (25 * 2 / 3 = 16,
    33 * 2 / 3 = 22); the r2c coefficient of e^{ikx} is n times its Fourier coefficient."""
    nx, ny = 48, 33
    o = oracle.Navier2D.new_periodic(nx, ny, 1e5, 1.0, 0.01, 2.0)
    xh, y = C.fourier_nodes(nx), C.cheb_nodes(ny)
    u_terms = [(_cos(8), 1), (_cos(8), 3)]
    w_terms = [(_cos(7), 19), (_cos(7), 20), (_sin(8), 19), (_sin(8), 20)]
    coef = np.zeros((nx // 2 + 1, ny), dtype=complex)
    phys = np.zeros((nx, ny))
    ex = np.exp(1j * np.outer(xh, np.arange(-nx, nx + 1)))  # e^{ikx} columns, k = -nx..nx
    for fu, py in u_terms:
        for fw, qy in w_terms:
            fx = _four_mul(fu, fw)
            for ky, ay in _cheb_mul(py, qy).items():
                tq = np.polynomial.chebyshev.chebval(y, np.eye(ny)[ky])
                for k, c in fx.items():
                    phys += np.real(np.outer(ex[:, k + nx] * c * ay, tq))
                    if k >= 0:
                        coef[k, ky] += nx * c * ay
    got = o.finish_conv(phys)
    want = coef.copy()
    want[16:, :] = 0.0
    want[:, 22:] = 0.0
    assert coef[15, 21] != 0 and coef[16, 21] != 0 and coef[15, 22] != 0
    assert np.abs(got - want).max() < 1e-12 * nx
    assert np.all(got[16:, :] == 0.0) and np.all(got[:, 22:] == 0.0)


@pytest.mark.parametrize("periodic", [False, True])
def test_conv_fields_analytic(oracle, periodic):
    """conv_temp / conv_ux / conv_uy = dealias(forward(u dq/dx + v dq/dy (+ v dT_bc/dy)))
    with the analytic derivatives (T_bc = -y/2 in both forms)."""
    a = 1.5
    nx, ny = (32, 33) if periodic else (40, 40)
    o = C.make(oracle.Navier2D, periodic, nx, ny, a, adiabatic=True)
    xh, y = C.grids(nx, ny, periodic)
    s, c = np.sin, np.cos
    if periodic:
        ux, uy = 0.3 * s(xh) * s(np.pi * y), 0.2 * c(2 * xh) * s(np.pi * y) ** 2
        uxx, uxy = 0.3 / a * c(xh) * s(np.pi * y), 0.3 * np.pi * s(xh) * c(np.pi * y)
        uyx = -0.4 / a * s(2 * xh) * s(np.pi * y) ** 2
        uyy = 0.2 * c(2 * xh) * np.pi * s(2 * np.pi * y)
        t = 0.1 * c(xh) * s(np.pi * y)
        tx, ty = -0.1 / a * s(xh) * s(np.pi * y), 0.1 * np.pi * c(xh) * c(np.pi * y)
    else:
        px = np.pi * xh
        ux, uy = 0.3 * s(px) * s(np.pi * y), 0.2 * s(2 * px) * s(np.pi * y) ** 2
        uxx, uxy = 0.3 * np.pi / a * c(px) * s(np.pi * y), 0.3 * np.pi * s(px) * c(np.pi * y)
        uyx = 0.4 * np.pi / a * c(2 * px) * s(np.pi * y) ** 2
        uyy = 0.2 * s(2 * px) * np.pi * s(2 * np.pi * y)
        t = 0.1 * c(px) * s(np.pi * y)  # cheb_neumann in x
        tx, ty = -0.1 * np.pi / a * s(px) * s(np.pi * y), 0.1 * np.pi * c(px) * c(np.pi * y)
    o.set_phys_forward("ux", ux)
    o.set_phys_forward("uy", uy)
    o.set_phys_forward("temp", t)
    cases = {
        "temp": ux * tx + uy * ty + uy * (-0.5),
        "ux": ux * uxx + uy * uxy,
        "uy": ux * uyx + uy * uyy,
    }
    for name, phys in cases.items():
        want = o.finish_conv(phys)
        got = o.conv(name)
        assert np.abs(got - want).max() < 1e-10 * np.abs(want).max(), name


@pytest.mark.parametrize("periodic", [False, True])
def test_update_advects_temperature(oracle, periodic):
    dev, ref = C.temp_advection(_orc(oracle), True, periodic)
    assert ref > 1e-5 and dev < 1e-5 * ref, (dev, ref)


def test_update_reynolds_stress_mean_flow(oracle):
    dev, ref = C.reynolds_stress_mean_flow(_orc(oracle), True)
    assert ref > 1e-6 and dev < 1e-5 * ref, (dev, ref)


"""Pins the oracle's adjoint descent (oracle/rbc_oracle.cpp NavierAdjoint, restating
src/navier/navier_adjoint.rs) on the CPU:

* the adjoint convective terms (navier_adjoint.rs:488-578) against the analytic products of
  each field's multiplier and the residuals' derivatives, through the pinned forward
  transform and dealias;
* the explicit temperature update (:629-648) in closed form: with the fields at rest the new
  temperature is dt ka lap(T1) (a sin-sin residual, an eigenfunction of the Laplacian) or
  dt uy1 (the adjoint buoyancy);
* the residual (:727-750): (NS step - state) / dt_navier, to_ortho'd, against an independent
  Navier2D step of the same state, and the smoothed residual against the Helmholtz operator
  (1 - lap / scale^2) applied to it (sign and weight of the smoother);
* the descent property of the method: the residual norm falls monotonically (a smoother of
  the wrong sign makes the iteration diverge), and the conduction state is a fixed point.
"""
import numpy as np
import pytest

import conv_cases as C
from conftest import rel_err


def _adj(oracle, periodic, adiabatic=True, ra=1e4, dt=0.02):
    nx, ny = (32, 33) if periodic else (40, 40)
    a = 2.0 if periodic else 1.5
    adj = oracle.NavierAdjoint(nx, ny, ra, 1.0, dt, a, adiabatic, periodic)
    return adj, nx, ny, a


def _spec(oracle, periodic, kind0, nx, ny, phys):
    k0 = oracle.FO if periodic else kind0
    return oracle.space_apply(k0, nx, oracle.CD, ny, "forward", phys)


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_conv_terms_analytic(oracle, periodic):
    O = oracle
    adj, nx, ny, a = _adj(O, periodic)
    xh, y = C.grids(nx, ny, periodic)
    s, c, pi = np.sin, np.cos, np.pi
    if periodic:
        X, dX = s(xh), c(xh) / a  # a smooth periodic x factor and its x-derivative
        Xb, dXb = c(2 * xh), -2 * s(2 * xh) / a
        kt = O.FO
    else:
        X, dX = s(pi * xh), pi / a * c(pi * xh)
        Xb, dXb = c(pi * xh), -pi / a * s(pi * xh)  # Neumann in x (adiabatic temp)
        kt = O.CN
    Y, dY = s(pi * y), pi * c(pi * y)
    # fields [0] (multipliers) and residuals [1] (differentiated)
    u0, v0, t0 = 0.3 * X * Y, 0.2 * X * Y ** 2, 0.1 * Xb * Y
    u1, u1x, u1y = 0.5 * X * Y, 0.5 * dX * Y, 0.5 * X * dY
    v1, v1x, v1y = 0.4 * X * Y ** 2, 0.4 * dX * Y ** 2, 0.8 * X * Y * dY
    t1, t1x, t1y = 0.7 * Xb * Y, 0.7 * dXb * Y, 0.7 * Xb * dY
    adj.fields.set_phys_forward("ux", u0)
    adj.fields.set_phys_forward("uy", v0)
    adj.fields.set_phys_forward("temp", t0)
    adj.set_residual("ux", _spec(O, periodic, O.CD, nx, ny, u1))
    adj.set_residual("uy", _spec(O, periodic, O.CD, nx, ny, v1))
    adj.set_residual("temp", _spec(O, periodic, kt, nx, ny, t1))
    tbc = -0.5 * y
    cases = {
        "ux": u0 * u1x + v0 * u1y + u0 * u1x + v0 * v1x + t0 * t1x + tbc * t1x,
        "uy": u0 * v1x + v0 * v1y + u0 * u1y + v0 * v1y + t0 * t1y + tbc * t1y,
        "temp": u0 * t1x + v0 * t1y,
    }
    for name, phys in cases.items():
        want = adj.fields.finish_conv(phys)
        got = adj.conv(name)
        assert np.abs(got - want).max() < 1e-10 * np.abs(want).max(), name


@pytest.mark.parametrize("which", ["diffusion", "buoyancy"])
def test_adjoint_temperature_update_closed_form(oracle, which):
    O = oracle
    adj, nx, ny, a = _adj(O, False, adiabatic=False, ra=1e6)
    xh, y = C.grids(nx, ny, False)
    g = np.sin(np.pi * xh) * np.sin(np.pi * y)
    zero = np.zeros((nx - 2, ny - 2))
    for f in ("temp", "ux", "uy"):
        adj.set_residual(f, zero)
    if which == "diffusion":
        adj.set_residual("temp", _spec(O, False, O.CD, nx, ny, 0.3 * g))
        ka = np.sqrt(8.0 / (1e6 * 1.0))
        want = 0.02 * ka * (-np.pi ** 2 * (1.0 / a ** 2 + 1.0)) * 0.3 * g
    else:
        adj.set_residual("uy", _spec(O, False, O.CD, nx, ny, 0.3 * g))
        want = 0.02 * 0.3 * g
    adj.advance()
    got = adj.fields.backward("temp")
    assert np.abs(got - want).max() < 1e-9 * np.abs(want).max()


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_residual_definition(oracle, periodic):
    """the unsmoothed residual is to_ortho((NS step - state) / 1e-2) of the state at the start
    of the update; the smoothed one is -(1 - lap / scale^2)^-1 of it"""
    O = oracle
    adj, nx, ny, a = _adj(O, periodic)
    adj.fields.set_temperature(0.5, 1.0, 1.0)
    adj.fields.set_velocity(0.5, 1.0, 1.0)
    adj.update(3)
    state = {f: adj.fields.get(f) for f in ("temp", "ux", "uy")}
    p_nav = adj.navier.get("pres")
    adj.update(1)
    ref = O.Navier2D(nx, ny, 1e4, 1.0, 1e-2, a, True, periodic)
    for f in state:
        ref.set(f, state[f])
    ref.set("pres", p_nav)
    ref.update(1)
    kx = {"temp": O.FO if periodic else O.CN, "ux": O.FO if periodic else O.CD, "uy": O.FO if periodic else O.CD}
    for f in ("ux", "uy", "temp"):
        r = (ref.get(f) - state[f]) / 1e-2
        want = O.space_apply(kx[f], nx, O.CD, ny, "to_ortho", r)
        assert rel_err(adj.unsmoothed(f), want) < 1e-13, f
        # smoother: (I - c lap) s = -unsmoothed with c = (1 / a^2, 1) (Hholtz, hholtz.rs)
        sm = O.solver_apply("hholtz", kx[f], nx, O.CD, ny, [1.0 / a ** 2, 1.0], -want)
        assert rel_err(adj.residual(f), sm) < 1e-12, f


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_descent_and_fixed_point(oracle, periodic):
    O = oracle
    adj, nx, ny, a = _adj(O, periodic)
    adj.fields.set_temperature(0.5, 1.0, 1.0)
    adj.fields.set_velocity(0.5, 1.0, 1.0)
    res = []
    for _ in range(20):
        adj.update(5)
        res.append(adj.norms()[:3].sum())
        assert not adj.exit()
    assert np.all(np.diff(res) < 0), res
    # the conduction state (fields zero, fieldbc only, hydrostatic pressure p = -y^2 / 4 =
    # -(T_0 + T_2) / 8 in the Navier-Stokes step) has no residual and stays put
    rest, nx, ny, _ = _adj(O, periodic)
    p = np.zeros((nx // 2 + 1, ny), dtype=complex) if periodic else np.zeros((nx, ny))
    p[0, 0] = p[0, 2] = -(nx if periodic else 1.0) / 8.0  # r2c coefficients carry nx
    rest.navier.set("pres", p)
    rest.update(2)
    assert rest.norms().max() < 1e-12 and rest.exit()
    for f in ("temp", "ux", "uy"):
        assert np.abs(rest.fields.get(f)).max() < 1e-12

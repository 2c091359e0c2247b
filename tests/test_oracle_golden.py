"""Pins Navier2D::update of the CPU oracle to the numpy-restatement fixtures (tests/golden).

No reference test exercises update() (the only step-level code is the assertion-free
doc example src/navier/navier.rs:139-152), so the step is pinned by an independent
restatement (numpy + scipy pocketfft + LAPACK), the fixtures written by
tests/golden/make_golden.py.  With the fixture's set-up loaded into the oracle
(fieldbc.vhat and the Poisson eigen-decomposition, the reference's pub fields), the two
restatements differ only by transform rounding: 1e-12 relative for temp/ux/uy/pres.
Two quantities go through the k_x = 0 (or x-eigenmode 0) Poisson row, which the
reference solves with its 1e-10 singular shift (poisson.rs:80-83) and which therefore
amplifies rounding: the pseudo pressure pres[1] (held to 1e-10) and, while the velocity
is still ~1e-5 after one step from rest, the periodic uy (held to 1e-10 at step 1).
"""
import os

import numpy as np
import pytest

from conftest import rel_err

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _load(name):
    return dict(np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False))


@pytest.mark.parametrize("name", ["confined_64x64_ra1e5", "periodic_64x65_ra1e6"])
def test_update_matches_numpy_restatement(oracle, name):
    import make_golden as M
    nx, ny, ra, pr, dt, aspect, periodic, amp, seed, kind, vel = M.CASES[name]
    fx = _load(name)
    o = oracle.Navier2D(nx, ny, ra, pr, dt, aspect, True, periodic)
    o.set("fieldbc", fx["fieldbc"])
    if "poisson_fwd" in fx:
        o.set_tensor(3, fx["poisson_lam"], fx["poisson_fwd"], fx["poisson_bwd"])
    else:
        np.testing.assert_array_equal(o.tensor_setup(3)[0], fx["poisson_lam"])
    if vel is not None:
        o.set_velocity(*vel)
    o.set_phys_forward("temp", fx["pert"])
    for step in range(1, 11):
        o.update()
        if step in (1, 10):
            for f in ("temp", "ux", "uy", "pres"):
                tol = 1e-10 if (periodic and step == 1 and f == "uy") else 1e-12
                assert rel_err(o.get(f), fx["%s_%d" % (f, step)]) < tol, (f, step)
            assert rel_err(o.get("pseu"), fx["pseu_%d" % step]) < 1e-10
    for key, fn in (("nu_10", o.eval_nu), ("nuvol_10", o.eval_nuvol), ("re_10", o.eval_re)):
        assert abs(fn() - float(fx[key])) <= 1e-12 * max(1.0, abs(float(fx[key]))), key


def test_independent_setup_conditioning(oracle):
    """Without sharing the set-up, two faithful restatements of the confined case differ
    at the 1e-11 level: eigenvalue rounding of inv(C)A and DCT noise of fieldbc amplified
    by d2/dx2 (both inherent in the reference's algorithm).  Bound it."""
    import make_golden as M
    name = "confined_64x64_ra1e5"
    nx, ny, ra, pr, dt, aspect, periodic, amp, seed, kind, vel = M.CASES[name]
    fx = _load(name)
    o = oracle.Navier2D(nx, ny, ra, pr, dt, aspect, True, periodic)
    o.set_velocity(*vel)
    o.set_phys_forward("temp", fx["pert"])
    o.update(10)
    for f in ("temp", "ux", "uy", "pres"):
        assert rel_err(o.get(f), fx["%s_10" % f]) < 1e-9, f


def hydrostatic_pressure(nx, ny, periodic):
    """Ortho coefficients of p = -y^2/4 = -(T_0 + T_2)/8 (dp/dy balances T_bc = -y/2).
    Periodic: r2c coefficients are unnormalised, so the k_x = 0 entry carries nx."""
    if periodic:
        p = np.zeros((nx // 2 + 1, ny), dtype=complex)
        p[0, 0] = p[0, 2] = -nx / 8.0
    else:
        p = np.zeros((nx, ny))
        p[0, 0] = p[0, 2] = -1.0 / 8.0
    return p


def test_conduction_state_is_fixed_point(oracle):
    """u = 0, T = T_bc with the hydrostatic pressure is a steady state of update():
    convection vanishes, d2/dy2 of -y/2 is zero and -dt dp/dy + dt T_bc = 0."""
    for periodic in (False, True):
        nx, ny = (32, 33) if periodic else (24, 24)
        o = oracle.Navier2D(nx, ny, 1e6, 1.0, 0.01, 2.0 if periodic else 1.0, True, periodic)
        o.set("pres", hydrostatic_pressure(nx, ny, periodic))
        o.update(5)
        for f in ("temp", "ux", "uy"):
            assert np.abs(o.get(f)).max() < 1e-13, (periodic, f)


def test_mirror_symmetry_preserved(oracle):
    """The mirror class x -> -x (T, uy even; ux odd) is invariant under the confined step.
    set_velocity(a, 2, 1): ux = a sin(2 pi x~) cos(pi y~) is odd about the centre."""
    nx = ny = 24
    o = oracle.Navier2D(nx, ny, 1e5, 1.0, 0.01, 1.0, True, False)
    o.set_velocity(0.2, 2.0, 1.0)
    x = np.linspace(0, 1, nx)
    pert = np.zeros((nx, ny))
    pert[1:-1, 1:-1] = 0.01 * np.outer(np.cos(np.pi * 2 * x[1:-1]), np.sin(np.pi * np.linspace(0, 1, ny)[1:-1]))
    pert = 0.5 * (pert + pert[::-1, :])
    o.set_phys_forward("temp", pert)
    o.update(5)
    t = o.backward("temp")
    u = o.backward("ux")
    v = o.backward("uy")
    assert np.abs(t - t[::-1, :]).max() < 1e-12 * max(1.0, np.abs(t).max())
    assert np.abs(u + u[::-1, :]).max() < 1e-12 * max(1.0, np.abs(u).max())
    assert np.abs(v - v[::-1, :]).max() < 1e-12 * max(1.0, np.abs(v).max())

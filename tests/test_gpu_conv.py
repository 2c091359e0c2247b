"""The GPU step's convective terms against closed forms (tests/conv_cases.py), independent of
the oracle: a temperature advected over one step (the sign and x scale of u.grad(T) and the
boundary field's term v dT_bc/dy) in both forms and in both precisions, and the mean flow
forced by a wave's Reynolds stress (u.grad(u) of the periodic step).  And Integrate::exit on a
NaN written into the state from the host (navier.rs:855-862, src/lib.rs:182-185).
"""
import numpy as np
import pytest

import conv_cases as C

pytestmark = pytest.mark.gpu


def _gpu(mixed):
    from rustpde_amd import Navier2D

    def new(periodic, nx, ny, a, adiabatic=False):
        nav = C.make(Navier2D, periodic, nx, ny, a, adiabatic)
        if mixed:
            nav.set_option(nav.OPT_MIXED, 1)
        return nav
    return new


@pytest.mark.parametrize("mixed", [False, True])
@pytest.mark.parametrize("periodic", [False, True])
def test_gpu_update_advects_temperature(periodic, mixed):
    print("case GPU temperature advection periodic=%s mixed=%s" % (periodic, mixed), flush=True)
    dev, ref = C.temp_advection(_gpu(mixed), False, periodic)
    print("deviation %.3e of %.3e" % (dev, ref), flush=True)
    tol = 1e-4 if mixed else 1e-5
    assert ref > 1e-5 and dev < tol * ref, (dev, ref)


@pytest.mark.parametrize("mixed", [False, True])
def test_gpu_update_reynolds_stress_mean_flow(mixed):
    print("case GPU Reynolds-stress mean flow mixed=%s" % mixed, flush=True)
    dev, ref = C.reynolds_stress_mean_flow(_gpu(mixed), False)
    print("deviation %.3e of %.3e" % (dev, ref), flush=True)
    tol = 1e-4 if mixed else 1e-5
    assert ref > 1e-6 and dev < tol * ref, (dev, ref)


@pytest.mark.parametrize("periodic", [False, True])
def test_exit_on_nan_written_from_host(periodic):
    """A NaN written into ux.vhat from the host makes exit() true (the divergence norm is
    NaN); integrate() then stops after the first step."""
    from rustpde_amd import Navier2D, integrate
    print("case exit on NaN periodic=%s" % periodic, flush=True)
    if periodic:
        nav = Navier2D.new_periodic(32, 33, 1e5, 1.0, 0.01, 2.0)
    else:
        nav = Navier2D.new(24, 24, 1e5, 1.0, 0.01, 1.0, True)
    nav.set_velocity(0.1, 1.0, 1.0)
    nav.update(2)
    assert nav.exit() is False
    a = nav.ux.vhat
    a.flat[3] = np.nan
    nav.ux.vhat = a
    assert nav.exit() is True
    t0 = nav.get_time()
    integrate(nav, t0 + 1.0, None)
    assert abs(nav.get_time() - (t0 + nav.get_dt())) < 1e-12  # stopped after one step
    assert nav.exit() is True
    nav.close()

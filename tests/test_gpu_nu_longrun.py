"""Time-averaged Nusselt number over 200 time units (the long-run bar: agreement within 0.5 %).

* confined C1 (64 x 64, Ra = 1e5, Pr = 1, aspect 1, adiabatic, dt = 0.02, set_velocity(0.2, 1, 1)
  and the seeded uniform perturbation): the FP64 GPU step against the oracle, Nu sampled once
  per time unit as integrate() does with save_intervall = 1 (src/lib.rs:155-187);
* periodic 128 x 129, Ra = 1e5, aspect 2, dt = 0.01, and C1: the mixed-precision step against
  the FP64 step.  Both cases settle into steady convection rolls, so their averages are deterministic
  (a chaotic case would only agree statistically).
"""
import numpy as np
import pytest

from helpers import init_pair

pytestmark = pytest.mark.gpu

MIXED = 7


def _nu_series(nav, dt, t_end, every=1.0):
    steps = int(round(every / dt))
    out = []
    for _ in range(int(round(t_end / every))):
        nav.update(steps)
        out.append(nav.eval_nu())
    return np.array(out)


def test_nu_200_confined_gpu_vs_oracle(oracle):
    from rustpde_amd import Navier2D
    print("case Nu 200 t.u. confined 64x64", flush=True)
    nx = ny = 64
    dt = 0.02
    g = Navier2D.new(nx, ny, 1e5, 1.0, dt, 1.0, True)
    o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, dt, 1.0, True)
    init_pair(g, o, nx, ny, False, 0.05, 1, "uniform", velocity=(0.2, 1.0, 1.0))
    ng = _nu_series(g, dt, 200.0)
    no = _nu_series(o, dt, 200.0)
    mg, mo = ng.mean(), no.mean()
    print("Nu mean gpu %.10f oracle %.10f (last %.6f / %.6f)" % (mg, mo, ng[-1], no[-1]), flush=True)
    assert np.all(np.isfinite(ng))
    assert abs(mg - mo) < 5e-3 * abs(mo)
    assert abs(mg - mo) < 1e-9 * abs(mo)  # in fact the FP64 paths agree to rounding
    assert mo > 1.5  # convection, not the conductive state


def test_nu_200_periodic_mixed_vs_fp64(oracle):
    from rustpde_amd import Navier2D
    print("case Nu 200 t.u. periodic 128x129 mixed vs fp64", flush=True)
    nx, ny, dt = 128, 129, 0.01
    a = Navier2D.new_periodic(nx, ny, 1e5, 1.0, dt, 2.0)
    b = Navier2D.new_periodic(nx, ny, 1e5, 1.0, dt, 2.0)
    o = oracle.Navier2D.new_periodic(nx, ny, 1e5, 1.0, dt, 2.0)  # only the shared set-up
    init_pair(a, o, nx, ny, True, 1e-2, 3, "gaussian")
    init_pair(b, o, nx, ny, True, 1e-2, 3, "gaussian")
    b.set_option(MIXED, 1)
    na = _nu_series(a, dt, 200.0)
    nb = _nu_series(b, dt, 200.0)
    ma, mb = na.mean(), nb.mean()
    print("Nu mean fp64 %.10f mixed %.10f (last %.6f / %.6f)" % (ma, mb, na[-1], nb[-1]), flush=True)
    assert np.all(np.isfinite(nb))
    assert abs(ma - mb) < 5e-3 * abs(ma)
    assert ma > 1.5


def test_nu_200_confined_mixed_vs_fp64(oracle):
    """The confined mixed step (FP32 matrix-core GEMMs for the convective terms) against the
    FP64 step over 200 time units of C1."""
    from rustpde_amd import Navier2D
    print("case Nu 200 t.u. confined 64x64 mixed vs fp64", flush=True)
    nx = ny = 64
    dt = 0.02
    a = Navier2D.new(nx, ny, 1e5, 1.0, dt, 1.0, True)
    b = Navier2D.new(nx, ny, 1e5, 1.0, dt, 1.0, True)
    o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, dt, 1.0, True)  # only the shared set-up
    init_pair(a, o, nx, ny, False, 0.05, 1, "uniform", velocity=(0.2, 1.0, 1.0))
    init_pair(b, o, nx, ny, False, 0.05, 1, "uniform", velocity=(0.2, 1.0, 1.0))
    b.set_option(MIXED, 1)
    na = _nu_series(a, dt, 200.0)
    nb = _nu_series(b, dt, 200.0)
    ma, mb = na.mean(), nb.mean()
    print("Nu mean fp64 %.10f mixed %.10f (last %.6f / %.6f)" % (ma, mb, na[-1], nb[-1]), flush=True)
    assert np.all(np.isfinite(nb))
    assert abs(ma - mb) < 5e-3 * abs(ma)
    assert ma > 1.5

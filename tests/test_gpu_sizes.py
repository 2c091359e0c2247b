"""Periodic steps at the transform lengths the other step tests leave out, against the oracle
after 10 steps from a moving state (set_velocity(0.05, 2, 1)):

* 512 x 513: x transforms of 512 points (eight per workgroup, at the spread transform
  stride), y transforms of 1024 points, chains of 256;
* 2048 x 129: x transforms of 2048 points (one per workgroup), y transforms of 256 points
  (two rows per workgroup: the guarded row loads).

FP64 within 1e-11 of the FP64 oracle (the 1e-12 of the smaller cases is below the FP64
rounding floor of the longer recurrences, as at 1024 x 1025); mixed precision within 1e-5,
with the one-launch convective stage by its size rule and forced on.
"""
import pytest

from helpers import compare, init_pair

pytestmark = pytest.mark.gpu

SHAPES = [(512, 513), (2048, 129)]


def _pair(oracle, nx, ny):
    from rustpde_amd import Navier2D
    g = Navier2D.new_periodic(nx, ny, 1e8, 1.0, 5e-4, 2.0)
    o = oracle.Navier2D.new_periodic(nx, ny, 1e8, 1.0, 5e-4, 2.0)
    init_pair(g, o, nx, ny, True, 1e-3, 17, "gaussian", velocity=(0.05, 2.0, 1.0))
    return g, o


@pytest.mark.parametrize("shape", SHAPES)
def test_sizes_fp64(oracle, shape):
    nx, ny = shape
    print("case periodic %dx%d FP64, 10 steps" % (nx, ny), flush=True)
    g, o = _pair(oracle, nx, ny)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    print("FP64 %dx%d errors: %s" % (nx, ny, e), flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-11, (k, e)
    assert abs(g.eval_nu() - o.eval_nu()) < 1e-11 * abs(o.eval_nu())
    g.close()


@pytest.mark.parametrize("conv3", [1, 2])
@pytest.mark.parametrize("shape", SHAPES)
def test_sizes_mixed(oracle, shape, conv3):
    nx, ny = shape
    print("case periodic %dx%d mixed conv3=%d, 10 steps" % (nx, ny, conv3), flush=True)
    g, o = _pair(oracle, nx, ny)
    g.set_option(g.OPT_MIXED, 1)
    g.set_option(g.OPT_CONV3, conv3)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    print("mixed %dx%d conv3=%d errors: %s" % (nx, ny, conv3, e), flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    assert max(e[k] for k in ("temp", "ux", "uy")) > 1e-15  # the FP32 path ran
    nu_g, nu_o = g.eval_nu(), o.eval_nu()
    assert abs(nu_g - nu_o) < 1e-5 * abs(nu_o)
    g.close()

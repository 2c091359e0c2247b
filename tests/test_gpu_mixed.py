"""Mixed-precision periodic step (RBC_OPT_MIXED: convective terms in FP32, the rest FP64).

Bar: every field within 1e-5 relative max-norm of the FP64 result after 10 steps, against
both the FP64 GPU step and the oracle on the same seeded inputs.  The FP32 rounding enters
only through dt * conv, so the measured differences are far below the bar (~1e-9).
"""
import numpy as np
import pytest

from helpers import compare, init_pair, rel_err

pytestmark = pytest.mark.gpu

MIXED = 7


def _periodic(oracle, nx, ny, ra, dt, seed):
    from rustpde_amd import Navier2D
    g = Navier2D.new_periodic(nx, ny, ra, 1.0, dt, 2.0)
    o = oracle.Navier2D.new_periodic(nx, ny, ra, 1.0, dt, 2.0)
    init_pair(g, o, nx, ny, True, 1e-3, seed, "gaussian", velocity=(0.05, 2.0, 1.0))
    return g, o


@pytest.mark.parametrize("shape", [(128, 129), (256, 257)])
def test_mixed_fields_vs_oracle(oracle, shape):
    nx, ny = shape
    print("case mixed vs oracle", shape, flush=True)
    g, o = _periodic(oracle, nx, ny, 1e7, 2e-3, 21)
    g.set_option(MIXED, 1)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    print("mixed errors", e, flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    assert abs(g.eval_nu() - o.eval_nu()) < 1e-5 * abs(o.eval_nu())


def test_mixed_vs_fp64_gpu_and_graph(oracle):
    """mixed against the FP64 GPU step; graph replays reproduce eager mixed steps bit for bit."""
    print("case mixed vs fp64 gpu", flush=True)
    a, _ = _periodic(oracle, 128, 129, 1e6, 4e-3, 5)
    b, _ = _periodic(oracle, 128, 129, 1e6, 4e-3, 5)
    c, _ = _periodic(oracle, 128, 129, 1e6, 4e-3, 5)
    b.set_option(MIXED, 1)
    c.set_option(MIXED, 1)
    c.set_option(0, 0)  # c: eager steps only
    a.update(10)
    b.update(10)
    c.update(10)
    for fa, fb, fc in ((a.temp, b.temp, c.temp), (a.ux, b.ux, c.ux), (a.uy, b.uy, c.uy),
                       (a.pres[0], b.pres[0], c.pres[0])):
        assert rel_err(fb.vhat, fa.vhat) < 1e-5
        assert np.array_equal(fb.vhat, fc.vhat)
    assert rel_err(b.temp.vhat, a.temp.vhat) > 1e-13  # the FP32 path is taken
    # switching back to FP64 continues with the FP64 path
    b.set_option(MIXED, 0)
    a.update(2)
    b.update(2)
    assert rel_err(b.temp.vhat, a.temp.vhat) < 1e-5

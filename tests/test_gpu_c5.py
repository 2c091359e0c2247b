"""The large periodic case, Navier2D::new_periodic(1024, 1025, Ra=1e9, Pr=1, dt=2e-4,
aspect=2), against the oracle at its full size.

FP64: 10 steps from the benchmark's seeded temperature noise with a live velocity field
(set_velocity(0.05, 2, 1)), with the oracle's set-up shared and with each side's own set-up
(the periodic set-up has no eigen-decomposition).  At this size FP64 rounding alone moves the
temperature by more than 1e-12: the FP64 oracle differs from the same algorithm evaluated in
long double (oracle/rbc_oracle_ld.cpp) by ~8e-13 after one step and ~2.5e-12 after ten (the
1023-long recurrences of the row solves amplify rounding).  So every field is compared with
the long double evaluation: within 1e-12, or within twice the FP64 oracle's own distance
from it where that exceeds 1e-12 (the FP64 rounding floor); and within 1e-11 of the FP64
oracle.  Mixed precision (the convective terms in FP32): within 1e-5 of the FP64 oracle after
10 steps from the same moving state, and Nu within 1e-5.
This size runs the longest recurrences (1023-long parity chains of 512) and the
instantiation of the pressure stage for the periodic Rayleigh-Benard bases.
"""
import numpy as np
import pytest

from helpers import compare, init_pair, rel_err

pytestmark = pytest.mark.gpu

NX, NY, RA, PR, DT, ASPECT = 1024, 1025, 1e9, 1.0, 2e-4, 2.0


def _pair(oracle, share):
    from rustpde_amd import Navier2D
    g = Navier2D.new_periodic(NX, NY, RA, PR, DT, ASPECT)
    o = oracle.Navier2D.new_periodic(NX, NY, RA, PR, DT, ASPECT)
    init_pair(g, o, NX, NY, True, 1e-3, 5, "gaussian", velocity=(0.05, 2.0, 1.0), share_setup=share)
    return g, o


@pytest.fixture(scope="module")
def c5_long_double(oracle):
    """fields of the long double evaluation after 10 steps from the same inputs"""
    import os
    from rustpde_amd import seeded
    print("case C5 long double oracle, 10 steps", flush=True)
    ld = oracle.NavierLD.new_periodic(NX, NY, RA, PR, DT, ASPECT, threads=min(16, os.cpu_count() or 1))
    ld.set_velocity(0.05, 2.0, 1.0)
    ld.set_phys_forward("temp", seeded.interior_perturbation(NX, NY, 1e-3, 5, kind="gaussian", periodic=True))
    ld.update(10)
    return {k: ld.get(k) for k in ("temp", "ux", "uy", "pres")}


@pytest.mark.parametrize("share", [True, False])
def test_c5_fp64_10_steps(oracle, c5_long_double, share):
    print("case C5 periodic 1024x1025 FP64 share_setup=%s, 10 steps" % share, flush=True)
    g, o = _pair(oracle, share)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    print("C5 FP64 GPU vs FP64 oracle: %s" % e, flush=True)
    for k, proxy in (("temp", g.temp), ("ux", g.ux), ("uy", g.uy), ("pres", g.pres[0])):
        e_g = rel_err(proxy.vhat, c5_long_double[k])
        e_o = rel_err(o.get(k), c5_long_double[k])
        print("C5 FP64 %s: GPU vs long double %.3e, FP64 oracle vs long double %.3e" % (k, e_g, e_o), flush=True)
        assert e_g < max(1e-12, 2.0 * e_o), (k, e_g, e_o)
        assert e[k] < 1e-11, (k, e)
    assert abs(g.eval_nu() - o.eval_nu()) < 1e-12 * abs(o.eval_nu())
    g.close()


def test_c5_mixed_10_steps(oracle):
    print("case C5 periodic 1024x1025 mixed, 10 steps from a moving state", flush=True)
    g, o = _pair(oracle, True)
    g.set_option(g.OPT_MIXED, 1)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    print("C5 mixed errors: %s" % e, flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    nu_g, nu_o = g.eval_nu(), o.eval_nu()
    print("C5 mixed Nu %.15e oracle %.15e" % (nu_g, nu_o), flush=True)
    assert abs(nu_g - nu_o) < 1e-5 * abs(nu_o)
    assert max(e[k] for k in ("temp", "ux", "uy")) > 1e-15  # the FP32 path ran
    g.close()

"""Mixed-precision steps (RBC_OPT_MIXED: convective terms in FP32, the rest FP64): the
periodic step (FP32 FFT/DCT kernels) and the confined step and ensembles (FP32 matrix-core
GEMMs for the convective terms' dense transforms).

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


def test_mixed_conv3_off_vs_oracle(oracle):
    """the mixed step with the convective y-stage as three launches (RBC_OPT_CONV3 0)
    against the oracle, and against the one-launch form (forced at this size with 2): they
    differ only by the pairing of rows inside the FP32 transforms"""
    print("case mixed conv3 off", flush=True)
    g, o = _periodic(oracle, 128, 129, 1e7, 2e-3, 21)
    h, _ = _periodic(oracle, 128, 129, 1e7, 2e-3, 21)
    g.set_option(MIXED, 1)
    g.set_option(8, 2)
    h.set_option(MIXED, 1)
    h.set_option(8, 0)
    g.update(10)
    h.update(10)
    o.update(10)
    e = compare(h, o)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    for fg, fh in ((g.temp, h.temp), (g.ux, h.ux), (g.uy, h.uy)):
        assert rel_err(fg.vhat, fh.vhat) < 1e-6
    assert rel_err(g.temp.vhat, h.temp.vhat) > 0.0  # the two forms differ by rounding: both ran
    e = compare(g, o)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)


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


def _confined(oracle, nx, ny, ra, dt, seed):
    from rustpde_amd import Navier2D
    g = Navier2D.new(nx, ny, ra, 1.0, dt, 1.0, True)
    o = oracle.Navier2D.new(nx, ny, ra, 1.0, dt, 1.0, True)
    init_pair(g, o, nx, ny, False, 0.05, seed, "uniform", velocity=(0.2, 1.0, 1.0))
    return g, o


@pytest.mark.parametrize("shape", [(64, 64), (48, 65), (96, 80)])
def test_confined_mixed_vs_oracle(oracle, shape):
    """Confined mixed step (the convective terms' dense transforms as FP32 matrix-core
    GEMMs) within 1e-5 of the oracle after 10 steps, and not bitwise the FP64 step."""
    nx, ny = shape
    print("case confined mixed vs oracle", shape, flush=True)
    g, o = _confined(oracle, nx, ny, 1e5, 0.02, 3)
    a, _ = _confined(oracle, nx, ny, 1e5, 0.02, 3)
    g.set_option(MIXED, 1)
    g.update(10)
    o.update(10)
    a.update(10)
    e = compare(g, o)
    print("confined mixed errors", e, flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    assert abs(g.eval_nu() - o.eval_nu()) < 1e-5 * abs(o.eval_nu())
    assert rel_err(g.temp.vhat, a.temp.vhat) > 1e-13  # the FP32 path is taken


def test_confined_mixed_pair_form(oracle):
    """With the banded stencils off (option 3 = 0) the confined step runs the ux/uy pair form
    instead of the field triple; its mixed variant (the FP32 convective GEMMs of the pair
    and of temp) stays within 1e-5 of the oracle and of the triple's mixed step."""
    print("case confined mixed pair form", flush=True)
    g, o = _confined(oracle, 48, 65, 1e5, 0.02, 5)
    t, _ = _confined(oracle, 48, 65, 1e5, 0.02, 5)
    a, _ = _confined(oracle, 48, 65, 1e5, 0.02, 5)
    g.set_option(MIXED, 1)
    g.set_option(3, 0)
    t.set_option(MIXED, 1)
    for nav in (g, o, t, a):
        nav.update(10)
    e = compare(g, o)
    print("confined mixed pair-form errors", e, flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-5, (k, e)
    assert rel_err(g.temp.vhat, t.temp.vhat) < 1e-5
    assert rel_err(g.temp.vhat, a.temp.vhat) > 1e-13  # the FP32 path is taken


def test_confined_mixed_graph_and_switch_back(oracle):
    """Graph replays of the confined mixed step equal eager mixed steps bit for bit; switching
    the option off continues with the FP64 step."""
    print("case confined mixed graph", flush=True)
    a, _ = _confined(oracle, 64, 64, 1e5, 0.02, 4)
    b, _ = _confined(oracle, 64, 64, 1e5, 0.02, 4)
    c, _ = _confined(oracle, 64, 64, 1e5, 0.02, 4)
    b.set_option(MIXED, 1)
    c.set_option(MIXED, 1)
    c.set_option(0, 0)  # c: eager steps only
    a.update(8)
    b.update(8)
    c.update(8)
    for fa, fb, fc in ((a.temp, b.temp, c.temp), (a.ux, b.ux, c.ux), (a.uy, b.uy, c.uy),
                       (a.pres[0], b.pres[0], c.pres[0])):
        assert rel_err(fb.vhat, fa.vhat) < 1e-5
        assert np.array_equal(fb.vhat, fc.vhat)
    b.set_option(MIXED, 0)
    a.update(2)
    b.update(2)
    assert rel_err(b.temp.vhat, a.temp.vhat) < 1e-5


def test_ensemble_mixed_matches_standalone_mixed():
    """An ensemble with the mixed step (rbc_ensemble_set_option) advances every instance as
    a standalone mixed handle does, bit for bit (each batch entry runs the arithmetic of its
    own launch), and within 1e-5 of the FP64 ensemble."""
    from rustpde_amd import Ensemble, Navier2D
    print("case ensemble mixed", flush=True)
    descs = [(48, 48, 1e4 * (1 + 3 * i), 1.0, 0.02, 1.0, True) for i in range(3)] + [(32, 40, 2e4, 1.0, 0.01, 1.0, True)]
    ens = Ensemble(descs)
    ref = Ensemble(descs)
    alone = [Navier2D.new(*d) for d in descs]
    for i in range(len(descs)):
        for h in (ens[i], ref[i], alone[i]):
            h.set_velocity(0.2, 1.0, 1.0)
        alone[i].set_option(MIXED, 1)
    ens.set_option(MIXED, 1)
    ens.update(6)
    ref.update(6)
    for i, h in enumerate(alone):
        h.update(6)
        for fe, fs, fr in ((ens[i].temp, h.temp, ref[i].temp), (ens[i].ux, h.ux, ref[i].ux),
                           (ens[i].uy, h.uy, ref[i].uy)):
            assert np.array_equal(fe.vhat, fs.vhat), i
            assert rel_err(fe.vhat, fr.vhat) < 1e-5, i
        h.close()
    ens.close()
    ref.close()

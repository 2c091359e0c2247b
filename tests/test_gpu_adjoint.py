"""The adjoint descent for steady states (Navier2DAdjoint, src/navier/navier_adjoint.rs) on the
GPU against the oracle's restatement, through the C ABI.

Periodic (Fourier x Chebyshev): no eigen-decomposition anywhere, each side builds its own
set-up; confined: the oracle's set-up (fieldbc, the Poisson and smoother eigen-decompositions,
the Navier-Stokes step's Poisson) given to the GPU handle.  After 10 adjoint iterations the
adjoint fields, the smoothed and unsmoothed residuals and the Navier-Stokes step's fields
agree with the oracle to 1e-12 relative (the pseudo pressure, through the 1e-10 singular
shift, to 1e-7); the residual norms and the exit decision agree; integrate() drives the
adjoint with its callback and the residual falls.
"""
import numpy as np
import pytest

from helpers import rel_err

pytestmark = pytest.mark.gpu


def _pair(oracle, periodic):
    from rustpde_amd import Navier2DAdjoint
    if periodic:
        args = (64, 65, 1e4, 1.0, 0.01, 2.0)
        g = Navier2DAdjoint.new_periodic(*args)
        o = oracle.NavierAdjoint.new_periodic(*args)
    else:
        args = (32, 33, 1e4, 1.0, 0.02, 1.0, True)
        g = Navier2DAdjoint.new(*args)
        o = oracle.NavierAdjoint.new(*args)
        # the oracle's set-up: fieldbc and tensors of both solvers' sets
        g.set_temp_bc(o.fields.get("fieldbc"))
        g.navier.set_temp_bc(o.navier.get("fieldbc"))
        g.navier.set_tensor(3, *o.navier.tensor_setup(3))
        g.set_tensor(3, *o.fields.tensor_setup(3))
        for s in range(3):
            g.set_tensor(s, *o.smoother_tensor(s))
    for x in (g, o.fields):
        x.set_temperature(0.5, 1.0, 1.0)
        x.set_velocity(0.5, 1.0, 1.0)
    return g, o


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_vs_oracle(oracle, periodic):
    print("case adjoint periodic=%s" % periodic, flush=True)
    g, o = _pair(oracle, periodic)
    for it in range(2):
        g.update(5)
        o.update(5)
        assert abs(g.get_time() - o.get_time()) < 1e-12
        errs = {}
        for name, k in (("temp", 0), ("ux", 0), ("uy", 0)):
            errs[name] = rel_err(getattr(g, name)[0].vhat, o.fields.get(name))
            errs[name + "1"] = rel_err(getattr(g, name)[1].vhat, o.residual(name))
            errs[name + "_uns"] = rel_err(getattr(g, name)[1].unsmoothed, o.unsmoothed(name))
            errs["ns_" + name] = rel_err(g.navier.__dict__[name].vhat, o.navier.get(name))
        errs["pres"] = rel_err(g.pres[0].vhat, o.fields.get("pres"))
        errs["pseu"] = rel_err(g.pres[1].vhat, o.fields.get("pseu"))
        errs["ns_pres"] = rel_err(g.navier.pres[0].vhat, o.navier.get("pres"))
        print("adjoint periodic=%s after %d iterations: %s" % (periodic, 5 * (it + 1), errs), flush=True)
        for k, e in errs.items():
            assert e < (1e-7 if k == "pseu" else 1e-12), (k, e)
        ng, no = g.norms(), o.norms()
        assert np.abs(ng - no).max() < 1e-12 * np.abs(no).max(), (ng, no)
        assert g.exit() == o.exit() == False
    for name, f in (("Nu", "eval_nu"), ("Nuvol", "eval_nuvol"), ("Re", "eval_re")):
        a, b = getattr(g, f)(), getattr(o.fields, f)()
        assert abs(a - b) <= 1e-11 * max(abs(b), 1e-3), (name, a, b)
    g.close()


def test_adjoint_integrate_descends(tmp_path):
    """integrate() with the adjoint: callback at the save interval (snapshot, diagnostics,
    residual norms) and a falling residual"""
    from rustpde_amd import Navier2DAdjoint, integrate
    print("case adjoint integrate", flush=True)
    g = Navier2DAdjoint.new(24, 24, 1e4, 1.0, 0.02, 1.0, True)
    g.set_temperature(0.5, 1.0, 1.0)
    g.set_velocity(0.5, 1.0, 1.0)
    g.data_dir = str(tmp_path)
    integrate(g, 1.0, 0.2)
    assert len(g.residuals) == 5 and len(g.diagnostics["Nu"]) == 5
    res = [r[2] for r in g.residuals]
    assert all(b < a for a, b in zip(res, res[1:])), res
    assert sorted(p.name for p in tmp_path.iterdir())[0].startswith("adjoint")
    g.close()

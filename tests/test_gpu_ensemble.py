"""Ensembles of confined instances (rbc_create_ensemble) against standalone handles.

An instance of an ensemble is Navier2D::new with its own parameters; the batched group
step must reproduce, instance by instance, the standalone GPU step (itself held to the
oracle at 1e-12 in test_gpu_navier.py): the same kernels run per instance, so the fields
agree to the last bit, which the test checks at 1e-14.
"""
import numpy as np
import pytest

from helpers import rel_err
from rustpde_amd import Ensemble, Navier2D, RbcError, seeded

pytestmark = pytest.mark.gpu

DESCS = [  # (nx, ny, ra, pr, dt, aspect, adiabatic)
    (24, 24, 1e4, 1.0, 0.02, 1.0, True),
    (32, 40, 3e5, 1.0, 0.01, 1.0, True),
    (24, 24, 1e6, 1.0, 0.02, 1.0, True),
    (24, 24, 3e5, 1.0, 0.02, 1.0, True),
    (32, 40, 1e7, 1.0, 0.01, 1.0, True),
    (20, 28, 1e5, 1.0, 0.02, 1.5, False),
]


def _seed(nav, d, i):
    nav.set_velocity(0.2, 1.0, 1.0)
    nav.temp.v = seeded.interior_perturbation(d[0], d[1], 1e-2 * (1 + i), 100 + i, kind="uniform", periodic=False)
    nav.temp.forward()


def test_ensemble_matches_standalone():
    print("case ensemble vs standalone", flush=True)
    ens = Ensemble(DESCS)
    assert len(ens) == len(DESCS) and ens.groups == 3
    solo = []
    for i, d in enumerate(DESCS):
        _seed(ens[i], d, i)
        s = Navier2D.new(*d)
        _seed(s, d, i)
        solo.append(s)
    for rep in range(2):
        ens.update(5)  # graph capture on the first call, replays afterwards
        for s in solo:
            s.update(5)
        for i, s in enumerate(solo):
            e = ens[i]
            assert abs(e.get_time() - s.get_time()) < 1e-12
            for fe, fs in ((e.temp, s.temp), (e.ux, s.ux), (e.uy, s.uy), (e.pres[0], s.pres[0])):
                assert rel_err(fe.vhat, fs.vhat) < 1e-14, (rep, i)
            assert abs(e.eval_nu() - s.eval_nu()) < 1e-12 * abs(s.eval_nu())
    ens.close()


def test_ensemble_instance_rules():
    print("case ensemble rules", flush=True)
    ens = Ensemble(DESCS[:2])
    with pytest.raises(RbcError):
        ens[0].update(1)  # instances advance with the ensemble
    ens.update(1)
    assert np.isfinite(ens[1].eval_nu())
    ens.close()


def _oracle_c4(oracle, descs, steps):
    """The oracle's fields after `steps` steps of every C4 instance (seeded as the benchmark
    seeds them: set_velocity(0.2, 1, 1), uniform 1e-2 perturbation with seed 1000 + i),
    with each instance's oracle set-up (fieldbc, Poisson eigen-decomposition)."""
    out = []
    for i, d in enumerate(descs):
        o = oracle.Navier2D.new(*d)
        o.set_velocity(0.2, 1.0, 1.0)
        o.set_phys_forward("temp", seeded.interior_perturbation(d[0], d[1], 1e-2, 1000 + i, kind="uniform",
                                                                periodic=False))
        setup = (o.get("fieldbc"), [o.tensor_setup(s) for s in range(4)])
        o.update(steps)
        out.append((setup, {k: o.get(k) for k in ("temp", "ux", "uy", "pres")}, o.eval_nu()))
    return out


def _c4_ensemble(descs, orc, mixed):
    ens = Ensemble(descs)
    for i, d in enumerate(descs):
        nav = ens[i]
        bc, tens = orc[i][0]
        nav.set_temp_bc(bc)
        for s, t in enumerate(tens):
            if t is not None:
                nav.set_tensor(s, *t)
        nav.set_velocity(0.2, 1.0, 1.0)
        nav.temp.v = seeded.interior_perturbation(d[0], d[1], 1e-2, 1000 + i, kind="uniform", periodic=False)
        nav.temp.forward()
    if mixed:
        ens.set_option(Navier2D.OPT_MIXED, 1)
    return ens


def test_c4_workload_vs_oracle(oracle):
    """The C4 workload itself (96 confined instances, N in {64, 96, 128, 192}, Ra
    log-uniform in [1e4, 1e7], dt = 0.01 * 64 / N) against the oracle instance by instance
    after 10 steps: the FP64 ensemble within 1e-12, the mixed ensemble within 1e-5 (fields)
    and 1e-5 (Nu)."""
    import bench
    descs = bench.c4_descs()
    sizes = sorted({d[0] for d in descs})
    assert sizes == [64, 96, 128, 192]
    print("case C4 workload: %d instances, sizes %s, Ra %.2e..%.2e" %
          (len(descs), sizes, min(d[2] for d in descs), max(d[2] for d in descs)), flush=True)
    orc = _oracle_c4(oracle, descs, 10)
    for mixed, tol in ((False, 1e-12), (True, 1e-5)):
        ens = _c4_ensemble(descs, orc, mixed)
        ens.update(10)
        worst = {}
        for i, d in enumerate(descs):
            nav = ens[i]
            for k, proxy in (("temp", nav.temp), ("ux", nav.ux), ("uy", nav.uy), ("pres", nav.pres[0])):
                e = rel_err(proxy.vhat, orc[i][1][k])
                worst[k] = max(worst.get(k, 0.0), e)
                assert e < tol, (mixed, i, d, k, e)
            nu = nav.eval_nu()
            worst["nu"] = max(worst.get("nu", 0.0), abs(nu - orc[i][2]) / abs(orc[i][2]))
            assert abs(nu - orc[i][2]) < max(tol, 1e-12) * abs(orc[i][2]), (mixed, i, nu, orc[i][2])
        print("C4 mixed=%s worst relative errors over the instances: %s" % (mixed, worst), flush=True)
        ens.close()

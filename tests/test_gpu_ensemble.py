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

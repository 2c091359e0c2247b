"""Writes the golden fixtures of Navier2D::update from the numpy restatement.

Run from the repository root:  python tests/golden/make_golden.py
Each .npz holds the seeded initial perturbation, the set-up the restatement computed
(fieldbc.vhat; the Poisson x-eigen-decomposition when axis 0 is Chebyshev, from
LAPACK through numpy.linalg) and temp/ux/uy/pres/pseu spectral coefficients after
1 and 10 steps.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
import np_restatement as R  # noqa: E402
from rustpde_amd import seeded  # noqa: E402

CASES = {
    # name: (nx, ny, ra, pr, dt, aspect, periodic, amp, seed, kind, velocity)
    "confined_64x64_ra1e5": (64, 64, 1e5, 1.0, 0.02, 1.0, False, 0.05, 1, "uniform", (0.2, 1.0, 1.0)),
    "periodic_64x65_ra1e6": (64, 65, 1e6, 1.0, 0.005, 2.0, True, 1e-3, 7, "gaussian", None),
}


def run(name):
    nx, ny, ra, pr, dt, aspect, periodic, amp, seed, kind, vel = CASES[name]
    n = R.Navier2D(nx, ny, ra, pr, dt, aspect, adiabatic=True, periodic=periodic)
    if vel is not None:
        n.set_velocity(*vel)
    pert = seeded.interior_perturbation(nx, ny, amp, seed, kind=kind, periodic=periodic)
    n.temp.v = pert.copy()
    n.temp.forward()
    out = {"pert": pert, "fieldbc": n.fieldbc.vhat}
    t = n.solver[3].t
    out["poisson_lam"] = t.lam
    if t.fwd is not None:
        out["poisson_fwd"], out["poisson_bwd"] = t.fwd, t.bwd
    for step in range(1, 11):
        n.update()
        if step in (1, 10):
            for fname, f in (("temp", n.temp), ("ux", n.ux), ("uy", n.uy), ("pres", n.pres[0]), ("pseu", n.pres[1])):
                out["%s_%d" % (fname, step)] = f.vhat
    out["nu_10"], out["nuvol_10"], out["re_10"] = n.eval_nu(), n.eval_nuvol(), n.eval_re()
    return out


if __name__ == "__main__":
    for name in CASES:
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **run(name))
        print("wrote", name)

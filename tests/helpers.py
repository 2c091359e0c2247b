"""Shared set-up of parity cases (same seeded inputs for the GPU path and the oracle)."""
import numpy as np

from rustpde_amd import seeded

FIELDS = ("temp", "ux", "uy", "pres")


def rel_err(a, b):
    den = max(float(np.abs(b).max()), 1e-300)
    return float(np.abs(a - b).max()) / den


def init_pair(gpu, orc, nx, ny, periodic, amp, seed, kind, velocity=None, share_setup=True):
    """Seed identical initial fields on the GPU Navier2D and the oracle.

    share_setup: give the GPU handle the oracle's fieldbc and tensor eigen-decompositions
    (the reference's pub fieldbc / FdmaTensor fields), so that both run the same set-up.
    """
    if share_setup:
        gpu.set_temp_bc(orc.get("fieldbc"))
        for s in range(4):
            t = orc.tensor_setup(s)
            if t is not None:
                gpu.set_tensor(s, *t)
    if velocity is not None:
        gpu.set_velocity(*velocity)
        orc.set_velocity(*velocity)
    pert = seeded.interior_perturbation(nx, ny, amp, seed, kind=kind, periodic=periodic)
    gpu.temp.v = pert
    gpu.temp.forward()
    orc.set_phys_forward("temp", pert)
    return pert


def compare(gpu, orc):
    out = {}
    for name, proxy in (("temp", gpu.temp), ("ux", gpu.ux), ("uy", gpu.uy), ("pres", gpu.pres[0]),
                        ("pseu", gpu.pres[1])):
        out[name] = rel_err(proxy.vhat, orc.get(name))
    return out

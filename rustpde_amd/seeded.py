"""Seeded synthetic inputs (SplitMix64 counter-based generator, vectorised with numpy).

The reference's initial condition is unseeded (ndarray-rand thread RNG,
src/navier/navier.rs:1079-1090); runs here inject seeded fields instead.  The same
arrays are fed to the GPU path and to the CPU oracle.
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    """n successive SplitMix64 outputs of state `seed` (uint64 array)."""
    with np.errstate(over="ignore"):
        idx = np.arange(1, n + 1, dtype=np.uint64)
        z = np.uint64(seed) + idx * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, shape, lo, hi):
    """U[lo, hi) doubles from the top 53 bits."""
    n = int(np.prod(shape))
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(shape)


def gaussian(seed, shape, sigma):
    """N(0, sigma^2) by Box-Muller on two uniform streams."""
    n = int(np.prod(shape))
    u1 = uniform(seed, (n,), 0.0, 1.0)
    u2 = uniform(seed ^ 0x5DEECE66D, (n,), 0.0, 1.0)
    u1 = np.where(u1 <= 0.0, 1e-300, u1)
    return (sigma * np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)).reshape(shape)


def interior_perturbation(nx, ny, amp, seed, kind="uniform", periodic=False):
    """Perturbation on interior points (zero on the Chebyshev boundaries).

    kind 'uniform': U[-amp, amp];  'gaussian': N(0, amp^2).  For periodic x every x
    point is interior; y boundaries j = 0, ny-1 stay zero.
    """
    i0, i1 = (0, nx) if periodic else (1, nx - 1)
    shape = (i1 - i0, ny - 2)
    p = np.zeros((nx, ny))
    if kind == "uniform":
        p[i0:i1, 1:ny - 1] = uniform(seed, shape, -amp, amp)
    else:
        p[i0:i1, 1:ny - 1] = gaussian(seed, shape, amp)
    return p

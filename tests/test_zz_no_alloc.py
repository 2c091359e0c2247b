"""rbc_solver_solve_noalloc (the solve without any device allocation: work arrays in the
library's fixed per-device scratch) against rbc_solver_solve on the same inputs: the same
results bit for bit, on calls in a row whose sizes shrink (the scratch keeps the previous
call's contents), at the scratch's largest fitting batch of the 512 x 512 Helmholtz batch, and
a batch beyond the scratch rejected with an error."""
import ctypes

import numpy as np
import pytest

from rustpde_amd import RbcError, _lib
from rustpde_amd.funspace import _Dev
from rustpde_amd.solver import HHOLTZ, HHOLTZ_ADI, POISSON, Solver

pytestmark = pytest.mark.gpu

CD, CN, FO = 1, 2, 4


def _rhs(sv, batch, seed):
    rng = np.random.default_rng(seed)
    shp = (batch,) + sv.shape_in
    a = rng.standard_normal(shp)
    if sv.is_complex:
        a = a + 1j * rng.standard_normal(shp)
    return np.ascontiguousarray(a)


def _solve(sv, rhs, noalloc):
    batch = rhs.shape[0]
    out = np.full((batch,) + sv.shape_out, np.nan, dtype=rhs.dtype)
    din, dout = _Dev(rhs.nbytes), _Dev(out.nbytes)
    L = _lib.load()
    _lib.check(L.rbc_memcpy_h2d(din.p, rhs.ctypes.data, rhs.nbytes))
    _lib.check(L.rbc_memcpy_h2d(dout.p, out.ctypes.data, out.nbytes))
    if noalloc:
        sv.solve_device_noalloc(din.p, dout.p, batch)
    else:
        sv.solve_device(din.p, dout.p, batch)
    _lib.check(L.rbc_memcpy_d2h(out.ctypes.data, dout.p, out.nbytes))
    return out


def test_noalloc_matches_solve_shrinking_calls():
    cases = [  # (kind, base0, base1, c, batch), each call smaller than the one before
        (HHOLTZ, (CD, 128), (CD, 129), (1e-3, 1e-3), 6),
        (POISSON, (FO, 64), (CN, 65), (1.0, 1.0), 3),
        (HHOLTZ_ADI, (CN, 48), (CD, 40), (2e-3, 2e-3), 2),
        (HHOLTZ, (CD, 32), (CN, 24), (1e-2, 1e-2), 1),
    ]
    for i, (kind, b0, b1, c, batch) in enumerate(cases):
        print("case noalloc kind=%d %s x %s batch=%d" % (kind, b0, b1, batch), flush=True)
        sv = Solver(kind, b0, b1, c)
        rhs = _rhs(sv, batch, 10 + i)
        a = _solve(sv, rhs, noalloc=True)
        b = _solve(sv, rhs, noalloc=False)
        assert np.all(np.isfinite(a))
        assert np.array_equal(a, b), (kind, b0, b1)
        again = _solve(sv, rhs, noalloc=True)  # a second call finds its own contents in the scratch
        assert np.array_equal(again, b)


def test_noalloc_largest_batch_and_rejection():
    print("case noalloc Hholtz 512 x 512, batch 8 (fits), then 9 (rejected), then 2", flush=True)
    sv = Solver(HHOLTZ, (CD, 512), (CD, 512), (1e-3, 1e-3))
    rhs = _rhs(sv, 8, 3)
    assert np.array_equal(_solve(sv, rhs, noalloc=True), _solve(sv, rhs, noalloc=False))
    big = _Dev(8)
    with pytest.raises(RbcError, match="exceed"):
        sv.solve_device_noalloc(big.p, big.p, 9)  # rejected before any buffer is touched
    small = _rhs(sv, 2, 4)
    assert np.array_equal(_solve(sv, small, noalloc=True), _solve(sv, small, noalloc=False))

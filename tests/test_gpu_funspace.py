"""GPU funspace Space2 operations (fs_* C entry points) against the oracle.

Bases and operations of funspace/src/space2.rs:136-356; tolerance 1e-12 relative to the
max-norm of the oracle's result (FP64 operators applied as GEMMs or, along axis 1 for
n1 - 1 a power of two, as FFT-based DCT-I kernels, vs the oracle's lane recurrences).
"""
import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

CASES = [  # (kind0, n0, kind1, n1)
    (1, 17, 1, 16),   # cheb_dirichlet x cheb_dirichlet
    (2, 24, 1, 21),   # cheb_neumann x cheb_dirichlet
    (0, 20, 0, 19),   # chebyshev x chebyshev
    (4, 32, 1, 33),   # fourier_r2c x cheb_dirichlet
    (4, 16, 2, 17),   # fourier_r2c x cheb_neumann
    (4, 64, 0, 65),   # fourier_r2c x chebyshev
    (0, 18, 3, 18),   # chebyshev x cheb_dirichlet_bc
]

DCT_CASES = [  # long axis-1 rows: FFT-based DCT-I kernels up to 1025, dense operators beyond
    (0, 12, 0, 513),   # chebyshev x chebyshev
    (1, 10, 2, 1025),  # cheb_dirichlet x cheb_neumann
    (4, 8, 1, 2049),   # fourier_r2c x cheb_dirichlet (dense GEMM path)
]


def _rand(shape, cplx, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(shape)
    return a + 1j * rng.standard_normal(shape) if cplx else a


@pytest.mark.parametrize("case", CASES)
def test_space2_ops(oracle, case):
    from rustpde_amd import Space2
    k0, n0, k1, n1 = case
    print("case fs", case, flush=True)
    sp = Space2((k0, n0), (k1, n1))
    cplx = sp.is_complex
    v = _rand(sp.shape_physical, False, 1)
    vh = _rand(sp.shape_spectral, cplx, 2)
    po = _rand(sp.shape_ortho, cplx, 3)
    checks = [
        ("forward", sp.forward(v), oracle.space_apply(k0, n0, k1, n1, "forward", v)),
        ("backward", sp.backward(vh), oracle.space_apply(k0, n0, k1, n1, "backward", vh)),
        ("to_ortho", sp.to_ortho(vh), oracle.space_apply(k0, n0, k1, n1, "to_ortho", vh)),
        ("from_ortho", sp.from_ortho(po), oracle.space_apply(k0, n0, k1, n1, "from_ortho", po)),
    ]
    for d in ((1, 0), (0, 1), (2, 0), (0, 2), (1, 1)):
        scale = (2.0, 1.0)
        checks.append(("gradient%s" % (d,), sp.gradient(vh, d, scale),
                       oracle.space_apply(k0, n0, k1, n1, "gradient", vh, deriv=d, scale=scale)))
    for name, got, ref in checks:
        assert got.shape == ref.shape, (name, got.shape, ref.shape)
        assert rel_err(got, ref) < 1e-12, (name, rel_err(got, ref))


def test_space2_batched(oracle):
    from rustpde_amd import Space2
    sp = Space2((4, 32), (1, 17))
    v = _rand((3,) + sp.shape_physical, False, 5)
    got = sp.forward(v)
    for b in range(3):
        assert rel_err(got[b], oracle.space_apply(4, 32, 1, 17, "forward", v[b])) < 1e-12


@pytest.mark.parametrize("case", DCT_CASES)
def test_space2_dct_rows(oracle, case):
    """forward / backward with long Chebyshev rows against the oracle, and the round trip."""
    from rustpde_amd import Space2
    k0, n0, k1, n1 = case
    print("case fs dct", case, flush=True)
    sp = Space2((k0, n0), (k1, n1))
    cplx = sp.is_complex
    rng = np.random.default_rng(7)
    # a smooth field (decaying spectrum), as the transforms see in the time step
    vh = _rand(sp.shape_spectral, cplx, 8) / (1.0 + np.arange(sp.shape_spectral[1]))[None, :] ** 2
    if cplx:  # the real transform keeps no imaginary part at k = 0 and k = n0 / 2
        vh[0] = vh[0].real
        vh[-1] = vh[-1].real
    v = sp.backward(vh)
    assert rel_err(v, oracle.space_apply(k0, n0, k1, n1, "backward", vh)) < 1e-12
    f = sp.forward(v)
    assert rel_err(f, oracle.space_apply(k0, n0, k1, n1, "forward", v)) < 1e-12
    assert rel_err(f, vh) < 1e-12  # forward(backward(vh)) = vh
    w = rng.standard_normal(sp.shape_physical)
    assert rel_err(sp.forward(w), oracle.space_apply(k0, n0, k1, n1, "forward", w)) < 1e-12


"""The drop-in's own set-up (Navier2D::new / new_periodic, navier.rs:219-307, 384-467) against
the oracle's, with nothing shared: the boundary field of bc_rbc (navier.rs:314-332,
474-492), the tensor solvers' eigen-decompositions (utils.rs:66-98, fdma_tensor.rs:105-153,
the singular shift of poisson.rs:80-83) and steps run from them.

Tolerances.  The periodic set-up has no eigen-decomposition (Fourier x is diagonal) and a
boundary field that is constant in x, so the steps agree to 1e-12 as with a shared set-up.
The confined set-up carries two quantities that no FP64 computation fixes beyond its
rounding: the eigenvalues of inv(C) A (determined to ~1e-12 relative at n = 62 by any FP64
eigensolver: LAPACK on X and on X^T differ by that much) and the x-rounding noise of the
boundary field, which d2/dx2 amplifies by O(n^4).  Two faithful set-ups therefore differ
by ~1e-12 after one step, and the 10-step comparison is held to 1e-10.
"""
import numpy as np
import pytest

from helpers import compare, init_pair, rel_err

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_c1_confined_own_setup(oracle):
    from rustpde_amd import Navier2D
    nx = ny = 64
    _flush("case C1 confined 64x64 own set-up, 10 steps")
    g = Navier2D.new(nx, ny, 1e5, 1.0, 0.02, 1.0, True)
    o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, 0.02, 1.0, True)
    init_pair(g, o, nx, ny, False, 0.05, 1, "uniform", velocity=(0.2, 1.0, 1.0), share_setup=False)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    _flush("C1 own set-up errors after 10 steps: %s" % e)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-10, (k, e)
    assert abs(g.eval_nu() - o.eval_nu()) < 1e-10 * abs(o.eval_nu())


def test_c3_periodic_own_setup(oracle):
    from rustpde_amd import Navier2D
    nx, ny = 256, 257
    _flush("case C3 periodic 256x257 own set-up, 10 steps")
    g = Navier2D.new_periodic(nx, ny, 1e7, 1.0, 0.002, 2.0)
    o = oracle.Navier2D.new_periodic(nx, ny, 1e7, 1.0, 0.002, 2.0)
    init_pair(g, o, nx, ny, True, 1e-3, 7, "gaussian", velocity=(0.05, 2.0, 1.0), share_setup=False)
    g.update(10)
    o.update(10)
    e = compare(g, o)
    _flush("C3 own set-up errors after 10 steps: %s" % e)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-12, (k, e)


@pytest.mark.parametrize("periodic", [False, True])
def test_fieldbc_matches(oracle, periodic):
    """bc_rbc: the boundary field's spectral and physical arrays (T_bc = -y/2)."""
    from rustpde_amd import Navier2D
    nx, ny = (64, 65) if periodic else (64, 64)
    _flush("case fieldbc periodic=%s" % periodic)
    if periodic:
        g = Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.01, 2.0)
        o = oracle.Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.01, 2.0)
    else:
        g = Navier2D.new(nx, ny, 1e6, 1.0, 0.01, 1.0, True)
        o = oracle.Navier2D.new(nx, ny, 1e6, 1.0, 0.01, 1.0, True)
    a, b = g.fieldbc.vhat, o.get("fieldbc")
    assert a.shape == b.shape
    assert np.abs(a - b).max() < 1e-14 * max(1.0, np.abs(b).max()), np.abs(a - b).max()
    y = -np.cos(np.pi * np.arange(ny) / (ny - 1))
    v = g.fieldbc.v
    assert np.abs(v - (-0.5 * y)[None, :]).max() < 1e-14


@pytest.mark.parametrize("adiabatic", [True, False])
def test_confined_poisson_tensor(oracle, adiabatic):
    """The pressure Poisson solver of the confined form (cheb_neumann x cheb_neumann): the
    x-operator has a zero column, so its eigenvector is exactly e_0 and its eigenvalue 0,
    shifted to exactly -1e-10 (poisson.rs:80-83); eigenvalues as the oracle's (sorted
    descending, utils.rs:80-96) and the decomposition reconstructs the same operator."""
    from rustpde_amd import Navier2D
    nx = ny = 64
    _flush("case confined Poisson tensor adiabatic=%s" % adiabatic)
    g = Navier2D.new(nx, ny, 1e5, 1.0, 0.02, 1.0, adiabatic)
    o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, 0.02, 1.0, adiabatic)
    lg, fg, bg = g.tensor_setup(3)
    lo, fo, bo = o.tensor_setup(3)
    m = nx - 2
    assert lg.shape == (m,) and fg.shape == (m, m) and bg.shape == (m, m)
    assert lg[0] == -1e-10 and lo[0] == -1e-10
    e0 = np.zeros(m)
    e0[0] = 1.0
    assert np.array_equal(np.abs(bg[:, 0]), e0), bg[:, 0]
    assert np.all(np.diff(lg) <= 0)
    assert rel_err(lg, lo) < 1e-11
    # Q diag(lam) Q^-1 C^-1 from both decompositions (independent of eigenvector scaling)
    rg = (bg * lg[None, :]) @ fg
    ro = (bo * lo[None, :]) @ fo
    assert rel_err(rg, ro) < 1e-11


def test_periodic_tensor_diagonal(oracle):
    """Periodic solvers: axis 0 is Fourier, so lam = diag(A_x) = -k^2 / sc^2 (the Hholtz
    solvers' c already folded in) and there are no fwd/bwd matrices."""
    from rustpde_amd import Navier2D
    nx, ny = 64, 65
    _flush("case periodic tensors")
    g = Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.01, 2.0)
    o = oracle.Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.01, 2.0)
    for s in range(4):
        tg, to = g.tensor_setup(s), o.tensor_setup(s)
        assert tg is not None and to is not None
        assert tg[1] is None and to[1] is None
        assert rel_err(tg[0], to[0]) < 1e-15, s

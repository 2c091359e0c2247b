"""The oracle's Statistics (src/navier/statistics.rs:130-153, :220-247) pinned without the
reference: the conduction state's closed form, linearity of the running average, the time
bookkeeping, and the Nusselt field recomputed from the statistics' own spectral arrays with
numpy's Chebyshev and FFT routines; plus the checkpoint container and broadcast helpers.
"""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as C

from helpers import rel_err
from rustpde_amd import seeded
from rustpde_amd.checkpoint import broadcast_2d, flow_filename, get_dx, coords, read_container, write_container

CASES = [(False, 16, 16), (True, 16, 17)]


def _new(oracle, periodic, nx, ny, ra=1e4):
    if periodic:
        return oracle.Navier2D.new_periodic(nx, ny, ra, 1.0, 0.01, 2.0)
    return oracle.Navier2D.new(nx, ny, ra, 1.0, 0.01, 1.0, True)


def _seed(o, nx, ny, periodic, seed):
    o.set_velocity(0.2, 1.0, 1.0)
    o.set_phys_forward("temp", seeded.interior_perturbation(nx, ny, 0.05, seed, kind="uniform", periodic=periodic))


@pytest.mark.parametrize("periodic,nx,ny", CASES)
def test_conduction_state_closed_form(oracle, periodic, nx, ny):
    """u = 0, T = T_bc = -y/2: dT/dy = -1/2, so the Nusselt field is 1 everywhere, t_avg is
    T_bc and the velocity arrays are zero."""
    o = _new(oracle, periodic, nx, ny)
    o.stats_enable(1.0, 10.0)
    assert o.stats_update()
    assert np.abs(o.stats_get(3, False) - 1.0).max() < 1e-12
    assert np.abs(o.stats_get(0, False) - o.get("fieldbc", False)).max() < 1e-14
    assert np.abs(o.stats_get(1)).max() == 0.0 and np.abs(o.stats_get(2)).max() == 0.0


@pytest.mark.parametrize("periodic,nx,ny", CASES)
def test_running_average_and_time_bookkeeping(oracle, periodic, nx, ny):
    """Three updates with temperatures q1, q2, q3 give the t_avg of one update with their
    mean (to_ortho is linear); avg_time = t_last - t_enable, tot_time = t_last, num_save = 3;
    an update at an earlier time changes nothing."""
    o = _new(oracle, periodic, nx, ny)
    o.set_time(0.5)
    o.stats_enable(1.0, 10.0)
    qs = [seeded.interior_perturbation(nx, ny, 0.05, 10 + k, kind="uniform", periodic=periodic) for k in range(3)]
    for k, q in enumerate(qs):
        o.set_phys_forward("temp", q)
        o.set_time(1.0 + k)
        assert o.stats_update()
    avg = o.stats_get(0)
    assert o.stats_info() == (2.5, 3.0, 3)
    o.set_time(2.5)
    assert not o.stats_update()
    assert o.stats_info() == (2.5, 3.0, 3) and np.array_equal(o.stats_get(0), avg)
    m = _new(oracle, periodic, nx, ny)
    m.set_phys_forward("temp", (qs[0] + qs[1] + qs[2]) / 3.0)
    m.stats_enable(1.0, 10.0)
    m.stats_update()
    assert rel_err(avg, m.stats_get(0)) < 1e-14


def _eval2(c, periodic, nx, ny):
    """Physical values of field-space coefficients c (Chebyshev T_l(y_j) along y at
    y_j = -cos(pi j / (ny - 1)); along x Chebyshev at the same nodes or the inverse real FFT
    with the 1/n normalisation), and of their y-derivative."""
    y = -np.cos(math.pi * np.arange(ny) / (ny - 1))
    dc = C.chebder(c, axis=1)
    vy = np.stack([C.chebval(y, row) for row in c])      # [n0, ny]
    dvy = np.stack([C.chebval(y, row) for row in dc])
    if periodic:
        return np.fft.irfft(vy, n=nx, axis=0), np.fft.irfft(dvy, n=nx, axis=0)
    x = -np.cos(math.pi * np.arange(nx) / (nx - 1))
    return C.chebval(x, vy).T, C.chebval(x, dvy).T  # chebval puts the evaluation axis last


@pytest.mark.parametrize("periodic,nx,ny", CASES)
def test_nusselt_field_numpy(oracle, periodic, nx, ny):
    """The Nusselt field (dT/dy (-1) + uy T / ka) 2 from t_avg and uy_avg after one update
    (t_avg = that), evaluated with numpy, matches the oracle's; its domain average is
    eval_nuvol."""
    o = _new(oracle, periodic, nx, ny)
    _seed(o, nx, ny, periodic, 5)
    o.update(2)
    o.stats_enable(1.0, 10.0)
    o.stats_update()
    t, dt = _eval2(o.stats_get(0), periodic, nx, ny)
    uy, _ = _eval2(o.stats_get(2), periodic, nx, ny)
    ka = math.sqrt(1.0 / ((1e4 / 8.0) * 1.0))
    nus = (dt * -1.0 + uy * t / ka) * 2.0
    assert rel_err(o.stats_get(3, False), nus) < 1e-11
    assert rel_err(uy, o.backward("uy")) < 1e-11
    x = coords(nx, periodic, 2.0 if periodic else 1.0)
    y = coords(ny, False, 1.0)
    wx, wy = get_dx(x, periodic), get_dx(y, False)
    # average.rs:25-57: weights dx / |x_last - x_first| along each axis (for Fourier x the
    # length is (n - 1) / n of the period, a property of the reference's average)
    lx, ly = abs(x[-1] - x[0]), abs(y[-1] - y[0])
    avg = float((wx[:, None] / lx * wy[None, :] / ly * o.stats_get(3, False)).sum())
    assert abs(avg - o.eval_nuvol()) < 1e-10 * max(1.0, abs(avg))


@pytest.mark.parametrize("suffix", [".npz", ".h5"])
def test_container_roundtrip(tmp_path, suffix):
    from rustpde_amd import h5
    if suffix == ".h5" and not h5.available():
        pytest.skip("libhdf5 not found on this host")
    f = str(tmp_path / ("snap" + suffix))
    rng = np.random.default_rng(3)
    a = rng.standard_normal((5, 7))
    z = rng.standard_normal((3, 4)) + 1j * rng.standard_normal((3, 4))
    write_container(f, {"temp/v": a, "temp/vhat": z, "time": 1.25, "num_save": 7})
    d = read_container(f)
    assert np.array_equal(d["temp/v"], a) and np.array_equal(d["temp/vhat"], z)
    assert d["time"].shape == (1,) and d["time"][0] == 1.25 and d["num_save"][0] == 7
    raw = h5.read(f) if suffix == ".h5" else dict(np.load(f))
    assert set(raw) == {"temp/v", "temp/vhat_re", "temp/vhat_im", "time", "num_save"}


def test_broadcast_2d_and_names():
    """field/read.rs:112-121: the overlap is copied, the rest of the target is kept."""
    old = np.arange(12.0).reshape(3, 4)
    new = np.full((4, 3), -1.0)
    b = broadcast_2d(old, new)
    assert np.array_equal(b[:3, :3], old[:3, :3]) and np.all(b[3] == -1.0)
    assert np.array_equal(broadcast_2d(np.ones((5, 5)), np.zeros((2, 3))), np.ones((2, 3)))
    from rustpde_amd.checkpoint import _SUFFIX
    assert flow_filename(1.5) == "flow00001.50" + _SUFFIX and flow_filename(123.456) == "flow00123.46" + _SUFFIX
    x = coords(9, False, 1.0)
    assert abs(get_dx(x, False).sum() - 2.0) < 1e-15 and np.allclose(get_dx(coords(8, True, 2.0), True), 2 * np.pi * 2 / 8)

"""Statistics on the device and checkpoint/restart, through the C ABI.

Statistics (src/navier/statistics.rs:130-153, :220-247): the four arrays and the time
bookkeeping against the oracle after several updates between steps, 1e-12 relative (FP64).
Checkpoint (Navier2D::write / read, navier.rs:961-1008): a run restarted from a snapshot
continues as the uninterrupted run (to 1e-13); a snapshot read at another resolution is
broadcast (field/read.rs:112-121); callback writes snapshots and statistics at their cadence.
"""
import os

import numpy as np
import pytest

from helpers import init_pair, rel_err

pytestmark = pytest.mark.gpu


def _pair(oracle, periodic, nx=None, ny=None):
    from rustpde_amd import Navier2D
    if periodic:
        nx, ny = nx or 128, ny or 129
        g = Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.004, 2.0)
        o = oracle.Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.004, 2.0)
        init_pair(g, o, nx, ny, True, 1e-3, 21, "gaussian", velocity=(0.05, 2.0, 1.0))
    else:
        nx, ny = nx or 48, ny or 65
        g = Navier2D.new(nx, ny, 1e5, 1.0, 0.01, 1.0, True)
        o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, 0.01, 1.0, True)
        init_pair(g, o, nx, ny, False, 0.05, 21, "uniform", velocity=(0.2, 1.0, 1.0))
    return g, o


@pytest.mark.parametrize("periodic", [False, True])
def test_statistics_vs_oracle(oracle, periodic):
    print("case statistics periodic=%s" % periodic, flush=True)
    g, o = _pair(oracle, periodic)
    st = g.init_statistics(0.01, 1.0)
    o.stats_enable(0.01, 1.0)
    for _ in range(4):
        g.update(2)
        o.update(2)
        assert st.update() and o.stats_update()
    for k, f in enumerate((st.t_avg, st.ux_avg, st.uy_avg, st.nusselt)):
        e_spec = rel_err(f.vhat, o.stats_get(k))
        e_phys = rel_err(f.v, o.stats_get(k, False))
        print("stat", k, e_spec, e_phys, flush=True)
        assert e_spec < 1e-12 and e_phys < 1e-12, (k, e_spec, e_phys)
    a, t, n = o.stats_info()
    assert abs(st.avg_time - a) < 1e-12 and abs(st.tot_time - t) < 1e-12 and st.num_save == n == 4
    g.time = g.get_time() - 1.0  # an earlier time: no update (statistics.rs:134-140)
    before = st.t_avg.vhat
    assert not st.update()
    assert np.array_equal(st.t_avg.vhat, before) and st.num_save == 4
    g.close()


@pytest.mark.parametrize("periodic", [False, True])
def test_restart_continues(oracle, periodic, tmp_path):
    print("case restart periodic=%s" % periodic, flush=True)
    from rustpde_amd import Navier2D, checkpoint
    a, o = _pair(oracle, periodic)
    b, _ = _pair(oracle, periodic)
    a.update(5)
    o.update(10)  # the oracle runs the 10 steps without a restart
    f = str(tmp_path / ("snap" + checkpoint._SUFFIX))  # an HDF5 file where libhdf5 is found
    a.temp.backward()
    a.ux.backward()
    temp_v = a.temp.v + a.fieldbc.v
    ux_v = a.ux.v
    a.write(f)
    a.update(5)
    b.read(f)
    assert b.get_time() == pytest.approx(5 * a.get_dt(), abs=1e-15)
    b.update(5)
    for fa, fb, name in ((a.temp, b.temp, "temp"), (a.ux, b.ux, "ux"), (a.uy, b.uy, "uy"),
                         (a.pres[0], b.pres[0], "pres")):
        # the first step after a read forms d/dy pres with its own kernel (periodic), so
        # the continuation agrees to rounding rather than bit for bit
        assert rel_err(fb.vhat, fa.vhat) < 1e-13
        assert rel_err(fb.vhat, o.get(name)) < 1e-12, name  # the restarted run against the oracle
    assert a.get_time() == b.get_time()
    # the snapshot's datasets (navier.rs:983-1013, field/write.rs:82-115): temp/v carries the
    # boundary field, v and vhat per group, coordinates and cell widths scaled by the aspect
    d = checkpoint.read_container(f)
    assert d["temp/v"].shape == (a.nx, a.ny) and d["time"].shape == (1,)
    assert np.array_equal(d["temp/v"], temp_v) and np.array_equal(d["ux/v"], ux_v)
    sx = a.scale[0]
    if periodic:
        assert np.allclose(d["x"], 2 * np.pi * np.arange(a.nx) / a.nx * sx, rtol=0, atol=1e-14)
        assert np.allclose(d["dx"], 2 * np.pi / a.nx * sx, rtol=0, atol=1e-14)
    else:
        assert np.allclose(d["x"], -np.cos(np.pi * np.arange(a.nx) / (a.nx - 1)) * sx, rtol=0, atol=1e-14)
        assert abs(d["dx"].sum() - 2.0 * sx) < 1e-13  # the cells partition [-sx, sx]
        assert d["dx"][0] == pytest.approx((d["x"][1] - d["x"][0]) / 2, abs=1e-15)
    assert np.allclose(d["y"], -np.cos(np.pi * np.arange(a.ny) / (a.ny - 1)), rtol=0, atol=1e-14)
    assert abs(d["dy"].sum() - 2.0) < 1e-13
    for k in ("time", "ra", "pr", "nu", "kappa"):
        assert d[k].shape == (1,)
    assert d["nu"][0] == a.nu and d["kappa"][0] == a.ka and "solid/mask" not in d
    # another resolution: the overlap of the spectral arrays is taken, the rest stays zero
    if periodic:
        c = Navier2D.new_periodic(160, 129, 1e6, 1.0, 0.004, 2.0)
    else:
        c = Navier2D.new(64, 65, 1e5, 1.0, 0.01, 1.0, True)
    c.read(f)
    old = d["temp/vhat"]
    got = c.temp.vhat
    n0, n1 = old.shape
    assert np.array_equal(got[:n0, :n1], old) and not np.any(got[n0:])
    c.update(2)
    assert np.isfinite(c.temp.vhat).all()
    for h in (a, b, c):
        h.close()


def test_callback_writes_at_cadence(oracle, tmp_path):
    print("case callback files", flush=True)
    from rustpde_amd import checkpoint, integrate
    g, _ = _pair(oracle, False)
    g.data_dir = str(tmp_path)
    g.write_intervall = 0.05
    g.init_statistics(0.02, 0.1)
    integrate(g, 0.1, 0.01)
    names = sorted(os.listdir(tmp_path))
    sfx = checkpoint._SUFFIX
    assert names == ["flow00000.05" + sfx, "flow00000.10" + sfx, "statistics" + sfx], names
    assert g.statistics.num_save == 5
    from rustpde_amd import Navier2D
    h = Navier2D.new(48, 65, 1e5, 1.0, 0.01, 1.0, True)
    st = h.init_statistics(0.02, 0.1)
    st.read(str(tmp_path / ("statistics" + sfx)))
    assert st.num_save == 5 and np.array_equal(st.t_avg.vhat, g.statistics.t_avg.vhat)
    g.close()
    h.close()

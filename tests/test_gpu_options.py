"""Equivalent implementations of the GPU step agree with the oracle and with each other.

Options (rbc_set_option): HIP-graph replay, FP64 MFMA GEMMs, FFT kernels for the
Fourier axis, banded kernels for stencils/preconditioners, FFT-based DCT-I kernels for
the Chebyshev y axis (active when ny - 1 is a power of two: the periodic 128x129 case
and the confined 48x65 case below), the batched and fused periodic step, and the
one-launch y-stage of the convective terms (conv3; its off form is the three launches it
replaces).  Tolerances as in
test_gpu_navier.py (1e-12 for T/ux/uy/p after the steps taken).
"""
import numpy as np
import pytest

from helpers import compare, init_pair

pytestmark = pytest.mark.gpu

OPTS = {"graph": 0, "mfma": 1, "fft": 2, "bands": 3, "dct": 4, "batch": 5, "fuse": 6, "conv3": 8}


def _pair(oracle, periodic):
    from rustpde_amd import Navier2D
    if periodic:
        nx, ny = 128, 129
        g = Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.004, 2.0)
        o = oracle.Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.004, 2.0)
        init_pair(g, o, nx, ny, True, 1e-3, 11, "gaussian", velocity=(0.05, 2.0, 1.0))
    else:
        nx, ny = 48, 65
        g = Navier2D.new(nx, ny, 1e5, 1.0, 0.01, 1.0, True)
        o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, 0.01, 1.0, True)
        init_pair(g, o, nx, ny, False, 0.05, 11, "uniform", velocity=(0.2, 1.0, 1.0))
    return g, o


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("off", [None, "graph", "mfma", "fft", "bands", "dct", "batch", "fuse", "conv3"])
def test_option_parity(oracle, periodic, off):
    print("case options periodic=%s off=%s" % (periodic, off), flush=True)
    g, o = _pair(oracle, periodic)
    if off is not None:
        g.set_option(OPTS[off], 0)
    g.update(6)
    o.update(6)
    e = compare(g, o)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-12, (k, e)
    assert e["pseu"] < 1e-7, e


@pytest.mark.parametrize("periodic", [False, True])
def test_graph_replay_matches_eager(oracle, periodic):
    """Graph replays (one step or eight per launch) interleaved with eager work reproduce the
    eager steps bit for bit, and the time advances as in eager steps."""
    print("case graph replay periodic=%s" % periodic, flush=True)
    ge, _ = _pair(oracle, periodic)
    gg, _ = _pair(oracle, periodic)
    ge.set_option(OPTS["graph"], 0)
    for rep, n in enumerate((3, 3, 3, 19)):
        ge.update(n)
        gg.update(n)  # 1 eager + 2 replays the first time, 3 after; 19: two 8-step replays and 3
        _ = gg.temp.v  # eager device work between replays
        for a, b in ((ge.temp, gg.temp), (ge.ux, gg.ux), (ge.uy, gg.uy), (ge.pres[0], gg.pres[0])):
            assert np.array_equal(a.vhat, b.vhat), rep
    assert ge.get_time() == gg.get_time()


@pytest.mark.parametrize("tile", [0, 1, 2, 3])
def test_gemm_tile_shapes_bitwise(oracle, monkeypatch, tile):
    """Every MFMA GEMM tile shape (64x64, 32x32, 64x32, 32x64; normally picked per launch
    shape by the autotuner) gives bit-identical steps: the accumulation order over K does
    not depend on the tile.  Confined 48x65 (dense x transforms and ADI operators) and an
    ensemble group of three instances (batched GEMMs)."""
    from rustpde_amd import Ensemble
    print("case gemm tile %d" % tile, flush=True)
    results = []
    for forced in (None, tile):
        if forced is None:
            monkeypatch.delenv("RUSTPDE_AMD_GEMM_TILE", raising=False)
        else:
            monkeypatch.setenv("RUSTPDE_AMD_GEMM_TILE", str(forced))
        g, o = _pair(oracle, False)
        g.update(5)
        ens = Ensemble([(32, 33, 1e4 * (1 + i), 1.0, 0.01, 1.0, True) for i in range(3)])
        for i in range(3):
            ens[i].set_velocity(0.2, 1.0, 1.0)
        ens.update(3)
        results.append([g.temp.vhat.copy(), g.pres[0].vhat.copy()] + [ens[i].uy.vhat.copy() for i in range(3)])
        ens.close()
        g.close()
    monkeypatch.delenv("RUSTPDE_AMD_GEMM_TILE", raising=False)
    for a, b in zip(results[0], results[1]):
        assert np.array_equal(a, b)


def test_confined_field_pair_bitwise(oracle):
    """The confined step's ux/uy pair (their transforms, products, implicit solves and
    projection as launches of twice the batch; option batch = 1) reproduces the step that
    treats the two fields one after the other (batch = 0) bit for bit: every batch entry
    runs the same arithmetic as its own launch would."""
    print("case confined field pair bitwise", flush=True)
    out = []
    for batch in (1, 0):
        g, _ = _pair(oracle, False)
        g.set_option(OPTS["batch"], batch)
        g.update(6)
        out.append([f.vhat.copy() for f in (g.temp, g.ux, g.uy, g.pres[0], g.pres[1])])
        g.close()
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_pressure_written_between_steps(oracle):
    """The periodic step takes d/dy pres from the previous step's pressure stage; a pressure
    written by the host between steps (here scaled and shifted) must be differentiated
    afresh, with and without graph replay, and the steps must still match the oracle."""
    print("case pressure written between steps", flush=True)
    for graph in (1, 0):
        g, o = _pair(oracle, True)
        g.set_option(OPTS["graph"], graph)
        g.update(4)
        o.update(4)
        p = o.get("pres") * 1.5 + 1e-3
        g.pres[0].vhat = p
        o.set("pres", p)
        g.update(4)
        o.update(4)
        e = compare(g, o)
        for k in ("temp", "ux", "uy", "pres"):
            assert e[k] < 1e-12, (graph, k, e)


@pytest.mark.parametrize("periodic", [False, True])
def test_single_step_calls_replay(oracle, periodic):
    """One step per call, as integrate() makes them: from the second call on the captured
    step is replayed.  The fields equal eager steps bit for bit, also with fields written by
    the host between calls (temp scaled, pres shifted), and match the oracle."""
    print("case single-step calls periodic=%s" % periodic, flush=True)
    ge, _ = _pair(oracle, periodic)
    gg, o = _pair(oracle, periodic)
    ge.set_option(OPTS["graph"], 0)
    for rep in range(2):
        for _ in range(5):
            ge.update(1)
            gg.update(1)
            o.update(1)
        for a, b in ((ge.temp, gg.temp), (ge.ux, gg.ux), (ge.uy, gg.uy), (ge.pres[0], gg.pres[0])):
            assert np.array_equal(a.vhat, b.vhat), rep
        t = o.get("temp") * 0.9
        p = o.get("pres") + 1e-3
        for g in (ge, gg):
            g.temp.vhat = t
            g.pres[0].vhat = p
        o.set("temp", t)
        o.set("pres", p)
    e = compare(gg, o)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-12, (k, e)
    assert abs(ge.get_time() - gg.get_time()) < 1e-15

"""Solid obstacles on the device (navier.rs:551-560, :579-584, :603-608) through the C ABI:
the penalised step against the oracle with the reference's mask builders, 1e-12 relative
(FP64) after 6 steps, eager and graph-replayed, in the batched periodic step and field by
field; removing the mask returns to the step without it (still matching the oracle)."""
import numpy as np
import pytest

from helpers import compare, init_pair
from rustpde_amd.checkpoint import coords
from rustpde_amd.solid_masks import solid_cylinder_inner, solid_roughness_sinusoid

pytestmark = pytest.mark.gpu


def _pair(oracle, periodic):
    from rustpde_amd import Navier2D
    if periodic:
        nx, ny = 128, 129
        g = Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.002, 2.0)
        o = oracle.Navier2D.new_periodic(nx, ny, 1e6, 1.0, 0.002, 2.0)
        init_pair(g, o, nx, ny, True, 1e-3, 31, "gaussian", velocity=(0.05, 2.0, 1.0))
        x = coords(nx, True, 2.0)
        masks = solid_roughness_sinusoid(x, coords(ny, False, 1.0), 0.1, 3.0)
    else:
        nx, ny = 48, 65
        g = Navier2D.new(nx, ny, 1e5, 1.0, 0.005, 1.0, True)
        o = oracle.Navier2D.new(nx, ny, 1e5, 1.0, 0.005, 1.0, True)
        init_pair(g, o, nx, ny, False, 0.05, 31, "uniform", velocity=(0.2, 1.0, 1.0))
        masks = solid_cylinder_inner(coords(nx, False, 1.0), coords(ny, False, 1.0), 0.2, 0.0, 0.3)
    return g, o, masks


@pytest.mark.parametrize("graph,batch", [(1, 1), (0, 1), (1, 0)])
@pytest.mark.parametrize("periodic", [False, True])
def test_solid_vs_oracle(oracle, periodic, graph, batch, tmp_path):
    from rustpde_amd import RbcError
    print("case solid periodic=%s graph=%d batch=%d" % (periodic, graph, batch), flush=True)
    g, o, masks = _pair(oracle, periodic)
    g.set_option(0, graph)
    g.set_option(g.OPT_BATCH, batch)  # periodic: the batched step takes the terms, or field by field
    g.solid = masks
    o.set_solid(masks)
    with pytest.raises(RbcError):  # the penalised step has no mixed-precision form
        g.set_option(g.OPT_MIXED, 1)
    g.update(6)
    o.update(6)
    e = compare(g, o)
    print("solid errors", e, flush=True)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-12, (k, e)
    assert e["pseu"] < 1e-7, e
    # a snapshot carries the mask (navier.rs:995-998)
    from rustpde_amd import checkpoint
    f = str(tmp_path / ("solid" + checkpoint._SUFFIX))
    g.write(f)
    assert np.array_equal(checkpoint.read_container(f)["solid/mask"], masks[0])
    # the obstacle changes the flow (the same step without it differs)
    h, _, _ = _pair(oracle, periodic)
    h.update(6)
    assert np.abs(h.ux.vhat - g.ux.vhat).max() > 1e-6 * np.abs(g.ux.vhat).max()
    # removed: the step without the obstacle, still the oracle's
    g.solid = None
    o.set_solid(None)
    g.update(3)
    o.update(3)
    e = compare(g, o)
    for k in ("temp", "ux", "uy", "pres"):
        assert e[k] < 1e-12, (k, e)
    g.close()
    h.close()

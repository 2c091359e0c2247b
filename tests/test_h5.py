"""HDF5 snapshots (rustpde_amd.h5, through the HDF5 C library) against the layout the
reference's hdf5-interface writes (hdf5-interface/src/lib.rs:111-172, 235-268):

* our files, inspected by the library's own h5dump: every dataset contiguous, no filter,
  f64 as IEEE F64LE, the usize scalar num_save as U64LE, complex arrays as <name>_re and
  <name>_im, groups along the dataset paths;
* a file written by another writer (the library's h5import tool, the same dataset paths
  and types as Navier2D::write of a periodic handle: temp/v, temp/vhat_re, temp/vhat_im,
  time) reads back through read_container with the values and the complex join.

Skipped on hosts without libhdf5 (the .npz container is then the only one) or without the
tools.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from rustpde_amd import h5
from rustpde_amd.checkpoint import read_container, write_container

pytestmark = pytest.mark.skipif(not h5.available(), reason="libhdf5 not found on this host")


def _tool(name):
    """an HDF5 command-line tool: beside the library found (lib -> bin), else on PATH"""
    lib = getattr(h5.load(), "_name", "") or ""
    p = os.path.join(os.path.dirname(os.path.dirname(lib)), "bin", name) if os.path.isabs(lib) else ""
    return p if p and os.path.exists(p) else shutil.which(name)


def test_layout_seen_by_h5dump(tmp_path):
    dump = _tool("h5dump")
    if not dump:
        pytest.skip("h5dump not found")
    f = str(tmp_path / "flow.h5")
    a = np.arange(6.0).reshape(2, 3)
    write_container(f, {"temp/v": a, "temp/vhat": a + 1j * a, "x": np.linspace(-1, 1, 2), "num_save": 3,
                        "time": 0.5})
    out = subprocess.run([dump, "-p", "-H", f], capture_output=True, text=True, timeout=60).stdout
    for ds in ("v", "vhat_re", "vhat_im", "x", "time", "num_save"):
        assert 'DATASET "%s"' % ds in out, ds
    assert 'GROUP "temp"' in out
    assert out.count("CONTIGUOUS") == 6 and out.count("CHUNKED") == 0
    assert out.count("H5T_IEEE_F64LE") == 5 and out.count("H5T_STD_U64LE") == 1
    vals = subprocess.run([dump, "-d", "/temp/v", "-y", "-w", "0", f], capture_output=True, text=True,
                          timeout=60).stdout
    assert "0, 1, 2, 3, 4, 5" in " ".join(vals.split())


def _h5import(tool, tmp_path, path, arr, out):
    txt = tmp_path / (path.replace("/", "_") + ".txt")
    cfg = tmp_path / (path.replace("/", "_") + ".cfg")
    txt.write_text(" ".join(repr(float(v)) for v in arr.ravel()) + "\n")
    cfg.write_text("PATH %s\nINPUT-CLASS TEXTFP\nINPUT-SIZE 64\nRANK %d\nDIMENSION-SIZES %s\n"
                   "OUTPUT-CLASS FP\nOUTPUT-SIZE 64\nOUTPUT-ARCHITECTURE IEEE\nOUTPUT-BYTE-ORDER LE\n"
                   % (path, arr.ndim, " ".join(str(n) for n in arr.shape)))
    return [str(txt), "-c", str(cfg)]


def test_reads_a_file_of_another_writer(tmp_path):
    tool = _tool("h5import")
    if not tool:
        pytest.skip("h5import not found")
    rng = np.random.default_rng(1)
    v = rng.standard_normal((4, 5))
    re, im = rng.standard_normal((3, 5)), rng.standard_normal((3, 5))
    t = np.array([2.5])
    f = str(tmp_path / "ref.h5")
    args = [tool]
    for path, arr in (("temp/v", v), ("temp/vhat_re", re), ("temp/vhat_im", im), ("time", t)):
        args += _h5import(tool, tmp_path, path, arr, f)
    args += ["-o", f]
    r = subprocess.run(args, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    d = read_container(f)
    assert np.array_equal(d["temp/v"], v)
    assert np.array_equal(d["temp/vhat"], re + 1j * im)
    assert d["time"].shape == (1,) and d["time"][0] == 2.5

"""HDF5 files through the HDF5 C library (libhdf5, loaded with ctypes), for the snapshots
and statistics files of Navier2D::write / read and Statistics::write / read.

The reference writes them with the hdf5 crate over the same library
(hdf5-interface/src/lib.rs:111-148): every dataset contiguous (no chunking, no filter),
f64 in the native layout (IEEE little endian here), usize scalars as u64, complex arrays as
two real datasets ``<name>_re`` and ``<name>_im`` (:150-172), groups created on the way
(``temp/v``).  ``write`` produces those datasets with the same calls (H5Fcreate,
H5Dcreate2 with intermediate groups, H5Dwrite); ``read`` walks every group of a file and
returns all datasets, so a file the reference wrote reads back here and the other way round.

The library is looked up in RUSTPDE_HDF5_LIB, then by ctypes.util.find_library("hdf5"), then
in /opt/conda/lib; ``available()`` tells whether one was found (callers fall back to the
NumPy container otherwise).  HDF5 1.10 or later (hid_t is 64-bit).
"""
import ctypes
import ctypes.util
import os

import numpy as np

_lib = None
_err = None
_hid = ctypes.c_int64
_herr = ctypes.c_int
_hsize = ctypes.c_uint64

H5F_ACC_RDONLY, H5F_ACC_TRUNC = 0x0000, 0x0002
H5P_DEFAULT = 0
H5S_ALL = 0
H5I_GROUP, H5I_DATASET = 2, 5
H5T_INTEGER, H5T_FLOAT = 0, 1
H5_INDEX_NAME, H5_ITER_INC = 0, 0

_ITER = ctypes.CFUNCTYPE(_herr, _hid, ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p)


def _candidates():
    env = os.environ.get("RUSTPDE_HDF5_LIB")
    if env:
        yield env
    found = ctypes.util.find_library("hdf5")
    if found:
        yield found
    for d in ("/opt/conda/lib", "/usr/lib/x86_64-linux-gnu", "/usr/local/lib"):
        yield os.path.join(d, "libhdf5.so")


def load():
    """The HDF5 library with the signatures used here; raises OSError if none is found."""
    global _lib, _err
    if _lib is not None:
        return _lib
    if _err is not None:
        raise OSError(_err)
    tried = []
    for path in _candidates():
        try:
            lib = ctypes.CDLL(path)
        except OSError as e:
            tried.append("%s (%s)" % (path, e))
            continue
        sig = {
            "H5open": (_herr, []),
            "H5Fcreate": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid, _hid]),
            "H5Fopen": (_hid, [ctypes.c_char_p, ctypes.c_uint, _hid]),
            "H5Fclose": (_herr, [_hid]),
            "H5Pcreate": (_hid, [_hid]),
            "H5Pset_create_intermediate_group": (_herr, [_hid, ctypes.c_uint]),
            "H5Pclose": (_herr, [_hid]),
            "H5Screate_simple": (_hid, [ctypes.c_int, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            "H5Sget_simple_extent_ndims": (ctypes.c_int, [_hid]),
            "H5Sget_simple_extent_dims": (ctypes.c_int, [_hid, ctypes.POINTER(_hsize), ctypes.POINTER(_hsize)]),
            "H5Sclose": (_herr, [_hid]),
            "H5Dcreate2": (_hid, [_hid, ctypes.c_char_p, _hid, _hid, _hid, _hid, _hid]),
            "H5Dopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Dwrite": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dread": (_herr, [_hid, _hid, _hid, _hid, _hid, ctypes.c_void_p]),
            "H5Dget_space": (_hid, [_hid]),
            "H5Dget_type": (_hid, [_hid]),
            "H5Dclose": (_herr, [_hid]),
            "H5Tget_class": (ctypes.c_int, [_hid]),
            "H5Tclose": (_herr, [_hid]),
            "H5Gopen2": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Gclose": (_herr, [_hid]),
            "H5Oopen": (_hid, [_hid, ctypes.c_char_p, _hid]),
            "H5Oclose": (_herr, [_hid]),
            "H5Iget_type": (ctypes.c_int, [_hid]),
            "H5Literate": (_herr, [_hid, ctypes.c_int, ctypes.c_int, ctypes.POINTER(_hsize), _ITER, ctypes.c_void_p]),
        }
        try:
            for name, (res, args) in sig.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
        except AttributeError as e:
            tried.append("%s (%s)" % (path, e))
            continue
        if lib.H5open() < 0:
            tried.append("%s (H5open failed)" % path)
            continue
        _lib = lib
        return lib
    _err = "libhdf5 not found: " + "; ".join(tried)
    raise OSError(_err)


def available():
    try:
        load()
        return True
    except OSError:
        return False


def _global(name):
    return _hid.in_dll(load(), name).value


def _check(v, what):
    if v < 0:
        raise OSError("HDF5: %s failed" % what)
    return v


def write(filename, flat):
    """Create (or overwrite) `filename` with the datasets {path: array}: real arrays as f64,
    integer arrays as u64 (i64 when negative), complex arrays as <path>_re / <path>_im, scalars
    as one-element datasets."""
    L = load()
    f = _check(L.H5Fcreate(filename.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "H5Fcreate " + filename)
    lcpl = _check(L.H5Pcreate(_global("H5P_CLS_LINK_CREATE_ID_g")), "H5Pcreate")
    try:
        _check(L.H5Pset_create_intermediate_group(lcpl, 1), "H5Pset_create_intermediate_group")
        items = []
        for k, v in flat.items():
            a = np.asarray(v)
            if a.ndim == 0:
                a = a.reshape(1)
            if np.iscomplexobj(a):
                items += [(k + "_re", a.real), (k + "_im", a.imag)]
            else:
                items.append((k, a))
        for k, a in items:
            if np.issubdtype(a.dtype, np.integer) or a.dtype == bool:
                neg = a.size and a.min() < 0
                a = np.ascontiguousarray(a, dtype=np.int64 if neg else np.uint64)
                ftype = _global("H5T_NATIVE_LLONG_g" if neg else "H5T_NATIVE_ULLONG_g")
            else:
                a = np.ascontiguousarray(a, dtype=np.float64)
                ftype = _global("H5T_NATIVE_DOUBLE_g")
            dims = (_hsize * a.ndim)(*a.shape)
            sp = _check(L.H5Screate_simple(a.ndim, dims, None), "H5Screate_simple")
            try:
                d = _check(L.H5Dcreate2(f, k.encode(), ftype, sp, lcpl, H5P_DEFAULT, H5P_DEFAULT), "H5Dcreate2 " + k)
                try:
                    _check(L.H5Dwrite(d, ftype, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data), "H5Dwrite " + k)
                finally:
                    L.H5Dclose(d)
            finally:
                L.H5Sclose(sp)
    finally:
        L.H5Pclose(lcpl)
        L.H5Fclose(f)


def _read_dataset(L, loc, name):
    d = _check(L.H5Dopen2(loc, name.encode(), H5P_DEFAULT), "H5Dopen2 " + name)
    try:
        sp = _check(L.H5Dget_space(d), "H5Dget_space")
        nd = L.H5Sget_simple_extent_ndims(sp)
        dims = (_hsize * max(nd, 1))()
        L.H5Sget_simple_extent_dims(sp, dims, None)
        L.H5Sclose(sp)
        shape = tuple(int(dims[i]) for i in range(nd)) if nd > 0 else (1,)
        t = L.H5Dget_type(d)
        cls = L.H5Tget_class(t)
        L.H5Tclose(t)
        if cls == H5T_INTEGER:
            out, mtype = np.empty(shape, dtype=np.int64), _global("H5T_NATIVE_LLONG_g")
        elif cls == H5T_FLOAT:
            out, mtype = np.empty(shape, dtype=np.float64), _global("H5T_NATIVE_DOUBLE_g")
        else:
            raise OSError("HDF5: dataset %s is neither integer nor floating point" % name)
        _check(L.H5Dread(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, out.ctypes.data), "H5Dread " + name)
        return out
    finally:
        L.H5Dclose(d)


def _walk(L, gid, prefix, out):
    names = []

    def collect(_g, name, _info, _data):
        names.append(name.decode())
        return 0

    cb = _ITER(collect)
    idx = _hsize(0)
    _check(L.H5Literate(gid, H5_INDEX_NAME, H5_ITER_INC, ctypes.byref(idx), cb, None), "H5Literate")
    for name in names:
        o = _check(L.H5Oopen(gid, name.encode(), H5P_DEFAULT), "H5Oopen " + name)
        try:
            kind = L.H5Iget_type(o)
        finally:
            L.H5Oclose(o)
        if kind == H5I_GROUP:
            g = _check(L.H5Gopen2(gid, name.encode(), H5P_DEFAULT), "H5Gopen2 " + name)
            try:
                _walk(L, g, prefix + name + "/", out)
            finally:
                L.H5Gclose(g)
        elif kind == H5I_DATASET:
            out[prefix + name] = _read_dataset(L, gid, name)


def read(filename):
    """Every dataset of `filename` as {path: array} (integer datasets as int64, the rest as
    float64; complex pairs are left as <path>_re / <path>_im)."""
    L = load()
    f = _check(L.H5Fopen(filename.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), "H5Fopen " + filename)
    try:
        out = {}
        _walk(L, f, "", out)
        return out
    finally:
        L.H5Fclose(f)

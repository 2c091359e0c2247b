"""C ABI of librustpde_amd.so (CPU-only checks: no device calls)."""
import ctypes
import re

from rustpde_amd import _lib


def declared_functions():
    text = open(_lib.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(?:int|double|void|const char\*)\s+\**\s*([a-z_0-9]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_header():
    assert set(declared_functions()) == set(_lib.SIGNATURES)


def test_version_and_error_string_without_device():
    lib = _lib.load()
    assert b"gfx950" in lib.rbc_version()
    assert isinstance(lib.rbc_last_error(), bytes)


def test_null_handles_are_rejected_without_device():
    lib = _lib.load()
    assert lib.rbc_update(None, 1) != 0
    assert b"null" in lib.rbc_last_error()
    out = ctypes.c_double()
    assert lib.rbc_eval(None, 0, ctypes.byref(out)) != 0
    assert lib.fs_forward(None, None, None, 1) != 0


def test_invalid_arguments_are_rejected_before_any_device_call():
    """Shape and kind errors come back as RBC_ERR_ARG with a message (the reference panics on
    them); the checks run before the library touches the device."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.rbc_create_periodic(17, 33, 1e5, 1.0, 0.01, 2.0, ctypes.byref(h)) != 0
    assert b"even" in lib.rbc_last_error()
    assert lib.rbc_create_confined(4, 4, 1e5, 1.0, 0.01, 1.0, 1, ctypes.byref(h)) != 0
    assert b">= 8" in lib.rbc_last_error()
    assert lib.rbc_create_ensemble(0, None, ctypes.byref(h)) != 0
    desc = (_lib.RbcDesc * 1)(_lib.RbcDesc(4, 4, 1e5, 1.0, 0.01, 1.0, 1))
    assert lib.rbc_create_ensemble(1, ctypes.cast(desc, ctypes.c_void_p), ctypes.byref(h)) != 0
    assert b">= 8" in lib.rbc_last_error()
    assert lib.fs_space2_create(1, 16, 4, 16, ctypes.byref(h)) != 0  # Fourier on axis 1
    assert b"axis 1" in lib.rbc_last_error()
    assert lib.fs_space2_create(9, 16, 1, 16, ctypes.byref(h)) != 0
    assert b"kind" in lib.rbc_last_error()
    assert lib.rbc_solver_create(7, 1, 16, 1, 16, 1e-3, 1e-3, ctypes.byref(h)) != 0
    assert b"solver" in lib.rbc_last_error()


def test_null_handles_are_rejected():
    """Entry points given a null handle return an error code with a message instead of
    dereferencing it."""
    lib = _lib.load()
    assert lib.rbc_set_option(None, 7, 1) != 0
    assert lib.rbc_ensemble_set_option(None, 7, 1) != 0
    assert b"null" in lib.rbc_last_error()
    assert lib.rbc_ensemble_update(None, 1) != 0
    assert lib.rbc_update(None, 1) != 0

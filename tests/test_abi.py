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

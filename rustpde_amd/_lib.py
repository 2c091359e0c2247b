"""Loader of librustpde_amd.so (the HIP/C-ABI library) with ctypes signatures.

The library is built in-tree (rustpde_amd/librustpde_amd.so, see __graft_entry__.build()).
There is no CPU fallback: if the library is missing, every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librustpde_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "rustpde_amd.h")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i, _d = ctypes.c_int, ctypes.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "rbc_last_error": (ctypes.c_char_p, []),
    "rbc_version": (ctypes.c_char_p, []),
    "rbc_create_confined": (_i, [_i, _i, _d, _d, _d, _d, _i, ctypes.POINTER(_vp)]),
    "rbc_create_periodic": (_i, [_i, _i, _d, _d, _d, _d, ctypes.POINTER(_vp)]),
    "rbc_destroy": (_i, [_vp]),
    "rbc_field_shape": (_i, [_vp, _i, _i, _ip, _ip, _ip]),
    "rbc_get_field": (_i, [_vp, _i, _i, _dp]),
    "rbc_set_field": (_i, [_vp, _i, _i, _dp]),
    "rbc_forward": (_i, [_vp, _i]),
    "rbc_backward": (_i, [_vp, _i]),
    "rbc_set_temp_bc": (_i, [_vp, _dp]),
    "rbc_set_velocity": (_i, [_vp, _d, _d, _d]),
    "rbc_set_temperature": (_i, [_vp, _d, _d, _d]),
    "rbc_update": (_i, [_vp, _i]),
    "rbc_synchronize": (_i, [_vp]),
    "rbc_get_time": (_d, [_vp]),
    "rbc_get_dt": (_d, [_vp]),
    "rbc_set_time": (_i, [_vp, _d]),
    "rbc_eval": (_i, [_vp, _i, _dp]),
    "rbc_exit": (_i, [_vp, _ip]),
    "rbc_set_solid": (_i, [_vp, _dp, _dp]),
    "rbc_stats_enable": (_i, [_vp, _d, _d]),
    "rbc_stats_update": (_i, [_vp, _ip]),
    "rbc_stats_get": (_i, [_vp, _i, _i, _dp]),
    "rbc_stats_set": (_i, [_vp, _i, _dp]),
    "rbc_stats_info": (_i, [_vp, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_longlong)]),
    "rbc_stats_set_info": (_i, [_vp, _d, _d, ctypes.c_longlong]),
    "rbc_tensor_info": (_i, [_vp, _i, _ip, _ip]),
    "rbc_get_tensor": (_i, [_vp, _i, _dp, _dp, _dp]),
    "rbc_set_tensor": (_i, [_vp, _i, _dp, _dp, _dp]),
    "rbc_set_option": (_i, [_vp, _i, _i]),
    "rbc_timing_enable": (_i, [_vp, _i]),
    "rbc_timing_read": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong), _dp, _dp]),
    "fs_space2_create": (_i, [_i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "fs_space2_destroy": (_i, [_vp]),
    "fs_space2_shape": (_i, [_vp, _ip, _ip, _ip, _ip, _ip, _ip, _ip]),
    "fs_forward": (_i, [_vp, _vp, _vp, _i]),
    "fs_backward": (_i, [_vp, _vp, _vp, _i]),
    "fs_to_ortho": (_i, [_vp, _vp, _vp, _i]),
    "fs_from_ortho": (_i, [_vp, _vp, _vp, _i]),
    "fs_gradient": (_i, [_vp, _vp, _vp, _i, _i, _d, _d, _i]),
    "rbc_solver_create": (_i, [_i, _i, _i, _i, _i, _d, _d, ctypes.POINTER(_vp)]),
    "rbc_solver_destroy": (_i, [_vp]),
    "rbc_solver_shape": (_i, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "rbc_solver_solve": (_i, [_vp, _vp, _vp, _i]),
    "rbc_solver_solve_noalloc": (_i, [_vp, _vp, _vp, _i]),
    "rbc_solver_set_mixed": (_i, [_vp, _i]),
    "rbc_solver_residual": (_i, [_vp, _vp, _vp, _i, _dp]),
    "rbc_solver_tensor_info": (_i, [_vp, _ip, _ip]),
    "rbc_solver_get_tensor": (_i, [_vp, _dp, _dp, _dp]),
    "rbc_solver_set_tensor": (_i, [_vp, _dp, _dp, _dp]),
    "rbc_adjoint_create_confined": (_i, [_i, _i, _d, _d, _d, _d, _i, ctypes.POINTER(_vp)]),
    "rbc_adjoint_create_periodic": (_i, [_i, _i, _d, _d, _d, _d, ctypes.POINTER(_vp)]),
    "rbc_adjoint_destroy": (_i, [_vp]),
    "rbc_adjoint_part": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "rbc_adjoint_update": (_i, [_vp, _i]),
    "rbc_adjoint_get_time": (_d, [_vp]),
    "rbc_adjoint_get_dt": (_d, [_vp]),
    "rbc_adjoint_set_time": (_i, [_vp, _d]),
    "rbc_adjoint_residual": (_i, [_vp, _i, _i, _dp]),
    "rbc_adjoint_norms": (_i, [_vp, _dp]),
    "rbc_adjoint_exit": (_i, [_vp, _ip]),
    "rbc_dev_malloc": (_i, [ctypes.POINTER(_vp), ctypes.c_ulonglong]),
    "rbc_dev_free": (_i, [_vp]),
    "rbc_memcpy_h2d": (_i, [_vp, _vp, ctypes.c_ulonglong]),
    "rbc_memcpy_d2h": (_i, [_vp, _vp, ctypes.c_ulonglong]),
    "rbc_memcpy_d2d": (_i, [_vp, _vp, ctypes.c_ulonglong]),
    "rbc_create_ensemble": (_i, [_i, _vp, ctypes.POINTER(_vp)]),
    "rbc_ensemble_destroy": (_i, [_vp]),
    "rbc_ensemble_size": (_i, [_vp, _ip, _ip]),
    "rbc_ensemble_instance": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
    "rbc_ensemble_update": (_i, [_vp, _i]),
    "rbc_ensemble_synchronize": (_i, [_vp]),
    "rbc_ensemble_set_option": (_i, [_vp, _i, _i]),
    "rbc_ensemble_timing_enable": (_i, [_vp, _i]),
    "rbc_ensemble_timing_read": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_longlong), _dp, _dp]),
}

_lib = None


class RbcDesc(ctypes.Structure):
    """rbc_desc of include/rustpde_amd.h: one confined instance of an ensemble."""
    _fields_ = [("nx", _i), ("ny", _i), ("ra", _d), ("pr", _d), ("dt", _d), ("aspect", _d), ("adiabatic", _i)]


class RbcError(RuntimeError):
    """Error status returned by librustpde_amd (message from rbc_last_error)."""


def load():
    """Load the library (once) and bind every exported signature; raises if absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RbcError(
                "librustpde_amd.so is not built (expected at %s); run __graft_entry__.build()" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != 0:
        raise RbcError("librustpde_amd error %d: %s" % (rc, load().rbc_last_error().decode()))
    return rc


def dptr(arr):
    return arr.ctypes.data_as(_dp)

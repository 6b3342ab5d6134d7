"""Loads ``libalgebra_linear_hip.so`` (built in-tree for gfx950) and declares its C ABI.

There is no CPU fallback: if the library is missing, importing the package fails with a
message that says how to build it, and calls on a machine without a GPU return
``AL_ERR_NO_DEVICE`` (raised as ``DeviceError``).
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libalgebra_linear_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "algebra_linear_hip.h")

AL_OK, AL_SINGULAR, AL_BAD_SHAPE, AL_CONVERGENCE = 0, 1, 2, 3
AL_ERR_DEVICE, AL_ERR_NO_DEVICE, AL_ERR_OOM, AL_ERR_ARG = -1, -2, -3, -4

_dp = ctypes.POINTER(ctypes.c_double)


class IrOptions(ctypes.Structure):
    """``al_ir_options``: refinement controls of the mixed-precision solve."""
    _fields_ = [("max_iter", ctypes.c_int), ("allow_fallback", ctypes.c_int), ("tol", ctypes.c_double),
                ("stop_tol", ctypes.c_double)]


class IrStats(ctypes.Structure):
    """``al_ir_stats``: what the mixed-precision solve did."""
    _fields_ = [("iterations", ctypes.c_int), ("fell_back", ctypes.c_int), ("factor_breakdown", ctypes.c_int),
                ("converged", ctypes.c_int), ("backward_error", ctypes.c_double),
                ("backward_error_low", ctypes.c_double)]


AL_FACTOR_FP64, AL_FACTOR_FP32, AL_FACTOR_BF16 = 0, 1, 2
AL_BATCH_GESV, AL_BATCH_POSV = 0, 1


class Request(ctypes.Structure):
    """``al_request``: one request to the solve server."""
    _fields_ = [("kind", ctypes.c_int32), ("factor", ctypes.c_int32), ("n", ctypes.c_size_t),
                ("nrhs", ctypes.c_size_t), ("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("x", ctypes.c_void_p),
                ("factor_id", ctypes.c_int64), ("keep", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Result(ctypes.Structure):
    """``al_result``: what the solve server did with a request."""
    _fields_ = [("status", ctypes.c_int32), ("steps", ctypes.c_int32), ("factor_id", ctypes.c_int64),
                ("batch_size", ctypes.c_int32), ("reserved", ctypes.c_int32), ("queue_ms", ctypes.c_double),
                ("service_ms", ctypes.c_double)]


class BatchEntry(ctypes.Structure):
    """``al_batch_entry``: one system of a batched solve."""
    _fields_ = [("a_off", ctypes.c_size_t), ("b_off", ctypes.c_size_t), ("n", ctypes.c_int32),
                ("nrhs", ctypes.c_int32), ("kind", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Complex(ctypes.Structure):
    """``al_complex``: the crate's Complex64 (num-complex ``Complex<f64>``, ``#[repr(C)]``)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]

    @classmethod
    def of(cls, v) -> "Complex":
        v = complex(v)
        return cls(v.real, v.imag)


_sp = ctypes.POINTER(ctypes.c_size_t)
_ip = ctypes.POINTER(ctypes.c_int32)
_d, _sz, _i, _v = ctypes.c_double, ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p

SIGNATURES = {
    "al_last_error": (ctypes.c_char_p, []),
    "al_version": (ctypes.c_char_p, []),
    "al_device_count": (_i, []),
    "al_dgemm": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz, _sz]),
    "al_dgemm_tn": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz, _sz]),
    "al_dgemm_nt": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz, _sz]),
    "al_dgemm_tt": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz, _sz]),
    "al_dtrsm": (_i, [_d, _dp, _sz, _dp, _sz, _sz, _sz, _i, _i, _i]),
    "al_dsyrk": (_i, [_d, _dp, _sz, _d, _dp, _sz, _sz, _sz, _i]),
    "al_dgetrf": (_i, [_dp, _sz, _sz, _sp, _sz, _dp]),
    "al_dgetrf_rect": (_i, [_dp, _sz, _sz, _sz, _sp, _sz]),
    "al_dgetrs": (_i, [_dp, _sz, _sp, _sz, _dp, _sz]),
    "al_dgesv": (_i, [_dp, _sz, _dp, _sz]),
    "al_dpotrf": (_i, [_dp, _sz, _sz]),
    "al_dgesv_nrhs": (_i, [_dp, _sz, _dp, _sz, _sz, _sz, _sp]),
    "al_dgetrs_nrhs": (_i, [_dp, _sz, _sp, _sz, _dp, _sz, _sz, _sz]),
    "al_dgemv": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz]),
    "al_dgemv_trans": (_i, [_d, _dp, _sz, _dp, _sz, _d, _dp, _sz, _sz, _sz]),
    "al_dlange": (_i, [ctypes.c_char, _dp, _sz, _sz, _sz, _dp]),
    "al_dtrsv": (_i, [_dp, _sz, _dp, _sz, _sz, _i, _i]),
    "al_dtrsv_transpose": (_i, [_dp, _sz, _dp, _sz, _sz, _i, _i]),
    "al_dgetri": (_i, [_dp, _sz, _sz, _sp, _sz]),
    "al_dsyr2k": (_i, [_i, _sz, _sz, _d, _dp, _sz, _dp, _sz, _d, _dp, _sz]),
    "al_dev_dgemv": (_i, [_i, _sz, _sz, _d, _v, _sz, _v, _d, _v, _v]),
    "al_dev_dlange": (_i, [ctypes.c_char, _sz, _sz, _v, _sz, _v, _v]),
    "al_dev_dtrsv": (_i, [_i, _i, _i, _sz, _v, _sz, _v, _v]),
    "al_dev_dgetri": (_i, [_sz, _v, _sz, _v, _v, _v]),
    "al_dev_dsyr2k": (_i, [_i, _sz, _sz, _d, _v, _sz, _v, _sz, _d, _v, _sz, _v]),
    "al_dev_dgemm": (_i, [_i, _i, _sz, _sz, _sz, _d, _v, _sz, _v, _sz, _d, _v, _sz, _v]),
    "al_dev_dsyrk": (_i, [_i, _sz, _sz, _d, _v, _sz, _d, _v, _sz, _v]),
    "al_dev_dtrsm": (_i, [_i, _i, _i, _sz, _sz, _d, _v, _sz, _v, _sz, _v]),
    "al_dev_dgetrf": (_i, [_sz, _sz, _v, _sz, _v, _v, _v]),
    "al_dev_dgetrs": (_i, [_sz, _sz, _v, _sz, _v, _v, _sz, _v]),
    "al_zgemm": (_i, [Complex, _dp, _sz, _dp, _sz, Complex, _dp, _sz, _sz, _sz, _sz]),
    "al_zgemm_tn": (_i, [Complex, _dp, _sz, _dp, _sz, Complex, _dp, _sz, _sz, _sz, _sz]),
    "al_zgemm_nt": (_i, [Complex, _dp, _sz, _dp, _sz, Complex, _dp, _sz, _sz, _sz, _sz]),
    "al_zgemm_tt": (_i, [Complex, _dp, _sz, _dp, _sz, Complex, _dp, _sz, _sz, _sz, _sz]),
    "al_zgetrf": (_i, [_dp, _sz, _sz, _sp, _sz]),
    "al_zgetrf_rect": (_i, [_dp, _sz, _sz, _sz, _sp, _sz]),
    "al_zgesv": (_i, [_dp, _sz, _dp, _sz]),
    "al_zpotrf": (_i, [_dp, _sz, _sz, _i]),
    "al_dev_zgemm": (_i, [_i, _i, _sz, _sz, _sz, Complex, _v, _sz, _v, _sz, Complex, _v, _sz, _v]),
    "al_dev_zgetrf": (_i, [_sz, _sz, _v, _sz, _v, _v, _v]),
    "al_dev_zgetrs": (_i, [_sz, _sz, _v, _sz, _v, _v, _sz, _v]),
    "al_dev_zpotrf": (_i, [_i, _sz, _v, _sz, _v, _v]),
    "al_dev_dpotrf": (_i, [_sz, _v, _sz, _v, _v]),
    "al_dev_dlaswp": (_i, [_sz, _v, _sz, _sz, _sz, _v, _v]),
    "al_dev_sgemm": (_i, [_i, _i, _sz, _sz, _sz, ctypes.c_float, _v, _sz, _v, _sz, ctypes.c_float, _v, _sz, _v]),
    "al_dev_gemm_bf16": (_i, [_sz, _sz, _sz, ctypes.c_float, _v, _sz, _v, _sz, ctypes.c_float, _v, _sz, _v]),
    "al_dev_batch_solve": (_i, [_sz, _v, _v, _sz, _v, _v, _v, _i, _v]),
    "al_dev_batch_solve_noalloc": (_i, [_sz, _v, _v, _v, _v, _v, _v]),
    "al_batch_tune_set": (_i, [_i, _i, _i]),
    "al_batch_tune_get": (_i, [_i, ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "al_server_create": (_v, [_sz, _i]),
    "al_server_submit": (_i, [_v, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    "al_server_wait": (_i, [_v, ctypes.c_uint64, ctypes.POINTER(Result)]),
    "al_server_release_factor": (_i, [_v, ctypes.c_int64]),
    "al_server_destroy": (None, [_v]),
    "al_ir_default_options": (None, [ctypes.POINTER(IrOptions)]),
    "al_dgesv_mixed": (_i, [_dp, _sz, _dp, _sz, _sz, _sz, _i, ctypes.POINTER(IrOptions), ctypes.POINTER(IrStats)]),
    "al_dev_dgesv_mixed": (_i, [_sz, _sz, _v, _sz, _v, _sz, _v, _sz, _i, ctypes.POINTER(IrOptions),
                                ctypes.POINTER(IrStats), _v]),
    "al_debug_set_wait_limit": (None, [ctypes.c_longlong]),
    "al_debug_coop_units_in_use": (_i, []),
    "al_debug_panel_pin_half": (_i, [_v]),
    "al_debug_set_panel_pinning": (None, [_i]),
    "al_prof_enable": (None, [_i]),
    "al_prof_reset": (None, []),
    "al_prof_query": (_i, [ctypes.c_char_p, ctypes.POINTER(_i), ctypes.POINTER(_d), ctypes.POINTER(_d),
                           ctypes.POINTER(_d)]),
}

_lib = None


def load() -> ctypes.CDLL:
    """The loaded library (raises ImportError if it has not been built).

    If PyTorch is installed it is imported first: torch bundles its own HIP runtime, and
    loading it after this library would put two HIP runtimes in one process (torch's would
    then see no device, and streams/pointers could not be shared).  Loaded first, torch's
    runtime satisfies this library's libamdhip64.so.7 dependency as well.
    """
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `make -C {HERE}` (hipcc, gfx950) or "
                "`python -c 'import __graft_entry__ as g; g.build()'`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def last_error() -> str:
    return load().al_last_error().decode()

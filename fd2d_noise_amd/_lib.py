"""Loader of the in-tree HIP library ``libfd2d_noise_amd.so``.

The library is built in place by ``fd2d_noise_amd/build.py``. There is no fallback: if it is
missing or cannot be loaded, every solver call raises.
"""
from __future__ import annotations

import ctypes
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfd2d_noise_amd.so")

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run fd2d_noise_amd/build.py (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        lib.fd2d_last_error.restype = ctypes.c_char_p
        lib.fd2d_last_error.argtypes = []
        lib.fd2d_version.restype = ctypes.c_char_p
        lib.fd2d_version.argtypes = []
        lib.fd2d_query_output_dims.restype = ctypes.c_int
        lib.fd2d_query_output_dims.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_double, ctypes.c_int64,
                                               ctypes.POINTER(ctypes.c_int64 * 3)]
        for name, at in (("fd2d_run_forward1_green_device", abi.GREEN_ARGTYPES + [ctypes.c_void_p]),
                         ("fd2d_run_forward2_correlation_device", abi.CORR_ARGTYPES + [ctypes.c_void_p]),
                         ("fd2d_run_noise_adjoint_device", abi.ADJ_ARGTYPES + [ctypes.c_void_p])):
            f = getattr(lib, name)
            f.argtypes = at
            f.restype = ctypes.c_int
        vp = ctypes.c_void_p
        for name, at in (("fd2d_resident_upload", [abi.PCArray, ctypes.c_int32, ctypes.POINTER(vp)]),
                         ("fd2d_resident_alloc", [ctypes.POINTER(ctypes.c_int64 * 3), ctypes.c_int32,
                                                  ctypes.POINTER(vp)]),
                         ("fd2d_resident_view", [vp, abi.PCArray, abi.PMArray]),
                         ("fd2d_resident_download", [vp, abi.PMArray, vp]),
                         ("fd2d_resident_free", [vp])):
            f = getattr(lib, name)
            f.argtypes = at
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


def last_error() -> str:
    return load().fd2d_last_error().decode()


def version() -> str:
    return load().fd2d_version().decode()


# Every symbol include/fd2d_noise.h declares.
EXPORTS = (
    "fd2d_run_forward1_green", "fd2d_run_forward2_correlation", "fd2d_run_noise_adjoint",
    "fd2d_run_forward1_green_device", "fd2d_run_forward2_correlation_device",
    "fd2d_run_noise_adjoint_device", "fd2d_query_output_dims", "fd2d_last_error", "fd2d_version",
    "fd2d_timing_enable", "fd2d_timing_read", "fd2d_set_steps_per_launch",
    "fd2d_make_adjoint_sources", "fd2d_make_adjoint_sources_device", "fd2d_dot_device", "fd2d_axpby_device",
    "fd2d_gaussblur2d_device", "fd2d_map_m_to_parameters_device", "fd2d_map_gradparameters_to_gradm_device",
    "fd2d_regularization_device", "fd2d_general_source_device", "fd2d_axpby_f32_device",
    "fd2d_engine_create", "fd2d_engine_submit", "fd2d_engine_advance", "fd2d_engine_done", "fd2d_engine_destroy",
    "fd2d_engine_progress", "fd2d_engine_cancel",
    "fd2d_run_batch_device", "fd2d_solve_device_bytes", "fd2d_release_cached_memory",
    "fd2d_make_adjoint_sources_queued", "fd2d_make_adjoint_sources_noalloc",
    "fd2d_resident_upload", "fd2d_resident_alloc", "fd2d_resident_view", "fd2d_resident_download",
    "fd2d_resident_free",
)


class Timing(ctypes.Structure):
    """fd2d_timing of include/fd2d_noise.h."""
    _fields_ = [("launches", (ctypes.c_int64 * 2) * 3), ("ms", (ctypes.c_double * 2) * 3),
                ("points", ctypes.c_int64 * 3)]


def timing_enable(every: int) -> None:
    load().fd2d_timing_enable(int(every))


def set_steps_per_launch(steps: int) -> None:
    """0 (default): tuned per grid shape by the first solve of that shape; 1: one time step per
    launch; 2: two consecutive steps per launch where possible. Results are the same in every case."""
    if load().fd2d_set_steps_per_launch(int(steps)) != abi.FD2D_OK:
        raise ValueError(last_error())


def timing_read() -> dict:
    t = Timing()
    load().fd2d_timing_read(ctypes.byref(t))
    out = {}
    for q, name in enumerate(("green", "correlation", "adjoint")):
        out[name] = {"launches": int(t.launches[q][0]), "steps": int(t.launches[q][1]),
                     "ms": float(t.ms[q][0]), "points": int(t.points[q])}
    return out

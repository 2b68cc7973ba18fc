"""Misfits and adjoint sources on the device (tools/measurement/make_adjoint_sources.m:31).

    misfit_n, adstf = make_adjoint_sources(u, u_0, du, t, veldis, measurement, src, rec, deriv_order)

Same arguments and outputs as the reference (src and rec only fix the receiver count there);
``window`` is the causal measurement window, [100, 200] s in the reference (lines 85-86). The work
runs in FP64 with compensated sums in fd2d_make_adjoint_sources (csrc/caller.hip). The device form
takes and returns torch CUDA tensors (MATLAB nr x nt arrays as (nt, nr) row-major tensors).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib, abi

MEASUREMENTS = {"waveform_difference": 0, "cc_time_shift": 1, "amplitude_difference": 2, "log_amplitude_ratio": 3}
_dp = ctypes.POINTER(ctypes.c_double)


def _bind():
    lib = _lib.load()
    f = lib.fd2d_make_adjoint_sources
    f.argtypes = [_dp, _dp, _dp, _dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                  ctypes.c_double, ctypes.c_double, _dp, _dp]
    f.restype = ctypes.c_int
    g = lib.fd2d_make_adjoint_sources_device
    g.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]
    g.restype = ctypes.c_int
    q = lib.fd2d_make_adjoint_sources_queued
    q.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p]
    q.restype = ctypes.c_int
    return f, g


def _codes(veldis, measurement, deriv_order):
    if measurement not in MEASUREMENTS:
        raise ValueError(f"measurement {measurement!r}")
    return (1 if veldis == "vel" else 0), MEASUREMENTS[measurement], (2 if deriv_order == "2nd" else 1)


def make_adjoint_sources(u, u_0, du, t, veldis, measurement, src=None, rec=None, deriv_order="1st",
                         window=(100.0, 200.0)):
    f, _ = _bind()
    u = np.asfortranarray(np.atleast_2d(u), dtype=np.float64)
    u0 = np.asfortranarray(np.atleast_2d(u_0), dtype=np.float64)
    nr, nt = u.shape
    d = None if du is None else np.asfortranarray(np.atleast_2d(du), dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    misfit = np.zeros(nr)
    adstf = np.zeros((nr, nt), order="F")
    v, m, o = _codes(veldis, measurement, deriv_order)
    p = lambda a: a.ctypes.data_as(_dp)
    st = f(p(u), p(u0), p(d) if d is not None else None, p(t), nr, nt, v, m, o, float(window[0]), float(window[1]),
           p(misfit), p(adstf))
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())
    return misfit, adstf


def make_adjoint_sources_device(u, u_0, du, t, veldis, measurement, deriv_order="1st", window=(100.0, 200.0),
                                misfit=None, adstf=None, t_host=None):
    """Device form: u, u_0, du (nt, nr) float64 CUDA tensors (MATLAB nr x nt), t an nt tensor. With
    t_host (the same time axis as a host float64 array) the work is only enqueued on the current
    stream (fd2d_make_adjoint_sources_queued); without it the call waits for the stream."""
    import torch
    lib = _lib.load()
    _bind()
    nt, nr = u.shape
    misfit = torch.zeros(nr, dtype=torch.float64, device=u.device) if misfit is None else misfit
    adstf = torch.zeros((nt, nr), dtype=torch.float64, device=u.device) if adstf is None else adstf
    v, m, o = _codes(veldis, measurement, deriv_order)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dup = du.data_ptr() if du is not None else None
    if t_host is not None:
        th = np.ascontiguousarray(t_host, dtype=np.float64)
        if th.size != nt:
            raise ValueError("t_host must hold nt samples")
        st = lib.fd2d_make_adjoint_sources_queued(u.data_ptr(), u_0.data_ptr(), dup, t.data_ptr(), th.ctypes.data, nr,
                                                  nt, v, m, o, float(window[0]), float(window[1]), misfit.data_ptr(),
                                                  adstf.data_ptr(), stream)
    else:
        st = lib.fd2d_make_adjoint_sources_device(u.data_ptr(), u_0.data_ptr(), dup, t.data_ptr(), nr, nt, v, m, o,
                                                  float(window[0]), float(window[1]), misfit.data_ptr(),
                                                  adstf.data_ptr(), stream)
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())
    return misfit, adstf

"""Device-resident calls of the three solvers (the ``*_device`` entry points of the C ABI).

Arrays are torch CUDA tensors whose memory is the MATLAB column-major array: a MATLAB nx x nz x n
array is a contiguous tensor of shape (n, nz, nx) (reversed dimensions). ``to_device`` and
``to_host`` convert between MATLAB-shaped numpy arrays and such tensors. Outputs are allocated by
the caller (``empty_like_dims``), so forward snapshots stay in HBM and can be handed to the adjoint
run without a host round trip. Calls run on torch's current stream and return after it has
synchronised.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib, abi
from .config import Config


def to_device(a, dtype=None, device="cuda") -> torch.Tensor:
    a = np.asfortranarray(a if dtype is None else np.asarray(a, dtype=dtype))
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device)


def to_host(t: torch.Tensor) -> np.ndarray:
    return t.detach().cpu().numpy().T


def release_cached_memory() -> None:
    """Returns the device blocks the library keeps for later solves and those torch's caching
    allocator holds unused (fd2d_release_cached_memory, torch.cuda.empty_cache); waits for the device."""
    _check(_lib.load().fd2d_release_cached_memory())
    torch.cuda.empty_cache()


def empty_like_dims(dims: Sequence[int], dtype, device="cuda") -> torch.Tensor:
    dims = [int(d) for d in dims]
    while len(dims) > 1 and dims[-1] == 1:
        dims.pop()
    shape = tuple(reversed(dims))
    try:
        return torch.zeros(shape, dtype=dtype, device=device)
    except torch.OutOfMemoryError:
        if torch.device(device).type != "cuda":
            raise
        # the free memory may sit in the library's cached blocks or in torch's, split into blocks
        # of other sizes: return both to the device and try once more
        release_cached_memory()
        return torch.zeros(shape, dtype=dtype, device=device)


def _mdims(t: torch.Tensor):
    return abi.matlab_dims(tuple(reversed(t.shape)))


def _c(t: Optional[torch.Tensor]):
    if t is None or t.numel() == 0:
        return None
    assert t.is_cuda and t.is_contiguous()
    d = abi.CArray()
    d.data = t.data_ptr()
    d.dims[:] = _mdims(t)
    return ctypes.pointer(d)


def _m(t: Optional[torch.Tensor]):
    if t is None or t.numel() == 0:
        return None
    assert t.is_cuda and t.is_contiguous()
    d = abi.MArray()
    d.data = t.data_ptr()
    d.dims[:] = _mdims(t)
    return ctypes.pointer(d)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(st):
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())


def run_forward1_green(mu, rho, src, rec, mode, dmu, G_in, *, config: Config, G_fft, G_out=None, seismograms=None):
    cc = config.to_c()
    st = _lib.load().fd2d_run_forward1_green_device(ctypes.byref(cc), _c(mu), _c(rho), _c(src), _c(rec), float(mode),
                                                    _c(dmu), _c(G_in), _m(G_fft), _m(G_out), _m(seismograms),
                                                    _stream())
    _check(st)


def run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in, *, config: Config,
                             seismograms=None, C_out=None):
    cc = config.to_c()
    st = _lib.load().fd2d_run_forward2_correlation_device(
        ctypes.byref(cc), _c(mu), _c(rho), _c(G_fft), _c(spectrum), _c(source_dist), _c(rec), float(mode), _c(dmu),
        _c(C_in), _m(seismograms), _m(C_out), _stream())
    _check(st)


def run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in, *,
                      config: Config, K=None, stf_fft=None, u_out=None, seismograms=None):
    cc = config.to_c()
    st = _lib.load().fd2d_run_noise_adjoint_device(
        ctypes.byref(cc), _c(mu), _c(rho), _c(u_fwd), _c(stf), _c(src), _c(rec), _c(spectrum), _c(source_dist),
        _c(G_fft), float(mode), _c(dmu), _c(u_in), _m(K), _m(stf_fft), _m(u_out), _m(seismograms), _stream())
    _check(st)

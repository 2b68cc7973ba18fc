"""Device-resident arrays behind the C ABI's ``fd2d_resident`` handles (include/fd2d_noise.h).

A ``Resident`` owns device memory holding one MATLAB-layout array (column-major, complex
interleaved), so models, Green spectra and snapshots stay on the GPU between solver calls without
torch: ``carray()`` / ``marray()`` are the device descriptors the ``*_device`` entry points and
``fd2d_solve`` take. The reference moves every array through the MEX boundary by value; this is
the handle a caller outside Python (a MEX shim, a service) would hold instead.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _lib, abi

ELEM_BYTES = {np.dtype(np.float32): 4, np.dtype(np.float64): 8, np.dtype(np.complex128): 16}


def _check(st: int) -> None:
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())


class Resident:
    def __init__(self, handle: ctypes.c_void_p, shape: Sequence[int], dtype):
        self._h = handle
        self.shape = tuple(int(d) for d in shape)
        self.dtype = np.dtype(dtype)

    @classmethod
    def upload(cls, a, dtype=np.float64) -> "Resident":
        """A resident copy of the MATLAB-shaped array ``a``."""
        arr = np.asfortranarray(np.asarray(a, dtype=dtype))
        d = abi.CArray()
        d.data = arr.ctypes.data if arr.size else None
        d.dims[:] = abi.matlab_dims(arr.shape)
        h = ctypes.c_void_p()
        _check(_lib.load().fd2d_resident_upload(ctypes.byref(d), ELEM_BYTES[arr.dtype], ctypes.byref(h)))
        return cls(h, arr.shape, arr.dtype)

    @classmethod
    def alloc(cls, shape: Sequence[int], dtype=np.float64) -> "Resident":
        """A zero-filled resident array (an output)."""
        dims = (ctypes.c_int64 * 3)(*abi.matlab_dims(shape))
        h = ctypes.c_void_p()
        _check(_lib.load().fd2d_resident_alloc(ctypes.byref(dims), ELEM_BYTES[np.dtype(dtype)], ctypes.byref(h)))
        return cls(h, shape, dtype)

    def carray(self) -> Optional[ctypes.pointer]:
        """Input descriptor (None for an empty array, as the entry points read ``[]``)."""
        if int(np.prod(self.shape)) == 0:
            return None
        d = abi.CArray()
        _check(_lib.load().fd2d_resident_view(self._h, ctypes.byref(d), None))
        return ctypes.pointer(d)

    def marray(self) -> Optional[ctypes.pointer]:
        if int(np.prod(self.shape)) == 0:
            return None
        d = abi.MArray()
        _check(_lib.load().fd2d_resident_view(self._h, None, ctypes.byref(d)))
        return ctypes.pointer(d)

    def download(self, stream: Optional[int] = None) -> np.ndarray:
        """The array in host memory, after the work queued on ``stream`` (a hipStream_t value)."""
        out = np.zeros(self.shape, dtype=self.dtype, order="F")
        d = abi.MArray()
        d.data = out.ctypes.data if out.size else None
        d.dims[:] = abi.matlab_dims(self.shape)
        _check(_lib.load().fd2d_resident_download(self._h, ctypes.byref(d), ctypes.c_void_p(stream)))
        return out

    def free(self) -> None:
        if self._h:
            _check(_lib.load().fd2d_resident_free(self._h))
            self._h = None

    def __enter__(self) -> "Resident":
        return self

    def __exit__(self, *exc) -> None:
        self.free()

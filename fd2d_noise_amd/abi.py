"""ctypes view of include/fd2d_noise.h and the MATLAB-shaped host calling convention.

Arrays cross the C ABI as ``{data, dims[3]}`` descriptors in MATLAB column-major order. This module
turns numpy arrays into descriptors, allocates outputs with the sizes the reference returns and
binds the three solver signatures of a shared library. ``HostSolvers`` works for any library that
exports the three host entry points under a common prefix.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from .config import Config

FD2D_OK, FD2D_EINVAL, FD2D_ENOMEM, FD2D_EDEVICE, FD2D_EUNSUPPORTED, FD2D_ENONFINITE = 0, 1, 2, 3, 4, 5
FD2D_GREEN, FD2D_CORRELATION, FD2D_ADJOINT = 1, 2, 3
# fd2d_solve_state: states of an engine's solve
(FD2D_SOLVE_RUNNING, FD2D_SOLVE_FINISHING, FD2D_SOLVE_COMPLETE, FD2D_SOLVE_CANCELLED,
 FD2D_SOLVE_FAILED) = 0, 1, 2, 3, 4

_STATUS = {FD2D_EINVAL: "EINVAL", FD2D_ENOMEM: "ENOMEM", FD2D_EDEVICE: "EDEVICE",
           FD2D_EUNSUPPORTED: "EUNSUPPORTED", FD2D_ENONFINITE: "ENONFINITE"}


class Fd2dError(RuntimeError):
    """A solver returned a non-zero fd2d_status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"fd2d {_STATUS.get(code, code)}: {message}")
        self.code = code


class CArray(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dims", ctypes.c_int64 * 3)]


class MArray(ctypes.Structure):
    _fields_ = [("data", ctypes.c_void_p), ("dims", ctypes.c_int64 * 3)]


PCArray = ctypes.POINTER(CArray)
PMArray = ctypes.POINTER(MArray)

GREEN_ARGTYPES = [ctypes.c_void_p, PCArray, PCArray, PCArray, PCArray, ctypes.c_double, PCArray, PCArray,
                  PMArray, PMArray, PMArray]
CORR_ARGTYPES = [ctypes.c_void_p, PCArray, PCArray, PCArray, PCArray, PCArray, PCArray, ctypes.c_double,
                 PCArray, PCArray, PMArray, PMArray]
ADJ_ARGTYPES = [ctypes.c_void_p, PCArray, PCArray, PCArray, PCArray, PCArray, PCArray, PCArray, PCArray,
                PCArray, ctypes.c_double, PCArray, PCArray, PMArray, PMArray, PMArray, PMArray]


def matlab_dims(shape: Sequence[int]) -> tuple:
    """MATLAB size of an array: vectors are columns, trailing singleton dimensions are 1."""
    shape = tuple(int(s) for s in shape)
    if len(shape) == 0:
        return (1, 1, 1)
    if len(shape) == 1:
        return (shape[0], 1, 1)
    if len(shape) == 2:
        return (shape[0], shape[1], 1)
    if len(shape) == 3:
        return shape
    raise ValueError(f"arrays of rank {len(shape)} are not solver arguments")


class _Keep:
    """Holds the numpy buffers that descriptors point into for the duration of a call."""

    def __init__(self):
        self.refs = []

    def carray(self, a, dtype) -> Optional[ctypes.pointer]:
        if a is None:
            return None
        arr = np.asarray(a)
        if arr.size == 0:
            return None
        arr = np.asfortranarray(arr.astype(dtype, copy=False))
        self.refs.append(arr)
        d = CArray()
        d.data = arr.ctypes.data
        d.dims[:] = matlab_dims(arr.shape)
        return ctypes.pointer(d)

    def marray(self, shape, dtype):
        if shape is None or int(np.prod(shape)) == 0:
            return None, None
        arr = np.zeros(shape, dtype=dtype, order="F")
        self.refs.append(arr)
        d = MArray()
        d.data = arr.ctypes.data
        d.dims[:] = matlab_dims(shape)
        return arr, ctypes.pointer(d)


def _scalar_single_zero():
    # The reference returns single(0.0) for a wavefield output it did not store
    # (run_forward1_green.m:109, run_forward2_correlation.m:99, run_noise_adjoint.m:138).
    return np.zeros((), dtype=np.float32)


class HostSolvers:
    """MATLAB-shaped calls of the three host entry points of one library."""

    def __init__(self, lib: ctypes.CDLL, prefix: str, last_error):
        self.lib = lib
        self._green = getattr(lib, prefix + "run_forward1_green")
        self._corr = getattr(lib, prefix + "run_forward2_correlation")
        self._adj = getattr(lib, prefix + "run_noise_adjoint")
        for f, at in ((self._green, GREEN_ARGTYPES), (self._corr, CORR_ARGTYPES), (self._adj, ADJ_ARGTYPES)):
            f.argtypes = at
            f.restype = ctypes.c_int
        self._last_error = last_error

    def _check(self, st: int):
        if st != FD2D_OK:
            raise Fd2dError(st, self._last_error())

    @staticmethod
    def _nrows(a) -> int:
        if a is None:
            return 0
        a = np.asarray(a)
        return 0 if a.size == 0 else int(matlab_dims(a.shape)[0])

    def run_forward1_green(self, mu, rho, src, rec, mode, dmu=None, G_in=None, *, config: Config,
                           nargout: int = 3):
        """[G_fft, G_out, seismograms] = run_forward1_green(mu, rho, src, rec, mode, dmu, G_in)."""
        c = config
        k = _Keep()
        cc = c.to_c()
        nr = self._nrows(rec)
        G_fft, pG = k.marray((c.nx, c.nz, c.n_sample), np.complex128)
        G_out, pO = (k.marray((c.nx, c.nz, c.n_fw_green()), np.float32)
                     if (mode != 0 and nargout >= 2) else (None, None))
        seis, pS = k.marray((nr, c.nt), np.float64) if nargout >= 3 else (None, None)
        st = self._green(ctypes.byref(cc), k.carray(mu, np.float64), k.carray(rho, np.float64),
                         k.carray(src, np.float64), k.carray(rec, np.float64), float(mode),
                         k.carray(dmu, np.float64), k.carray(G_in, np.float32), pG, pO, pS)
        self._check(st)
        if G_fft is None:
            G_fft = np.zeros((c.nx, c.nz, c.n_sample), np.complex128, order="F")
        if G_out is None:
            G_out = _scalar_single_zero()
        if seis is None:
            seis = np.zeros((nr, c.nt))
        return G_fft, G_out, seis

    def run_forward2_correlation(self, mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu=None,
                                 C_in=None, *, config: Config, nargout: int = 2):
        """[seismograms, C_out] = run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec,
        mode, dmu, C_in)."""
        c = config
        k = _Keep()
        cc = c.to_c()
        nr = self._nrows(rec)
        seis, pS = k.marray((nr, c.n_corr), np.float64)
        C_out, pC = (k.marray((c.nx, c.nz, c.n_fw_corr()), np.float32)
                     if (mode != 0 and nargout >= 2) else (None, None))
        st = self._corr(ctypes.byref(cc), k.carray(mu, np.float64), k.carray(rho, np.float64),
                        k.carray(G_fft, np.complex128), k.carray(spectrum, np.float64),
                        k.carray(source_dist, np.float64), k.carray(rec, np.float64), float(mode),
                        k.carray(dmu, np.float64), k.carray(C_in, np.float32), pS, pC)
        self._check(st)
        if seis is None:
            seis = np.zeros((nr, c.n_corr))
        if C_out is None:
            C_out = _scalar_single_zero()
        return seis, C_out

    def run_noise_adjoint(self, mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode,
                          dmu=None, u_in=None, *, config: Config, nargout: int = 4):
        """[K, stf_fft, u_out, seismograms] = run_noise_adjoint(mu, rho, u_fwd, stf, src, rec,
        spectrum, source_dist, G_fft, mode, dmu, u_in)."""
        c = config
        k = _Keep()
        cc = c.to_c()
        nr = self._nrows(rec)
        nb_cols = max(c.n_basis_fct, 1)
        imode = int(mode)
        K, pK = k.marray((c.nx, c.nz, nb_cols + 1), np.float64)
        sf, pF = (k.marray((c.nx, c.nz, c.n_sample), np.complex128)
                  if (imode in (1, 12, 13) and nargout >= 2) else (None, None))
        uo, pU = (k.marray((c.nx, c.nz, c.n_fw_corr()), np.float32)
                  if (imode in (11, 13) and nargout >= 3) else (None, None))
        seis, pS = k.marray((nr, c.n_corr), np.float64) if nargout >= 4 else (None, None)
        st = self._adj(ctypes.byref(cc), k.carray(mu, np.float64), k.carray(rho, np.float64),
                       k.carray(u_fwd, np.float32), k.carray(stf, np.complex128), k.carray(src, np.float64),
                       k.carray(rec, np.float64), k.carray(spectrum, np.float64),
                       k.carray(source_dist, np.float64), k.carray(G_fft, np.complex128), float(mode),
                       k.carray(dmu, np.float64), k.carray(u_in, np.float32), pK, pF, pU, pS)
        self._check(st)
        if sf is None:
            sf = np.zeros((c.nx, c.nz, c.n_sample), np.complex128, order="F")
        if uo is None:
            uo = _scalar_single_zero()
        if seis is None:
            seis = np.zeros((nr, c.n_corr))
        return K, sf, uo, seis

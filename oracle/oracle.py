"""ctypes binding of oracle/_build/libfd2d_oracle.so (TEST INFRASTRUCTURE ONLY).

The oracle exports the same solver signatures as the product's C ABI with an ``oracle_`` prefix,
so the MATLAB-shaped calling code of fd2d_noise_amd.abi drives both.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from fd2d_noise_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libfd2d_oracle.so")

_lib = None
_solvers = None


def build() -> str:
    subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    return LIB_PATH


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_last_error.restype = ctypes.c_char_p
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_get_threads.restype = ctypes.c_int
        dp = ctypes.POINTER(ctypes.c_double)
        lib.oracle_dx_v.argtypes = [dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        lib.oracle_dz_v.argtypes = [dp, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        lib.oracle_div_s.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, dp]
        lib.oracle_absbound.argtypes = [ctypes.c_void_p, dp]
        lib.oracle_interpol_matrix.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        lib.oracle_interpol_matrix.restype = ctypes.c_int
        _lib = lib
    return _lib


def solvers() -> abi.HostSolvers:
    global _solvers
    if _solvers is None:
        lib = load()
        _solvers = abi.HostSolvers(lib, "oracle_", lambda: lib.oracle_last_error().decode())
    return _solvers


def set_threads(n: int) -> None:
    load().oracle_set_threads(int(n))


class primitives:
    """The code/propagation/ primitives on float64 MATLAB-shaped arrays."""

    @staticmethod
    def _p(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))

    @staticmethod
    def dx_v(v, dx, order):
        v = np.asfortranarray(v, dtype=np.float64)
        nx, nz = v.shape
        out = np.zeros((nx - 1, nz), order="F")
        load().oracle_dx_v(primitives._p(v), dx, nx, nz, order, primitives._p(out))
        return out

    @staticmethod
    def dz_v(v, dz, order):
        v = np.asfortranarray(v, dtype=np.float64)
        nx, nz = v.shape
        out = np.zeros((nx, nz - 1), order="F")
        load().oracle_dz_v(primitives._p(v), dz, nx, nz, order, primitives._p(out))
        return out

    @staticmethod
    def div_s(sxy, szy, dx, dz, order):
        sxy = np.asfortranarray(sxy, dtype=np.float64)
        szy = np.asfortranarray(szy, dtype=np.float64)
        nx, nz = szy.shape[0], sxy.shape[1]
        out = np.zeros((nx, nz), order="F")
        load().oracle_div_s(primitives._p(sxy), primitives._p(szy), dx, dz, nx, nz, order, primitives._p(out))
        return out

    @staticmethod
    def absbound(config):
        out = np.zeros((config.nx, config.nz), order="F")
        cc = config.to_c()
        load().oracle_absbound(ctypes.byref(cc), primitives._p(out))
        return out

    @staticmethod
    def interpol_matrix(n_sample, n_basis_fct):
        out = np.zeros((n_sample, max(n_basis_fct, 1)), order="F")
        st = load().oracle_interpol_matrix(n_sample, n_basis_fct, primitives._p(out))
        if st:
            raise abi.Fd2dError(st, load().oracle_last_error().decode())
        return out

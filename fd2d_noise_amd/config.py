"""Solver configuration: the tuples that fd2d_noise's solvers read inside every call.

The MATLAB solvers call ``input_parameters()`` (input/input_parameters.m:2),
``input_interferometry()`` (input/input_interferometry.m:2) and ``absorb_specs()``
(input/absorb_specs.m:2) themselves; here the same values travel with the call as a
:class:`Config`. The defaults are the reference's defaults: the "tiny setup for gradient
validation" (input_parameters.m:8-16, order 4 at :45, fw_nth = 1 at :53, n_basis_fct = 0 at :87),
f_sample = 0.02:0.0015:0.2 with freq_samp = 5 (input_interferometry.m:16-21) and a 5000 m taper on
all four sides (absorb_specs.m:8-13).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Tuple

import numpy as np


def matlab_colon(a: float, step: float, b: float) -> np.ndarray:
    """``a:step:b`` with MATLAB's element count (floor((b-a)/step) + 1, with a small tolerance)."""
    n = int(np.floor((b - a) / step + 1e-10)) + 1
    return a + step * np.arange(n, dtype=np.float64)


@dataclasses.dataclass
class Config:
    # input_parameters()
    Lx: float = 6.0e4
    Lz: float = 6.0e4
    nx: int = 50
    nz: int = 50
    dt: float = 0.09
    nt: int = 150
    order: int = 4
    n_basis_fct: int = 0
    fw_nth: int = 1
    # input_interferometry()
    f_sample: np.ndarray = dataclasses.field(default_factory=lambda: matlab_colon(0.02, 0.0015, 0.2))
    freq_samp: int = 5
    # absorb_specs()
    absorb_width: float = 5000.0
    absorb_left: int = 1
    absorb_right: int = 1
    absorb_top: int = 1
    absorb_bottom: int = 1

    # ---- derived quantities, as the reference computes them ----------------------------------
    @property
    def n_sample(self) -> int:
        return int(len(self.f_sample))

    @property
    def w_sample(self) -> np.ndarray:
        return 2.0 * np.pi * np.asarray(self.f_sample, dtype=np.float64)

    @property
    def dw(self) -> float:
        w = self.w_sample
        return float(w[1] - w[0]) if len(w) > 1 else 0.0

    @property
    def dx(self) -> float:
        return self.Lx / (self.nx - 1)

    @property
    def dz(self) -> float:
        return self.Lz / (self.nz - 1)

    @property
    def n_corr(self) -> int:
        """Steps of the correlation and adjoint runs: length(-(nt-1)dt:dt:(nt-1)dt)."""
        return 2 * self.nt - 1

    def x(self) -> np.ndarray:
        return np.arange(self.nx) * self.dx

    def z(self) -> np.ndarray:
        return np.arange(self.nz) * self.dz

    def t_green(self) -> np.ndarray:
        return np.arange(self.nt) * self.dt

    def t_corr(self) -> np.ndarray:
        return (np.arange(1, self.n_corr + 1) - self.nt) * self.dt

    def n_fw_green(self) -> int:
        return sum(1 for n in range(self.nt, 2 * self.nt) if n % self.fw_nth == 0)

    def n_fw_corr(self) -> int:
        return self.n_corr // self.fw_nth

    def input_parameters(self) -> Tuple:
        return (self.Lx, self.Lz, self.nx, self.nz, self.dt, self.nt, self.order, None, None,
                self.n_basis_fct, self.fw_nth)

    def input_interferometry(self) -> Tuple:
        return (np.asarray(self.f_sample), self.n_sample, self.w_sample, self.dw, self.freq_samp)

    def absorb_specs(self) -> Tuple:
        return (self.absorb_width, self.absorb_left, self.absorb_right, self.absorb_top,
                self.absorb_bottom)

    def replace(self, **kw) -> "Config":
        return dataclasses.replace(self, **kw)

    # ---- C ABI -------------------------------------------------------------------------------
    def to_c(self) -> "CConfig":
        w = np.ascontiguousarray(self.w_sample, dtype=np.float64)
        c = CConfig(
            Lx=self.Lx, Lz=self.Lz, nx=self.nx, nz=self.nz, dt=self.dt, nt=self.nt, order=self.order,
            n_basis_fct=self.n_basis_fct, fw_nth=self.fw_nth, n_sample=self.n_sample,
            freq_samp=self.freq_samp,
            w_sample=w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if len(w) else None,
            dw=self.dw, absorb_width=self.absorb_width, absorb_left=self.absorb_left,
            absorb_right=self.absorb_right, absorb_top=self.absorb_top,
            absorb_bottom=self.absorb_bottom)
        c._keep = w  # the struct borrows the frequency vector
        return c


class CConfig(ctypes.Structure):
    """``fd2d_config`` of include/fd2d_noise.h."""
    _fields_ = [
        ("Lx", ctypes.c_double), ("Lz", ctypes.c_double),
        ("nx", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("nt", ctypes.c_int32), ("order", ctypes.c_int32), ("n_basis_fct", ctypes.c_int32),
        ("fw_nth", ctypes.c_int32),
        ("n_sample", ctypes.c_int32), ("freq_samp", ctypes.c_int32),
        ("w_sample", ctypes.POINTER(ctypes.c_double)),
        ("dw", ctypes.c_double),
        ("absorb_width", ctypes.c_double),
        ("absorb_left", ctypes.c_int32), ("absorb_right", ctypes.c_int32),
        ("absorb_top", ctypes.c_int32), ("absorb_bottom", ctypes.c_int32),
    ]

"""fd2d_noise_amd: gfx950 (MI355X) engine for fd2d_noise's 2-D SH noise-correlation solvers.

The product path is the HIP library ``libfd2d_noise_amd.so`` behind the C ABI of
include/fd2d_noise.h; this package is its Python mirror of the MATLAB call surface.
"""
from .abi import Fd2dError
from .config import Config
from .solvers import run_forward1_green, run_forward2_correlation, run_noise_adjoint

__all__ = ["Config", "Fd2dError", "run_forward1_green", "run_forward2_correlation", "run_noise_adjoint"]

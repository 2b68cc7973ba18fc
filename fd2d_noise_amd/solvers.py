"""MATLAB call surface of the three solvers, executed by the gfx950 HIP engine.

    G_fft, G_out, seis = run_forward1_green(mu, rho, src, rec, mode, dmu, G_in, config=cfg)
    seis, C_out = run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in,
                                           config=cfg)
    K, stf_fft, u_out, seis = run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist,
                                                G_fft, mode, dmu, u_in, config=cfg)

Positional arguments and outputs are those of code/run_forward1_green.m:1,
code/run_forward2_correlation.m:1 and code/run_noise_adjoint.m:1; ``None`` or an empty array is
MATLAB's ``[]``. Arrays are numpy arrays in MATLAB shape (nx x nz fields, n x 2 positions, ...).
``nargout`` drops trailing outputs the way MATLAB's nargout does (they are then not copied back).
Errors raise :class:`fd2d_noise_amd.abi.Fd2dError` with the library's message.
"""
from __future__ import annotations

from . import _lib, abi
from .config import Config

_host = None


def _solvers() -> abi.HostSolvers:
    global _host
    if _host is None:
        _host = abi.HostSolvers(_lib.load(), "fd2d_", _lib.last_error)
    return _host


def run_forward1_green(mu, rho, src, rec, mode, dmu=None, G_in=None, *, config: Config, nargout: int = 3):
    return _solvers().run_forward1_green(mu, rho, src, rec, mode, dmu, G_in, config=config, nargout=nargout)


def run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu=None, C_in=None, *,
                             config: Config, nargout: int = 2):
    return _solvers().run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in,
                                               config=config, nargout=nargout)


def run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu=None, u_in=None, *,
                      config: Config, nargout: int = 4):
    return _solvers().run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu,
                                        u_in, config=config, nargout=nargout)

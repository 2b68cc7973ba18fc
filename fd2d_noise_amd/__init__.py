"""fd2d_noise_amd: gfx950 (MI355X) engine for fd2d_noise's 2-D SH noise-correlation solvers.

The product path is the HIP library ``libfd2d_noise_amd.so`` behind the C ABI of
include/fd2d_noise.h; this package is its Python mirror of the MATLAB call surface:

    G_fft, G_out, seis = run_forward1_green(mu, rho, src, rec, mode, dmu, G_in, config=cfg)
    seis, C_out = run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in,
                                           config=cfg)
    K, stf_fft, u_out, seis = run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist,
                                                G_fft, mode, dmu, u_in, config=cfg)

Positional arguments and outputs are those of code/run_forward1_green.m:1,
code/run_forward2_correlation.m:1 and code/run_noise_adjoint.m:1; ``None`` or an empty array is
MATLAB's ``[]``. Arrays are numpy arrays in MATLAB shape (nx x nz fields, n x 2 positions, ...).
``nargout`` drops trailing outputs the way MATLAB's nargout does (they are then not copied back).
Errors raise :class:`Fd2dError` with the library's message. The three functions are the methods of
one ``abi.HostSolvers`` over the library, bound on first use: the library is loaded then, and a
missing build fails loudly there. Device-resident forms of the same calls are in ``device``.
"""
from .abi import Fd2dError
from .config import Config

__all__ = ["Config", "Fd2dError", "run_forward1_green", "run_forward2_correlation", "run_noise_adjoint"]

_SOLVERS = ("run_forward1_green", "run_forward2_correlation", "run_noise_adjoint")
_host = None


def __getattr__(name):
    global _host
    if name in _SOLVERS:
        if _host is None:
            from . import _lib, abi
            _host = abi.HostSolvers(_lib.load(), "fd2d_", _lib.last_error)
        return getattr(_host, name)
    raise AttributeError(f"module {__name__!r} has no attribute {name!r}")

"""Batched solves: many Green / correlation / adjoint runs advanced together on the device.

The reference calls its three MEX solvers one at a time, per station and per basis function, from
its objective and gradient loops (inversion/optlib_callback_fct/eval_noise_objective_and_gradient.m
:54-160, code/run_noise_source_kernel.m, code/run_noise_structure_kernel.m). At small grids one
solve fills a fraction of the GPU, so the engine here steps every active solve in the same launch
(one grouped launch per FD order and field set), with results bit-identical to the single-solve
entry points of ``device.py``.

    specs = [batch.green(mu, rho, src_i, rec, 0, None, None, config=cfg, G_fft=G_i) for ...]
    batch.run_batch(specs)                      # all to completion

    eng = batch.Engine()                        # continuous batching
    t = eng.submit(spec); eng.advance(100); eng.done(t)

Arguments follow ``device.py`` (torch CUDA tensors in MATLAB column-major memory, caller-allocated
outputs). A spec keeps its tensors alive; the engine keeps each submitted spec until it is done.
"""
from __future__ import annotations

import ctypes
from typing import List, Optional, Sequence

import torch

from . import _lib, abi
from .config import Config
from .device import _c, _m, _check, _stream

PCArray = abi.PCArray
PMArray = abi.PMArray


class CSolve(ctypes.Structure):
    """``fd2d_solve`` of include/fd2d_noise.h."""
    _fields_ = [
        ("solver", ctypes.c_int32),
        ("mode", ctypes.c_double),
        ("cfg", ctypes.c_void_p),
        ("mu", PCArray), ("rho", PCArray), ("src", PCArray), ("rec", PCArray),
        ("G_fft", PCArray), ("spectrum", PCArray), ("source_dist", PCArray), ("stf", PCArray),
        ("u_fwd", PCArray), ("dmu", PCArray), ("born_in", PCArray),
        ("out", PMArray * 4),
    ]


def _bind():
    lib = _lib.load()
    if getattr(lib, "_batch_bound", False):
        return lib
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    for name, at in (
            ("fd2d_engine_create", [vp, ctypes.POINTER(vp)]),
            ("fd2d_engine_submit", [vp, ctypes.POINTER(CSolve), ctypes.POINTER(i64)]),
            ("fd2d_engine_advance", [vp, i32, ctypes.POINTER(i32)]),
            ("fd2d_engine_done", [vp, i64, ctypes.POINTER(i32)]),
            ("fd2d_engine_destroy", [vp]),
            ("fd2d_run_batch_device", [ctypes.POINTER(CSolve), i32, vp]),
            ("fd2d_solve_device_bytes", [ctypes.POINTER(CSolve), ctypes.POINTER(i64)])):
        f = getattr(lib, name)
        f.argtypes = at
        f.restype = ctypes.c_int
    lib._batch_bound = True
    return lib


class SolveSpec:
    """One solve: the C struct plus every object its pointers refer to."""

    def __init__(self, solver: int, mode: float, config: Config, inputs: dict, outputs: Sequence):
        self.solver = solver
        self.config = config
        self.cc = config.to_c()
        self.tensors = [t for t in list(inputs.values()) + list(outputs) if t is not None]
        self._desc = {k: _c(v) for k, v in inputs.items()}
        self._outs = [_m(t) for t in outputs] + [None] * (4 - len(outputs))
        s = CSolve()
        s.solver = solver
        s.mode = float(mode)
        s.cfg = ctypes.cast(ctypes.pointer(self.cc), ctypes.c_void_p)
        for k, d in self._desc.items():
            setattr(s, k, d)
        for q, d in enumerate(self._outs):
            s.out[q] = d
        self.c = s
        self.outputs = list(outputs)

    def device_bytes(self) -> int:
        """Upper bound of the device memory the library allocates for this solve."""
        b = ctypes.c_int64(0)
        _check(_bind().fd2d_solve_device_bytes(ctypes.byref(self.c), ctypes.byref(b)))
        return int(b.value)


def green(mu, rho, src, rec, mode, dmu, G_in, *, config: Config, G_fft, G_out=None, seismograms=None) -> SolveSpec:
    """run_forward1_green (code/run_forward1_green.m:1) as a batch item."""
    return SolveSpec(abi.FD2D_GREEN, mode, config,
                     dict(mu=mu, rho=rho, src=src, rec=rec, dmu=dmu, born_in=G_in), [G_fft, G_out, seismograms])


def correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in, *, config: Config, seismograms=None,
                C_out=None) -> SolveSpec:
    """run_forward2_correlation (code/run_forward2_correlation.m:1) as a batch item."""
    return SolveSpec(abi.FD2D_CORRELATION, mode, config,
                     dict(mu=mu, rho=rho, G_fft=G_fft, spectrum=spectrum, source_dist=source_dist, rec=rec, dmu=dmu,
                          born_in=C_in), [seismograms, C_out])


def adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in, *, config: Config, K=None,
            stf_fft=None, u_out=None, seismograms=None) -> SolveSpec:
    """run_noise_adjoint (code/run_noise_adjoint.m:1) as a batch item."""
    return SolveSpec(abi.FD2D_ADJOINT, mode, config,
                     dict(mu=mu, rho=rho, u_fwd=u_fwd, stf=stf, src=src, rec=rec, spectrum=spectrum,
                          source_dist=source_dist, G_fft=G_fft, dmu=dmu, born_in=u_in), [K, stf_fft, u_out, seismograms])


def run_batch(specs: Sequence[SolveSpec]) -> None:
    """Runs every solve to completion on torch's current stream (one engine)."""
    arr = (CSolve * max(len(specs), 1))()
    for q, s in enumerate(specs):
        arr[q] = s.c
    _check(_bind().fd2d_run_batch_device(arr, len(specs), _stream()))


class Engine:
    """A batch engine on torch's current stream (fd2d_engine_* of include/fd2d_noise.h)."""

    def __init__(self):
        self._lib = _bind()
        self._h = ctypes.c_void_p()
        _check(self._lib.fd2d_engine_create(_stream(), ctypes.byref(self._h)))
        self._live = {}

    def submit(self, spec: SolveSpec) -> int:
        t = ctypes.c_int64(0)
        _check(self._lib.fd2d_engine_submit(self._h, ctypes.byref(spec.c), ctypes.byref(t)))
        self._live[int(t.value)] = spec
        return int(t.value)

    def advance(self, rounds: int = 1) -> int:
        n = ctypes.c_int32(0)
        _check(self._lib.fd2d_engine_advance(self._h, int(rounds), ctypes.byref(n)))
        for t in [t for t in self._live if self.done(t)]:
            del self._live[t]
        return int(n.value)

    def done(self, ticket: int) -> bool:
        d = ctypes.c_int32(0)
        _check(self._lib.fd2d_engine_done(self._h, int(ticket), ctypes.byref(d)))
        return bool(d.value)

    @property
    def active(self) -> int:
        return len(self._live)

    def run_until_done(self) -> None:
        while self.advance(1 << 20):
            pass

    def close(self) -> None:
        if self._h:
            _check(self._lib.fd2d_engine_destroy(self._h))
            self._h = ctypes.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

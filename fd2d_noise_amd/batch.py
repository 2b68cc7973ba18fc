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
from typing import Optional, Sequence

import numpy as np
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
            ("fd2d_engine_progress", [vp, i64, ctypes.POINTER(i32), ctypes.POINTER(i64), ctypes.POINTER(i64)]),
            ("fd2d_engine_cancel", [vp, i64, ctypes.POINTER(i32)]),
            ("fd2d_engine_destroy", [vp]),
            ("fd2d_run_batch_device", [ctypes.POINTER(CSolve), i32, vp]),
            ("fd2d_solve_device_bytes", [ctypes.POINTER(CSolve), ctypes.POINTER(i64)]),
            ("fd2d_release_cached_memory", [])):
        f = getattr(lib, name)
        f.argtypes = at
        f.restype = ctypes.c_int
    lib._batch_bound = True
    return lib


HOST_INPUTS = ("src", "rec", "spectrum")  # inputs the library also reads from host memory


def _c_host(a: np.ndarray):
    """Descriptor of a host array (MATLAB column-major float64)."""
    d = abi.CArray()
    d.data = a.ctypes.data
    d.dims[:] = abi.matlab_dims(a.shape)
    return ctypes.pointer(d)


def _desc(name, v):
    if isinstance(v, np.ndarray):
        if name not in HOST_INPUTS:
            raise TypeError(f"{name} must be a device tensor (host arrays are accepted for {HOST_INPUTS})")
        return _c_host(v)
    return _c(v)


class SolveSpec:
    """One solve: the C struct plus every object its pointers refer to. src, rec and spectrum may be
    host arrays (float64, MATLAB layout): the library then reads them without waiting for the
    stream."""

    def __init__(self, solver: int, mode: float, config: Config, inputs: dict, outputs: Sequence):
        self.solver = solver
        self.config = config
        self.cc = config.to_c()
        inputs = {k: (np.asfortranarray(v, dtype=np.float64) if isinstance(v, np.ndarray) else v)
                  for k, v in inputs.items()}
        self.tensors = [t for t in list(inputs.values()) + list(outputs) if t is not None]
        self._desc = {k: _desc(k, v) for k, v in inputs.items()}
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

    def run_plain(self) -> None:
        """Runs this solve alone through its own *_device entry point (device.py)."""
        from . import device as D
        fn = {abi.FD2D_GREEN: D.run_forward1_green, abi.FD2D_CORRELATION: D.run_forward2_correlation,
              abi.FD2D_ADJOINT: D.run_noise_adjoint}[self.solver]
        args = [D.to_device(a) if isinstance(a, np.ndarray) else a for a in self.call_args]
        fn(*args, config=self.config, **self.call_outputs)

    def time_steps(self) -> int:
        """Time steps of this solve: nt (Green), 2nt-1 (correlation, adjoint) or nt for the
        half-length adjoint modes 0 / 1 with a distributed source (run_noise_adjoint.m:165-167)."""
        c = self.config
        if self.solver == abi.FD2D_GREEN:
            return c.nt
        if self.solver == abi.FD2D_ADJOINT and int(self.c.mode) in (0, 1):
            stf = self._desc.get("stf")
            if stf is not None and stf.contents.dims[2] != 1:
                return c.nt
        return c.n_corr

    def device_bytes(self) -> int:
        """Upper bound of the device memory the library allocates for this solve."""
        b = ctypes.c_int64(0)
        _check(_bind().fd2d_solve_device_bytes(ctypes.byref(self.c), ctypes.byref(b)))
        return int(b.value)


def _with_call(spec: SolveSpec, args, outputs: dict) -> SolveSpec:
    spec.call_args, spec.call_outputs = args, outputs
    return spec


def green(mu, rho, src, rec, mode, dmu, G_in, *, config: Config, G_fft, G_out=None, seismograms=None) -> SolveSpec:
    """run_forward1_green (code/run_forward1_green.m:1) as a batch item."""
    return _with_call(SolveSpec(abi.FD2D_GREEN, mode, config,
                                dict(mu=mu, rho=rho, src=src, rec=rec, dmu=dmu, born_in=G_in),
                                [G_fft, G_out, seismograms]),
                      (mu, rho, src, rec, mode, dmu, G_in), dict(G_fft=G_fft, G_out=G_out, seismograms=seismograms))


def correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in, *, config: Config, seismograms=None,
                C_out=None) -> SolveSpec:
    """run_forward2_correlation (code/run_forward2_correlation.m:1) as a batch item."""
    return _with_call(SolveSpec(abi.FD2D_CORRELATION, mode, config,
                                dict(mu=mu, rho=rho, G_fft=G_fft, spectrum=spectrum, source_dist=source_dist, rec=rec,
                                     dmu=dmu, born_in=C_in), [seismograms, C_out]),
                      (mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in),
                      dict(seismograms=seismograms, C_out=C_out))


def adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in, *, config: Config, K=None,
            stf_fft=None, u_out=None, seismograms=None) -> SolveSpec:
    """run_noise_adjoint (code/run_noise_adjoint.m:1) as a batch item."""
    return _with_call(SolveSpec(abi.FD2D_ADJOINT, mode, config,
                                dict(mu=mu, rho=rho, u_fwd=u_fwd, stf=stf, src=src, rec=rec, spectrum=spectrum,
                                     source_dist=source_dist, G_fft=G_fft, dmu=dmu, born_in=u_in),
                                [K, stf_fft, u_out, seismograms]),
                      (mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in),
                      dict(K=K, stf_fft=stf_fft, u_out=u_out, seismograms=seismograms))


def solve_bytes(solver: int, config: Config, mode: float = 0.0, n_src: int = 0, n_rec: int = 0,
                G_fft: bool = False, spectrum_cols: int = 1, source_dist: bool = False, stf_dims=None,
                dmu: bool = False, fft_out: bool = False) -> int:
    """Device bytes the library allocates for a solve of this shape (fd2d_solve_device_bytes on
    size-only descriptors): the figure an admission policy charges before the arrays exist."""
    c = config
    keep = []

    def desc(dims):
        d = abi.CArray()
        d.data = 1  # marks the array as given; only its sizes are read
        d.dims[:] = tuple(int(x) for x in dims)
        keep.append(d)
        return ctypes.pointer(d)

    s = CSolve()
    s.solver, s.mode = solver, float(mode)
    cc = c.to_c()
    s.cfg = ctypes.cast(ctypes.pointer(cc), ctypes.c_void_p)
    if n_src:
        s.src = desc((n_src, 2, 1))
    if n_rec:
        s.rec = desc((n_rec, 2, 1))
    if G_fft:
        s.G_fft = desc((c.nx, c.nz, c.n_sample))
    if spectrum_cols:
        s.spectrum = desc((c.n_sample, spectrum_cols, 1))
    if source_dist:
        s.source_dist = desc((c.nx, c.nz, max(c.n_basis_fct, 1)))
    if stf_dims is not None:
        s.stf = desc(stf_dims)
    if dmu:
        s.dmu = desc((c.nx, c.nz, 1))
    m = abi.MArray()
    if fft_out or solver == abi.FD2D_GREEN:
        m.data = 1
        m.dims[:] = (c.nx, c.nz, c.n_sample)
        keep.append(m)
        s.out[1 if solver == abi.FD2D_ADJOINT else 0] = ctypes.pointer(m)
    b = ctypes.c_int64(0)
    _check(_bind().fd2d_solve_device_bytes(ctypes.byref(s), ctypes.byref(b)))
    return int(b.value)


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
        """The solve no longer runs and its queued work has completed (complete, cancelled or failed)."""
        d = ctypes.c_int32(0)
        _check(self._lib.fd2d_engine_done(self._h, int(ticket), ctypes.byref(d)))
        return bool(d.value)

    def progress(self, ticket: int):
        """(state, steps issued, steps total) of a solve; the step counts are host-side and -1 once
        the solve is complete. state is one of abi.FD2D_SOLVE_*."""
        st, a, b = ctypes.c_int32(0), ctypes.c_int64(0), ctypes.c_int64(0)
        _check(self._lib.fd2d_engine_progress(self._h, int(ticket), ctypes.byref(st), ctypes.byref(a),
                                              ctypes.byref(b)))
        return int(st.value), int(a.value), int(b.value)

    def state(self, ticket: int) -> int:
        return self.progress(ticket)[0]

    def cancel(self, ticket: int) -> bool:
        """Stops a running solve (its outputs are left undefined); False if it had finished."""
        c = ctypes.c_int32(0)
        _check(self._lib.fd2d_engine_cancel(self._h, int(ticket), ctypes.byref(c)))
        self._live.pop(int(ticket), None)
        return bool(c.value)

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


def available_bytes() -> int:
    """Device bytes a batch can plan with: the free device memory after the library has returned
    its cached blocks, plus what torch's caching allocator holds but does not use."""
    _check(_bind().fd2d_release_cached_memory())
    free, _ = torch.cuda.mem_get_info()
    return int(free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated())


def run_programs(programs: Sequence, rounds: int = 64, costs: Optional[Sequence[int]] = None,
                 budget: Optional[int] = None) -> list:
    """Runs solve programs together through one engine and returns their results in order. A
    program is a generator that yields the solves of its next stage (a list of SolveSpec), is
    resumed once all of them are complete, and returns its result; the stages of all programs share
    the engine's grouped launches. With costs (peak device bytes of each program) and budget, a
    program starts only when its cost fits beside the programs running (the first always starts),
    in order, as earlier ones finish; a generator allocates nothing before it starts."""
    progs = list(programs)
    results = [None] * len(progs)
    pending = {}
    running = {}  # started programs -> cost
    nxt = 0
    with Engine() as eng:
        def resume(q, exc=None):
            try:
                specs = progs[q].throw(exc) if exc is not None else next(progs[q])
            except StopIteration as fin:
                results[q] = fin.value
                pending.pop(q, None)
                running.pop(q, None)
                progs[q] = None
                return
            pending[q] = (specs, [eng.submit(s) for s in specs])

        def admit():
            nonlocal nxt
            while nxt < len(progs):
                cost = 0 if costs is None else int(costs[nxt])
                if budget is not None and running and sum(running.values()) + cost > budget:
                    return
                running[nxt] = cost
                nxt += 1
                resume(nxt - 1)
        admit()
        while pending:
            eng.advance(rounds)
            for q in [q for q, (_, ts) in pending.items() if all(eng.done(t) for t in ts)]:
                # the finished stage's specs are released before the program runs on, so that the
                # arrays it drops are free for its next stage
                _, ts = pending.pop(q)
                bad = [k for k, t in enumerate(ts) if eng.state(t) == abi.FD2D_SOLVE_FAILED]
                # a failed solve is raised inside its program (which may handle it); the other
                # programs are not affected
                resume(q, abi.Fd2dError(abi.FD2D_ENONFINITE, f"solve {bad[0]} of the stage became NaN or Inf")
                       if bad else None)
            admit()
    torch.cuda.synchronize()
    return results

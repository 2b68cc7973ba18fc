"""Serving loop: correlation, kernel and Hessian-vector requests with continuous batching.

The reference's L-BFGS driver asks for per-station correlations, gradients (kernels) and
Hessian-vector products one call at a time (inversion/optlib_callback_fct/
eval_objective_and_gradient.m:54-120, eval_hessian_vector_product_2.m:44-183). Here each request is
a small program over solves: a generator that yields the solves of its next stage (which may run
together), does its FP64 caller work on the device between stages (adjoint sources, sums) and
returns its result. The server

  * admits requests against an HBM budget that counts what each request holds at its peak: its
    wavefield and spectrum arrays (stored forward snapshots included) and the library's internal
    buffers of its solves (fd2d_solve_device_bytes). A request that cannot fit even alone, or a
    Hessian-vector product on a model that does not store every step, is rejected at submission;
    one that fits but not now waits, and later requests that fit may start before it (backfill)
    a bounded number of times;
  * keeps models (mu, rho) and Green-function spectra resident (the spectra in an LRU cache,
    counted in the budget), so a correlation request for a station whose G_fft is cached skips the
    Green run, as the reference reuses G_fft_<i>.mat (eval_objective_and_gradient.m:44-53);
  * submits every stage to one batch engine, so solves of different requests share launches, and
    requests that arrive while others run join at the next round (continuous batching);
  * records queueing + service latency per request (p50 / p99 from ``latencies()``).

``run_alone`` runs the same request programs one solve at a time through the single-solve entry
points; the results of both paths are bit-identical.
"""
from __future__ import annotations

import collections
import dataclasses
import time
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import abi, batch as B, device as D
from .config import Config
from .measurement import make_adjoint_sources_device

f32, f64, cd = torch.float32, torch.float64, torch.complex128


@dataclasses.dataclass
class Model:
    """A resident model: mu, rho as device tensors of one grid, plus the noise source."""
    config: Config
    mu: torch.Tensor
    rho: torch.Tensor
    spectrum: torch.Tensor
    source_dist: torch.Tensor
    spectrum_host: Optional[np.ndarray] = None  # the same spectrum in host memory (read without a wait)
    t: Optional[torch.Tensor] = None             # the correlation time axis, resident


@dataclasses.dataclass(eq=False)
class Request:
    """kind: 'correlation' (Green + correlation), 'kernel' (correlation + structure kernel:
    eval_objective_and_gradient.m:52-98 for one station) or 'hvp' (the structure Hessian-vector
    product of eval_hessian_vector_product_2.m for one station). src is the reference station
    (1 x 2, metres), rec the receivers (n x 2); observed correlations (n x N, MATLAB layout as an
    (N, n) device tensor) for 'kernel' and 'hvp'; dmu (nx x nz device tensor) for 'hvp'."""
    kind: str
    model: str
    src: np.ndarray
    rec: np.ndarray
    observed: Optional[torch.Tensor] = None
    dmu: Optional[torch.Tensor] = None
    measurement: str = "waveform_difference"
    veldis: str = "dis"
    window: tuple = (100.0, 200.0)
    tag: Any = None


# ---------------------------------------------------------------------------------------------
# Request programs
# ---------------------------------------------------------------------------------------------
def _host(a) -> np.ndarray:
    return np.asfortranarray(np.atleast_2d(np.asarray(a, dtype=np.float64)))


def _adjoint_stf(corr, observed, du, m: "Model", req: Request, order: str):
    _, adstf = make_adjoint_sources_device(corr, observed, du, m.t, req.veldis, req.measurement, order, req.window,
                                           t_host=m.config.t_corr())
    stf = torch.zeros(adstf.shape, dtype=cd, device=adstf.device)
    torch.view_as_real(stf)[..., 0].copy_(adstf)
    return stf


def correlation_program(srv: "Server", req: Request):
    m = srv.models[req.model]
    c = m.config
    src, rec = _host(req.src), _host(req.rec)
    key = (req.model, tuple(src.ravel()))
    G = srv.cached_spectrum(key)
    if G is None:
        G = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
        yield [B.green(m.mu, m.rho, src, None, 0, None, None, config=c, G_fft=G)]
        srv.store_spectrum(key, G)
    corr = D.empty_like_dims((req.rec.shape[0], c.n_corr), f64)
    yield [B.correlation(m.mu, m.rho, G, m.spectrum_host, m.source_dist, rec, 0, None, None, config=c,
                         seismograms=corr)]
    return {"correlations": corr}


def kernel_program(srv: "Server", req: Request):
    """Green (snapshots), correlation (snapshots), misfit and adjoint sources, adjoint mode 1
    against C, adjoint mode 0 against G with the transformed adjoint state; K = K1 + K2."""
    m = srv.models[req.model]
    c = m.config
    src, rec = _host(req.src), _host(req.rec)
    nr = req.rec.shape[0]
    G = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    Gout = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32)
    yield [B.green(m.mu, m.rho, src, None, 1, None, None, config=c, G_fft=G, G_out=Gout)]
    srv.store_spectrum((req.model, tuple(src.ravel())), G)
    corr = D.empty_like_dims((nr, c.n_corr), f64)
    C = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
    yield [B.correlation(m.mu, m.rho, G, m.spectrum_host, m.source_dist, rec, 1, None, None, config=c,
                         seismograms=corr, C_out=C)]
    stf = _adjoint_stf(corr, req.observed, None, m, req, "1st")
    nbk = max(c.n_basis_fct, 1) + 1
    K1 = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    A1 = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    yield [B.adjoint(m.mu, m.rho, C, stf, rec, None, None, None, None, 1, None, None, config=c, K=K1, stf_fft=A1)]
    del C
    K2 = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    yield [B.adjoint(m.mu, m.rho, Gout, A1, rec, None, m.spectrum_host, m.source_dist, None, 0, None, None,
                     config=c, K=K2)]
    return {"correlations": corr, "K": K1 + K2}


def hvp_program(srv: "Server", req: Request):
    """Structure Hessian-vector product of one station (eval_hessian_vector_product_2.m:57-170 with
    type 'structure' and dm on mu only): Green (mode 1) and Born Green (mode 5), correlation and
    Born correlation, 1st- and 2nd-order adjoint sources, then YELLOW (13), GRAY (11) and BLACK (12,
    Born with u_dag) and RED (10, Born with p). Needs fw_nth = 1."""
    m = srv.models[req.model]
    c = m.config
    if c.fw_nth != 1:
        raise ValueError("the Hessian-vector product needs fw_nth = 1 (eval_hessian_vector_product_2.m:75,154)")
    src, rec = _host(req.src), _host(req.rec)
    nr, N = req.rec.shape[0], c.n_corr
    mu, rho, spec, sd, dmu = m.mu, m.rho, m.spectrum_host, m.source_dist, req.dmu
    nbk = max(c.n_basis_fct, 1) + 1
    G = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    Gout = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32)
    yield [B.green(mu, rho, src, None, 1, None, None, config=c, G_fft=G, G_out=Gout)]  # :57
    dG = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    dGout = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32)
    c_ref = D.empty_like_dims((nr, N), f64)
    C = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
    yield [B.green(mu, rho, src, None, 5, dmu, Gout, config=c, G_fft=dG, G_out=dGout),  # :61
           B.correlation(mu, rho, G, spec, sd, rec, 1, None, None, config=c, seismograms=c_ref, C_out=C)]  # :71
    c_pert = D.empty_like_dims((nr, N), f64)
    dC = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
    yield [B.correlation(mu, rho, dG, spec, sd, rec, 1, dmu, C, config=c, seismograms=c_pert, C_out=dC)]  # :75
    stf1 = _adjoint_stf(c_ref, req.observed, c_pert, m, req, "1st")  # :99-105
    stf2 = _adjoint_stf(c_ref, req.observed, c_pert, m, req, "2nd")
    H = torch.zeros((nbk, c.nz, c.nx), dtype=f64, device="cuda")
    KY = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    uT = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    u_dag = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
    yield [B.adjoint(mu, rho, dC, stf1, rec, None, spec, None, dG, 13, None, None, config=c, K=KY, stf_fft=uT,
                     u_out=u_dag)]  # YELLOW :113-121
    KG = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    p = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
    KB = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    kT = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
    yield [B.adjoint(mu, rho, dGout, uT, rec, None, spec, sd, None, 11, None, None, config=c, K=KG, u_out=p),  # GRAY
           B.adjoint(mu, rho, C, stf2, rec, None, spec, None, G, 12, dmu, u_dag, config=c, K=KB, stf_fft=kT)]  # BLACK
    KR = D.empty_like_dims((c.nx, c.nz, nbk), f64)
    yield [B.adjoint(mu, rho, Gout, kT, rec, None, spec, sd, None, 10, dmu, p, config=c, K=KR)]  # RED :168-170
    # H = sum of the kernels; with type 'structure' only the structure map of YELLOW and BLACK
    # (:120, :157-159); GRAY and RED enter whole
    for K, part in ((KY, "structure"), (KG, None), (KB, "structure"), (KR, None)):
        if part == "structure":
            K[:nbk - 1].zero_()
        H += K
    return {"correlations": c_ref, "H": H}


PROGRAMS: Dict[str, Callable] = {"correlation": correlation_program, "kernel": kernel_program, "hvp": hvp_program}


def request_bytes(srv: "Server", req: Request) -> int:
    """Peak device bytes of a request: its arrays plus the library buffers of its largest stage."""
    return request_bytes_for(req.kind, srv.models[req.model].config, req.rec.shape[0])


def request_bytes_for(kind: str, c: Config, nr: int) -> int:
    """Peak device bytes of a request of this kind on a model with config c and nr receivers."""
    P, K, N = c.nx * c.nz, c.n_sample, c.n_corr
    spec_b = 16 * P * K
    g_snap, c_snap = 4 * P * c.n_fw_green(), 4 * P * c.n_fw_corr()
    green = B.solve_bytes(abi.FD2D_GREEN, c, 1, 1, 0)
    corr = B.solve_bytes(abi.FD2D_CORRELATION, c, 1, 0, nr, True, 1, True)
    adj1 = B.solve_bytes(abi.FD2D_ADJOINT, c, 1, nr, nr, False, 0, False, (nr, N, 1), fft_out=True)
    adj0 = B.solve_bytes(abi.FD2D_ADJOINT, c, 0, nr, nr, False, 1, True, (c.nx, c.nz, K))
    traces = 8 * nr * N * 4
    if kind == "correlation":
        return spec_b + max(green, corr) + traces
    if kind == "kernel":
        return 2 * spec_b + g_snap + c_snap + 16 * P * 2 + max(green, corr, adj1, adj0) + traces
    # hvp: G, dG, uT, kT spectra; G, dG snapshots; C, dC, u_dag, p snapshots; 4 kernels; two
    # solves of a stage at once
    return 4 * spec_b + 2 * g_snap + 4 * c_snap + 4 * 16 * P + 2 * max(green, corr, adj1, adj0) + traces


def kernel_fw_nth(c: Config, nr: int, budget: int) -> int:
    """The smallest snapshot interval fw_nth (at least c.fw_nth) at which a kernel request on this
    model fits in `budget` bytes: the stored forward wavefields (4 B per grid point per stored step
    of the Green and the correlation run) dominate a kernel request, and the reference's own knob
    for them is fw_nth (input_parameters.m). Raises ValueError when no interval fits."""
    f = max(1, int(c.fw_nth))
    while f <= c.nt:
        if request_bytes_for("kernel", c.replace(fw_nth=f), nr) <= budget:
            return f
        f += max(1, f // 8)
    raise ValueError(f"no fw_nth lets a kernel request on a {c.nx} x {c.nz} grid fit in {budget / 1e9:.1f} GB")


class RequestRejected(Exception):
    """A request the server cannot run: it does not fit the budget even alone, or its model cannot
    serve its kind (a Hessian-vector product needs fw_nth = 1)."""


@dataclasses.dataclass(eq=False)
class _Live:
    req: Request
    gen: Any
    bytes: int
    t_arrival: float
    bypassed: int = 0  # later requests admitted while this one waited at the head of the queue
    t_admit: float = 0.0
    tickets: List[int] = dataclasses.field(default_factory=list)
    specs: List[Any] = dataclasses.field(default_factory=list)
    pending: List[Any] = dataclasses.field(default_factory=list)  # solves of the stage not yet submitted
    t_done: float = 0.0
    result: Any = None
    error: Optional[BaseException] = None  # set instead of result when the request failed or was cancelled


class Server:
    """The serving loop on torch's current stream. ``budget_bytes`` defaults to 80% of the device
    memory available at start; ``rounds`` is the number of engine rounds between completion checks.

    Admission: a request whose peak bytes exceed the budget, or that its model cannot serve, is
    rejected at submission (it ends at once with RequestRejected and is listed in ``rejected``). A
    request that fits the budget but not beside the running ones waits (deferred). Requests start
    in arrival order, except that a later request that fits now may start while the head of the
    queue waits (backfill), at most ``max_bypass`` times per head; after that nothing starts
    before the head, so a large request is not starved by a stream of small ones."""

    def __init__(self, budget_bytes: Optional[int] = None, rounds: int = 32, spectrum_cache_bytes: Optional[int] = None,
                 max_bypass: int = 8):
        # default: 80% of the device memory available at start (free, plus what torch's caching
        # allocator holds unused, after the library has returned its cached blocks)
        self.budget = int(budget_bytes if budget_bytes is not None else 0.8 * B.available_bytes())
        self.max_bypass = int(max_bypass)
        self.rejected: List[_Live] = []
        self.cache_cap = int(spectrum_cache_bytes if spectrum_cache_bytes is not None else self.budget // 4)
        self.rounds = rounds
        self.models: Dict[str, Model] = {}
        self.engine = B.Engine()
        self.queue: collections.deque = collections.deque()
        self.live: List[_Live] = []
        self.done: List[_Live] = []
        self.in_use = 0
        self.gcache: "collections.OrderedDict[Any, torch.Tensor]" = collections.OrderedDict()
        self.gcache_bytes = 0
        self.cache_hits = 0
        self.gp_updates = 0  # grid points x time steps of every solve submitted

    # ---- resident state ------------------------------------------------------------------------
    def add_model(self, name: str, config: Config, mu, rho, spectrum, source_dist) -> None:
        dev = lambda a: a if torch.is_tensor(a) else D.to_device(a)
        spec_h = D.to_host(spectrum) if torch.is_tensor(spectrum) else np.asarray(spectrum, dtype=np.float64)
        self.models[name] = Model(config, dev(mu), dev(rho), dev(spectrum), dev(source_dist),
                                  np.asfortranarray(spec_h, dtype=np.float64),
                                  torch.as_tensor(config.t_corr(), device="cuda"))

    def cached_spectrum(self, key):
        G = self.gcache.get(key)
        if G is not None:
            self.gcache.move_to_end(key)
            self.cache_hits += 1
        return G

    def store_spectrum(self, key, G) -> None:
        if key in self.gcache:
            return
        b = G.numel() * G.element_size()
        if b > self.cache_cap:
            return
        # the cache stays within its cap and, with the admitted requests, within the budget
        while self.gcache and (self.gcache_bytes + b > self.cache_cap or
                               self.in_use + self.gcache_bytes + b > self.budget):
            _, old = self.gcache.popitem(last=False)
            self.gcache_bytes -= old.numel() * old.element_size()
        if self.in_use + self.gcache_bytes + b > self.budget:
            return
        self.gcache[key] = G
        self.gcache_bytes += b

    # ---- requests --------------------------------------------------------------------------------
    def submit(self, req: Request, t_arrival: Optional[float] = None) -> _Live:
        if req.kind not in PROGRAMS:
            raise ValueError(f"unknown request kind {req.kind!r}")
        if req.model not in self.models:
            raise KeyError(f"model {req.model!r} is not resident")
        live = _Live(req, None, request_bytes(self, req), time.perf_counter() if t_arrival is None else t_arrival)
        c = self.models[req.model].config
        reason = None
        if req.kind == "hvp" and c.fw_nth != 1:
            reason = (f"a Hessian-vector product needs fw_nth = 1 (eval_hessian_vector_product_2.m:75,154); model "
                      f"{req.model!r} stores every {c.fw_nth}th step")
        elif live.bytes > self.budget:
            reason = f"the request needs {live.bytes / 1e9:.1f} GB at its peak, more than the budget {self.budget / 1e9:.1f} GB"
        if reason is not None:
            live.error, live.t_done = RequestRejected(reason), live.t_arrival
            self.rejected.append(live)
            return live
        self.queue.append(live)
        return live

    def _room_for(self, b: int) -> bool:
        """Whether b more bytes fit beside the running requests, dropping cached spectra (oldest
        first) to make room."""
        while self.in_use + self.gcache_bytes + b > self.budget and self.gcache:
            _, old = self.gcache.popitem(last=False)
            self.gcache_bytes -= old.numel() * old.element_size()
        return self.in_use + self.gcache_bytes + b <= self.budget

    def _admit(self) -> None:
        q = 0
        while q < len(self.queue):
            head, live = self.queue[0], self.queue[q]
            if q > 0 and head.bypassed >= self.max_bypass:
                return  # the head has waited long enough: nothing more starts before it
            if not self._room_for(live.bytes):
                q += 1  # deferred; a later request may fit
                continue
            del self.queue[q]
            if q > 0:
                head.bypassed += 1
            self.in_use += live.bytes
            live.t_admit = time.perf_counter()
            live.gen = PROGRAMS[live.req.kind](self, live.req)
            self.live.append(live)
            self._resume(live)

    def _finish(self, live: _Live, result=None, error: Optional[BaseException] = None) -> None:
        live.result, live.error = result, error
        live.t_done = time.perf_counter()
        self.in_use -= live.bytes
        self.live.remove(live)
        self.done.append(live)
        if live.gen is not None and error is not None:
            live.gen.close()
        live.gen = None

    def _resume(self, live: _Live) -> None:
        """Runs the request program to its next stage and submits that stage's solves."""
        try:
            specs = next(live.gen)
        except StopIteration as fin:
            live.gen = None
            self._finish(live, fin.value)
            return
        except Exception as e:  # one request's failure must not stop the others
            # the request's own caller work rejected its data, its arrays for the next stage do not
            # fit the device memory free now, or its program failed: the request ends with the
            # error, the others go on
            self._finish(live, error=e)
            return
        live.specs, live.tickets, live.pending = specs, [], list(specs)
        self.gp_updates += sum(s.config.nx * s.config.nz * s.time_steps() for s in specs)
        self._submit_pending(live)

    def _submit_pending(self, live: _Live) -> bool:
        """Submits the solves of the request's stage that are not yet running. A solve whose library
        buffers do not fit the device memory free now (FD2D_ENOMEM: the budget counts what requests
        hold, but torch's caching allocator may keep the freed arrays of finished requests, split
        into blocks of other sizes) is retried once after torch returns its cached memory; if it still
        does not fit, it waits for the next step, when running requests may have freed theirs. A solve
        the engine refuses for another reason (its arrays do not match its configuration) ends the
        request with that error. True when nothing of the stage is left to submit."""
        emptied = False
        while live.pending:
            try:
                live.tickets.append(self.engine.submit(live.pending[0]))
            except abi.Fd2dError as e:
                if e.code != abi.FD2D_ENOMEM:
                    self._abandon(live, e)
                    return True
                if emptied:
                    return False
                torch.cuda.empty_cache()  # the library has returned its own cached blocks already
                emptied = True
                continue
            live.pending.pop(0)
        return True

    def _abandon(self, live: _Live, error: BaseException) -> None:
        """Stops the request's running solves and ends it with `error`; other requests go on."""
        for t in live.tickets:
            self.engine.cancel(t)
        live.tickets, live.pending = [], []
        self._finish(live, error=error)

    def cancel(self, live: _Live) -> bool:
        """Cancels a queued or running request (it ends with an error); False if it had finished."""
        if live in self.queue:
            self.queue.remove(live)
            live.error, live.t_done = RuntimeError("cancelled"), time.perf_counter()
            self.done.append(live)
            return True
        if live in self.live:
            self._abandon(live, RuntimeError("cancelled"))
            return True
        return False

    def step(self) -> None:
        """Admits what fits, advances the engine, resumes requests whose stage has completed and
        ends those with a failed (non-finite) solve."""
        self._admit()
        if not self.live:
            return
        waiting = [live for live in self.live if live.pending and not self._submit_pending(live)]
        if waiting and len(waiting) == len(self.live) and self.engine.active == 0:
            # nothing runs that could free memory: the first waiting request cannot run now
            self._abandon(waiting[0], abi.Fd2dError(abi.FD2D_ENOMEM, "the device memory its solves need is not free"))
            return
        self.engine.advance(self.rounds)
        for live in list(self.live):
            if live.pending:
                continue  # its stage waits for device memory
            states = [self.engine.state(t) for t in live.tickets]
            if abi.FD2D_SOLVE_FAILED in states:
                q = states.index(abi.FD2D_SOLVE_FAILED)
                self._abandon(live, abi.Fd2dError(
                    abi.FD2D_ENONFINITE, f"solve {q} of the request's stage became NaN or Inf (time step beyond "
                    "the stability limit of the stencil?)"))
            elif all(self.engine.done(t) for t in live.tickets):
                live.tickets, live.specs = [], []
                self._resume(live)

    def busy(self) -> bool:
        return bool(self.queue or self.live)

    def run_until_idle(self) -> None:
        while self.busy():
            self.step()
        torch.cuda.synchronize()

    def serve(self, arrivals: List[tuple]) -> None:
        """arrivals: (offset seconds, Request) in time order; each request is submitted once the
        wall clock passes its offset, while the loop keeps stepping the running ones."""
        t0 = time.perf_counter()
        pending = collections.deque(sorted(arrivals, key=lambda a: a[0]))
        while pending or self.busy():
            now = time.perf_counter() - t0
            while pending and pending[0][0] <= now:
                off, req = pending.popleft()
                self.submit(req, t0 + off)
            if self.busy():
                self.step()
            elif pending:
                time.sleep(min(0.002, max(0.0, pending[0][0] - now)))
        torch.cuda.synchronize()

    def latencies(self) -> np.ndarray:
        """Arrival to result of every finished request (rejected ones are not in done)."""
        return np.array([d.t_done - d.t_arrival for d in self.done])

    def waits(self) -> np.ndarray:
        """Arrival to admission of every finished request: the time it was deferred."""
        return np.array([max(0.0, d.t_admit - d.t_arrival) for d in self.done if d.t_admit])

    def close(self) -> None:
        self.engine.close()


class _Alone:
    """The resident models of a server without its spectrum cache."""

    def __init__(self, srv: Server):
        self.models = srv.models

    def cached_spectrum(self, key):
        return None

    def store_spectrum(self, key, G) -> None:
        pass


def run_alone(srv: Server, req: Request):
    """The request program with every solve run alone through its single-solve entry point, in
    order and without the spectrum cache (the plain path the server is compared with)."""
    gen = PROGRAMS[req.kind](_Alone(srv), req)
    try:
        specs = next(gen)
        while True:
            for s in specs:
                s.run_plain()
            specs = gen.send(None)
    except StopIteration as fin:
        return fin.value

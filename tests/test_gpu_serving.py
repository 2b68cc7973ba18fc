"""The serving loop: every request's result is the same alone and in the mix, bit for bit.

A stream of correlation, kernel and Hessian-vector requests on two resident models of different
grid size and FD order arrives while earlier requests run; the server batches their solves through
one engine, reuses cached Green-function spectra and admits by HBM budget. Each result must equal
the request run alone with every solve through its single-solve entry point (serving.run_alone).
The solves themselves are held to the oracle by test_gpu_parity.py and the HVP composition by
test_gpu_caller.py.
"""
import numpy as np
import pytest

from fd2d_noise_amd import Config

from helpers import blob_distribution, hetero_model, interior_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _server(torch, budget=None):
    from fd2d_noise_amd import device as D, serving as S
    srv = S.Server(budget_bytes=budget, rounds=16)
    rng = np.random.default_rng(42)
    cfgs = {"a": Config(order=4, nx=60, nz=48, nt=120, fw_nth=1, freq_samp=3),
            "b": Config(order=8, nx=72, nz=70, nt=100, fw_nth=1, Lx=71e3, Lz=69e3)}
    for name, c in cfgs.items():
        mu, rho = hetero_model(rng, c)
        spec = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
        srv.add_model(name, c, mu, rho, spec, blob_distribution(rng, c))
    reqs = []
    for q, (kind, name) in enumerate([("correlation", "a"), ("kernel", "b"), ("hvp", "a"), ("correlation", "b"),
                                      ("correlation", "a"), ("kernel", "a"), ("hvp", "b"), ("correlation", "b")]):
        c = cfgs[name]
        src = interior_points(rng, c, 1, 0.4) if q not in (4, 7) else reqs[q - 4].src  # repeats: cache hits
        rec = interior_points(rng, c, 3)
        obs = torch.as_tensor(0.05 * rng.standard_normal((c.n_corr, 3)), device="cuda")
        dmu = D.to_device(0.02 * srv.models[name].mu.cpu().numpy().T * np.sign(rng.standard_normal((c.nx, c.nz))))
        win = (0.2 * c.t_corr()[-1], 0.6 * c.t_corr()[-1])
        reqs.append(S.Request(kind, name, src, rec, observed=obs, dmu=dmu if kind == "hvp" else None, window=win,
                              tag=q))
    return srv, reqs


def _check(torch, srv, reqs, lives):
    from fd2d_noise_amd import serving as S
    for live in lives:
        alone = S.run_alone(srv, live.req)
        assert live.result is not None and set(alone) == set(live.result)
        for k in alone:
            x, y = alone[k], live.result[k]
            assert torch.isfinite(x).all(), f"request {live.req.tag} {k} not finite"
            same = torch.equal(x, y)
            print(f"request {live.req.tag} {live.req.kind} {k} {tuple(x.shape)}: "
                  f"{'identical' if same else 'DIFFERENT'}", flush=True)
            assert same
            if k in ("K", "H"):
                assert float(x.abs().max()) > 0.0


def test_serving_mix_matches_alone(torch_mod):
    torch = torch_mod
    srv, reqs = _server(torch)
    arrivals = [(0.02 * q, r) for q, r in enumerate(reqs)]
    srv.serve(arrivals)
    lat = srv.latencies()
    print(f"{len(lat)} requests, p50 {np.percentile(lat, 50) * 1e3:.1f} ms, p99 {np.percentile(lat, 99) * 1e3:.1f} ms, "
          f"spectrum cache hits {srv.cache_hits}", flush=True)
    assert len(srv.done) == len(reqs) and srv.in_use == 0
    _check(torch, srv, reqs, srv.done)
    srv.close()


def test_serving_budget_defers_admission(torch_mod):
    torch = torch_mod
    from fd2d_noise_amd import serving as S
    srv, reqs = _server(torch)
    srv.close()
    need = max(S.request_bytes(srv, r) for r in reqs[:4])
    srv2, reqs2 = _server(torch, budget=int(need * 1.01))
    for r in reqs2[:4]:
        srv2.submit(r)
    deferred = False
    while srv2.busy():
        srv2.step()
        deferred |= bool(srv2.queue) and bool(srv2.live)
        assert srv2.in_use + srv2.gcache_bytes <= srv2.budget
    torch.cuda.synchronize()
    assert deferred and len(srv2.done) == 4
    _check(torch, srv2, reqs2[:4], srv2.done)
    with pytest.raises(KeyError):
        srv2.submit(S.Request("correlation", "zzz", reqs2[0].src, reqs2[0].rec))
    srv2.close()


def test_serving_rejects_infeasible_requests(torch_mod):
    """A request that cannot fit the budget even alone, and a Hessian-vector product on a model
    that does not store every step, are rejected at submission with the reason; the others run."""
    torch = torch_mod
    from fd2d_noise_amd import serving as S
    srv, reqs = _server(torch)
    m = srv.models["a"]
    srv.add_model("a4", m.config.replace(fw_nth=4), m.mu, m.rho, m.spectrum, m.source_dist)
    hv = reqs[2]
    assert hv.kind == "hvp" and hv.model == "a"
    bad = S.Request("hvp", "a4", hv.src, hv.rec, observed=hv.observed, dmu=hv.dmu, window=hv.window, tag="a4")
    r1 = srv.submit(bad)
    assert isinstance(r1.error, S.RequestRejected) and "fw_nth = 1" in str(r1.error)
    small = S.Server(budget_bytes=1024)
    small.models = srv.models
    r2 = small.submit(reqs[0])
    assert isinstance(r2.error, S.RequestRejected) and "budget" in str(r2.error)
    assert small.rejected == [r2] and not small.busy()
    small.close()
    srv.serve([(0.0, reqs[0]), (0.0, S.Request("kernel", "a4", reqs[5].src, reqs[5].rec, observed=reqs[5].observed,
                                                window=reqs[5].window, tag="k4"))])
    assert srv.rejected == [r1] and len(srv.done) == 2 and all(d.error is None for d in srv.done)
    # a kernel on the model that stores every 4th step is the reference's kernel at fw_nth = 4
    _check(torch, srv, reqs, srv.done)
    srv.close()


@pytest.mark.parametrize("max_bypass", [8, 0])
def test_serving_backfill(torch_mod, max_bypass):
    """A kernel request that does not fit beside a running one waits; a correlation request that
    arrives after it and fits starts first (backfill), unless max_bypass = 0."""
    torch = torch_mod
    from fd2d_noise_amd import serving as S
    srv, reqs = _server(torch)
    ka, kb, ca = reqs[5], reqs[1], reqs[0]
    assert (ka.kind, kb.kind, ca.kind) == ("kernel", "kernel", "correlation")
    b = [S.request_bytes(srv, r) for r in (ka, kb, ca)]
    srv.close()
    assert b[2] < b[1]
    budget = max(b[1], b[0] + b[2])  # room for ka and ca, or for kb alone, not for ka and kb
    srv2, reqs2 = _server(torch, budget=budget)
    srv2.max_bypass = max_bypass
    ka, kb, ca = reqs2[5], reqs2[1], reqs2[0]
    lives = [srv2.submit(r) for r in (ka, kb, ca)]
    srv2.run_until_idle()
    t = {lv.req.tag: lv.t_admit for lv in lives}
    print(f"max_bypass {max_bypass}: admitted order {sorted(t, key=t.get)}", flush=True)
    assert t[ka.tag] < t[kb.tag]
    assert (t[ca.tag] < t[kb.tag]) == (max_bypass > 0)
    assert len(srv2.done) == 3 and srv2.in_use == 0
    _check(torch, srv2, reqs2, srv2.done)
    srv2.close()


def test_serving_requests_match_oracle(torch_mod, orc):
    """Every served request of the mix against the FP64 oracle restatement of its program
    (oracle/requests.py): correlations within 1e-5 relative L2, kernels and Hessian-vector products
    within 1e-4."""
    torch = torch_mod
    from fd2d_noise_amd import device as D
    from oracle import requests as OR
    from conftest import rel_l2
    srv, reqs = _server(torch)
    srv.serve([(0.0, r) for r in reqs])
    tol = {"correlations": 1e-5, "K": 1e-4, "H": 1e-4}
    for live in sorted(srv.done, key=lambda d: d.req.tag):
        req = live.req
        m = srv.models[req.model]
        c = m.config
        args = (orc, c, D.to_host(m.mu), D.to_host(m.rho), m.spectrum_host, D.to_host(m.source_dist),
                np.atleast_2d(req.src), req.rec)
        if req.kind == "correlation":
            ref = OR.correlation(*args)
        elif req.kind == "kernel":
            ref = OR.kernel(*args, D.to_host(req.observed), req.measurement, req.veldis, req.window)
        else:
            ref = OR.hvp(*args, D.to_host(req.observed), D.to_host(req.dmu), req.measurement, req.veldis, req.window)
        for k, r in ref.items():
            e = rel_l2(D.to_host(live.result[k]), r)
            print(f"request {req.tag} {req.kind} on {req.model}: {k} rel L2 {e:.2e}", flush=True)
            assert e < tol[k], f"request {req.tag} {k}: {e:.3e}"
    srv.close()


class _Scarce:
    """The server's engine, except that a submission for which refuse(call number, solve) holds
    fails with FD2D_ENOMEM, as when device memory the budget does not count (torch's cached
    blocks, another process) is not free at that moment."""

    def __init__(self, engine, refuse):
        self._engine, self._refuse, self.calls, self.refused = engine, refuse, 0, 0

    def submit(self, spec):
        from fd2d_noise_amd import abi
        self.calls += 1
        if self._refuse(self.calls, spec):
            self.refused += 1
            raise abi.Fd2dError(abi.FD2D_ENOMEM, "device allocation refused")
        return self._engine.submit(spec)

    def __getattr__(self, name):
        return getattr(self._engine, name)


def test_serving_waits_for_device_memory(torch_mod):
    """Solves refused for want of device memory, also after torch returns its cached blocks, wait
    and are submitted on a later step; every request is still served, identical to alone."""
    torch = torch_mod
    srv, reqs = _server(torch)
    eng = srv.engine
    # while other solves run, a first try and its retry fail in turn (memory they hold is freed when
    # they end, so with nothing running every submission succeeds)
    scarce = _Scarce(eng, lambda n, spec: eng.active > 0 and n % 5 in (1, 2))
    srv.engine = scarce
    for r in reqs:
        srv.submit(r)
    waited = 0
    while srv.busy():
        srv.step()
        waited += sum(1 for lv in srv.live if lv.pending)
    torch.cuda.synchronize()
    print(f"{scarce.refused} of {scarce.calls} submissions refused, {waited} request-steps waiting", flush=True)
    assert scarce.refused > 0 and waited > 0
    assert len(srv.done) == len(reqs) and all(d.error is None for d in srv.done) and srv.in_use == 0
    srv.engine = scarce._engine
    _check(torch, srv, reqs, srv.done)
    srv.close()


def test_serving_abandons_request_that_never_fits(torch_mod):
    """A request whose solves are refused while nothing else runs (so nothing can free memory) ends
    with FD2D_ENOMEM; the server goes on serving the next one."""
    torch = torch_mod
    from fd2d_noise_amd import abi
    srv, reqs = _server(torch)
    scarce = _Scarce(srv.engine, lambda n, spec: spec.config.nx == 72)  # every solve on model "b"
    srv.engine = scarce
    kb, ca = reqs[1], reqs[0]
    assert (kb.model, ca.model) == ("b", "a")
    lb = srv.submit(kb)
    srv.run_until_idle()
    assert isinstance(lb.error, abi.Fd2dError) and lb.error.code == abi.FD2D_ENOMEM and lb.result is None
    assert srv.in_use == 0 and not srv.live
    la = srv.submit(ca)
    srv.run_until_idle()
    assert la.error is None
    srv.engine = scarce._engine
    _check(torch, srv, reqs, [la])
    srv.close()


def test_serving_ends_request_whose_solve_is_refused(torch_mod):
    """A solve the engine refuses for a reason other than memory (here FD2D_EINVAL for every solve
    on model "b") ends its request with that error; the other requests are served."""
    torch = torch_mod
    from fd2d_noise_amd import abi
    srv, reqs = _server(torch)
    eng = srv.engine

    class _Refusing(_Scarce):
        def submit(self, spec):
            if spec.config.nx == 72:
                raise abi.Fd2dError(abi.FD2D_EINVAL, "solve refused")
            return self._engine.submit(spec)

    srv.engine = _Refusing(eng, None)
    kb, ca, cb = reqs[1], reqs[0], reqs[3]
    assert (kb.model, ca.model, cb.model) == ("b", "a", "b")
    lives = [srv.submit(r) for r in (kb, ca, cb)]
    srv.run_until_idle()
    assert [lv.error.code if lv.error else None for lv in lives] == [abi.FD2D_EINVAL, None, abi.FD2D_EINVAL]
    assert srv.in_use == 0 and not srv.live
    srv.engine = eng
    _check(torch, srv, reqs, [lives[1]])
    srv.close()

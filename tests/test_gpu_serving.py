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
    small = S.Server(budget_bytes=1024)
    small.models = srv2.models
    with pytest.raises(ValueError):
        small.submit(reqs2[0])
    small.close()
    srv2.close()

"""Parity at the grids of the batched-stations, inversion and serving configurations, through the
batch engine and the serving loop, against the FP64 oracle.

* Batched stations (config3): 1200 x 1200 nodes on 2000 km (dx = 1.67 km), FD order 4, dt = 0.12 s,
  128 frequencies in 0.01-0.2 Hz, 4 source basis functions with a flat spectrum. Truncated to
  nt = 100 (100 Green and 199 correlation steps), two of the workload's stations run together
  through one engine, with receivers 10-40 km from each station so that the truncated runs reach
  them. Green spectra and correlation traces within 1e-5 relative L2.
* Inversion (config4): 600 x 600 nodes on 2000 km (dx = 3.34 km), order 4, dt = 0.25 s,
  91 frequencies, fw_nth = 1, structure inversion. Truncated to nt = 160 with a 3-station ring of
  radius 60 km around a +8% vs anomaly (the workload's 600 km ring is beyond the reach of a
  truncated run). The batched objective (1e-4), gradient (1e-3) and traces (1e-5), and a Rademacher
  Hessian-vector probe (1e-4), against the oracle composition (inversion/optlib_callback_fct/
  eval_objective_and_gradient.m:54-120, eval_hessian_vector_product_2.m:44-183).
* Serving (the full 200^2-4000^2 range): the admission decisions on the real model sizes (fw_nth per
  model, rejection of snapshot-infeasible Hessian-vector products, a 4000^2 kernel request queued and
  cancelled before it runs) and a 200^2 correlation request against the oracle.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fd2d_noise_amd import _lib
    _lib.load()
    return torch


def _near(rng, c, p, n, lo, hi):
    ang = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(lo, hi, n)
    return np.stack([p[0] + r * np.cos(ang), p[1] + r * np.sin(ang)], axis=1)


def test_config3_batched_stations(torch_mod, orc):
    torch = torch_mod
    from fd2d_noise_amd import batch as B, device as D
    from fd2d_noise_amd.workloads import config3_workload
    w = config3_workload(nt=100, n_stations=2)
    c = w["config"]
    rng = np.random.default_rng(303)
    st = w["stations"]
    rec = np.concatenate([_near(rng, c, st[i], 3, 10e3, 40e3) for i in range(len(st))], axis=0)
    mu, rho, spec, sd = w["mu"], w["rho"], w["spectrum"], w["source_dist"]
    mu_d, rho_d, sd_d = D.to_device(mu), D.to_device(rho), D.to_device(sd)
    G = [D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128) for _ in st]
    corr = [D.empty_like_dims((rec.shape[0], c.n_corr), torch.float64) for _ in st]
    B.run_batch([B.green(mu_d, rho_d, st[i:i + 1], None, 0, None, None, config=c, G_fft=G[i]) for i in range(len(st))])
    B.run_batch([B.correlation(mu_d, rho_d, G[i], spec, sd_d, rec, 0, None, None, config=c, seismograms=corr[i])
                 for i in range(len(st))])
    torch.cuda.synchronize()
    for i in range(len(st)):
        G_ref, _, _ = orc.run_forward1_green(mu, rho, st[i:i + 1], None, 0, None, None, config=c)
        corr_ref, _ = orc.run_forward2_correlation(mu, rho, G_ref, spec, sd, rec, 0, None, None, config=c)
        eg, ec = rel_l2(D.to_host(G[i]), G_ref), rel_l2(D.to_host(corr[i]), corr_ref)
        print(f"config3 station {i}: G_fft {eg:.2e}, correlation {ec:.2e}", flush=True)
        assert eg < 1e-5 and ec < 1e-5


def test_config4_grid_gradient_and_hvp(torch_mod, orc):
    torch = torch_mod
    from fd2d_noise_amd import inversion as V
    from fd2d_noise_amd.workloads import MU0, config4_workload
    from oracle import inversion as OI
    w = config4_workload(nt=160)
    c = w["config"]
    P = c.nx * c.nz
    centre = np.array([c.Lx / 2, c.Lz / 2])
    ang = 2 * np.pi * np.arange(3) / 3
    ring = np.stack([centre[0] + 60e3 * np.cos(ang), centre[1] + 60e3 * np.sin(ang)], axis=1)
    X, Z = c.x()[:, None], c.z()[None, :]
    blob = np.exp(-((X - centre[0]) ** 2 + (Z - centre[1]) ** 2) / (2 * 40e3 ** 2))
    usr = V.UsrPar(config=c, ref_stat=ring, array=ring, c_data=np.zeros((6, c.n_corr)), t=c.t_corr(),
                   rho=w["rho"], spectrum=w["spectrum"], type="structure", mu_0=MU0)
    # vs + 8% at the centre: mu = mu_0 (1 + m_structure) with (1.08)^2 - 1 at the peak
    m_true = np.concatenate([np.ones(P), (1.08 ** 2 - 1) * blob.ravel(order="F")])
    _, _, usr.c_data = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    j_ref, g_ref, c_ref = OI.eval_objective_and_gradient(m, usr, orc)
    prob = V.NoiseProblem(usr, batched=True)
    m_d = torch.as_tensor(m, device="cuda")
    j, g, c_all = prob.eval_objective_and_gradient(m_d)
    eg, ec = rel_l2(g.cpu().numpy(), g_ref), rel_l2(c_all.cpu().numpy().T, c_ref)
    print(f"config4 grid: j {j:.6e} (oracle {j_ref:.6e}), gradient {eg:.2e}, traces {ec:.2e}", flush=True)
    assert j == pytest.approx(j_ref, rel=1e-4) and eg < 1e-3 and ec < 1e-5
    z = np.random.default_rng(404).choice([-1.0, 1.0], P)
    dmp = np.concatenate([np.zeros(P), 0.01 * MU0 * z])
    H_ref = OI.eval_hessian_vector_product(m, dmp, usr, orc)
    H = prob.eval_hessian_vector_products(m_d, [dmp])[0].cpu().numpy()
    for part, sl in (("source", slice(0, P)), ("structure", slice(P, 2 * P))):
        e = rel_l2(H[sl], H_ref[sl])
        print(f"config4 grid: H dm {part} maps {e:.2e}", flush=True)
        assert e < 1e-4


def test_serving_full_range_admission(torch_mod, orc):
    torch = torch_mod
    from fd2d_noise_amd import device as D, serving as S
    from fd2d_noise_amd.workloads import SERVING_RANGE_SIZES, serving_models
    from oracle import requests as OR
    models = serving_models(sizes=SERVING_RANGE_SIZES, orders=(8,))
    # the budget is 80% of the available memory: return what earlier tests left cached
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    from fd2d_noise_amd import _lib
    _lib.load().fd2d_release_cached_memory()
    srv = S.Server(rounds=32)
    fw = {}
    for name, m in models.items():
        c = m["config"]
        try:
            f = S.kernel_fw_nth(c, 8, srv.budget // 2)
        except ValueError:
            f = S.kernel_fw_nth(c, 8, srv.budget)
        m["config"] = c = c.replace(fw_nth=f)
        fw[c.nx] = f
        srv.add_model(name, c, m["mu"], m["rho"], m["spectrum"], m["source_dist"])
    print(f"budget {srv.budget / 1e9:.0f} GB, fw_nth by size {fw}", flush=True)
    assert fw[200] == 1 and fw[4000] > 1
    assert all(fw[a] <= fw[b] for a, b in zip(SERVING_RANGE_SIZES, SERVING_RANGE_SIZES[1:]))

    def req(kind, n):
        name = f"n{n}_o8"
        m, c = models[name], models[name]["config"]
        obs = torch.zeros((c.n_corr, 3), dtype=torch.float64, device="cuda")
        dmu = D.to_device(0.01 * m["mu"]) if kind == "hvp" else None
        return S.Request(kind, name, m["stations"][:1], m["stations"][1:4], observed=obs, dmu=dmu, tag=(kind, n))
    hv_big = srv.submit(req("hvp", 4000))
    assert isinstance(hv_big.error, S.RequestRejected) and "fw_nth = 1" in str(hv_big.error)
    hv_894 = srv.submit(req("hvp", 894))
    assert isinstance(hv_894.error, S.RequestRejected) and "budget" in str(hv_894.error)
    k_big = srv.submit(req("kernel", 4000))
    assert k_big.error is None and k_big in srv.queue and k_big.bytes <= srv.budget
    assert srv.cancel(k_big) and not srv.busy()
    corr = srv.submit(req("correlation", 200))
    srv.run_until_idle()
    assert corr.error is None and corr.result is not None
    m, c = srv.models["n200_o8"], srv.models["n200_o8"].config
    ref = OR.correlation(orc, c, D.to_host(m.mu), D.to_host(m.rho), m.spectrum_host, D.to_host(m.source_dist),
                         np.atleast_2d(corr.req.src), corr.req.rec)
    e = rel_l2(D.to_host(corr.result["correlations"]), ref["correlations"])
    print(f"200^2 correlation request: {e:.2e}", flush=True)
    assert e < 1e-5
    srv.close()

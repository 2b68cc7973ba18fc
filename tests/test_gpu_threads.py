"""Re-entrancy: solver calls from several host threads at once (the C ABI's threading contract:
no global mutable state but the device context, the block cache and the tuning table, which are
locked). Each thread's outputs must equal the same call made alone, bit for bit; the host entry
points and the device entry points on per-thread streams are both exercised."""
import threading

import numpy as np
import pytest

import fd2d_noise_amd as fd

from helpers import blob_distribution, hetero_model, interior_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def _run_threads(fns):
    out, errs = [None] * len(fns), []

    def work(q):
        try:
            out[q] = fns[q]()
        except BaseException as e:  # reported below
            errs.append(e)
    ts = [threading.Thread(target=work, args=(q,)) for q in range(len(fns))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errs, errs
    assert all(not t.is_alive() for t in ts)
    return out


def test_host_entry_points_from_four_threads(torch_mod):
    c = fd.Config(order=4, fw_nth=1)
    rng = np.random.default_rng(950)
    mu, rho = hetero_model(rng, c)
    rec = interior_points(rng, c, 3)
    srcs = [interior_points(rng, c, 1, 0.4) for _ in range(4)]
    alone = [fd.run_forward1_green(mu, rho, s, rec, 0, config=c) for s in srcs]
    together = _run_threads([lambda s=s: fd.run_forward1_green(mu, rho, s, rec, 0, config=c) for s in srcs])
    for a, b in zip(alone, together):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[2], b[2])


def test_device_entry_points_on_per_thread_streams(torch_mod):
    torch = torch_mod
    from fd2d_noise_amd import device as D
    c = fd.Config(order=8, nx=80, nz=70, nt=200, fw_nth=1, Lx=79e3, Lz=69e3)
    rng = np.random.default_rng(951)
    mu, rho = hetero_model(rng, c)
    rec = interior_points(rng, c, 3)
    spectrum = D.to_device(np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None])
    sd = D.to_device(blob_distribution(rng, c))
    mu_d, rho_d, rec_d = D.to_device(mu), D.to_device(rho), D.to_device(rec)
    Gs = []
    for _ in range(3):
        G, _, _ = fd.run_forward1_green(mu, rho, interior_points(rng, c, 1, 0.4), rec, 0, config=c)
        Gs.append(D.to_device(G))

    def corr(G):
        out = D.empty_like_dims((rec.shape[0], c.n_corr), torch.float64)
        D.run_forward2_correlation(mu_d, rho_d, G, spectrum, sd, rec_d, 0, None, None, config=c, seismograms=out)
        return out

    alone = [D.to_host(corr(G)) for G in Gs]

    def on_stream(G):
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            r = corr(G)
        s.synchronize()
        return D.to_host(r)
    together = _run_threads([lambda G=G: on_stream(G) for G in Gs])
    for a, b in zip(alone, together):
        assert np.array_equal(a, b)

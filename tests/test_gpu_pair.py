"""Two time steps per launch (temporal blocking) against one step per launch, through the C ABI.

A pair launch recomputes the halo of the first step inside each tile with the operations of a
single step, so both launch shapes give the same wavefields; the comparison allows 1e-6 relative
L2 for the compiler's freedom to contract multiply-adds differently in the two kernels (measured:
identical). One step per launch is the default and is held to the oracle by test_gpu_parity.py.
"""
import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import Config, _lib

from conftest import rel_l2
from helpers import blob_distribution, hetero_model, interior_points

pytestmark = pytest.mark.gpu
TOL_SAME = 1e-6


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    yield
    _lib.set_steps_per_launch(0)


def both(fn):
    out = []
    for k in (1, 2):
        _lib.set_steps_per_launch(k)
        out.append(fn())
    _lib.set_steps_per_launch(0)
    return out


def _cmp(name, a, b):
    for q, (x, y) in enumerate(zip(a, b)):
        x, y = np.asarray(x), np.asarray(y)
        assert np.all(np.isfinite(y)), f"{name}[{q}]: unstable set-up"
        err = rel_l2(x, y) if np.any(y != 0) else float(np.abs(x).max())
        print(f"{name}[{q}]: one vs two steps per launch {err:.2e}")
        assert x.shape == y.shape and err <= TOL_SAME


@pytest.mark.parametrize("order,nx,nz", [(2, 50, 50), (4, 67, 45), (6, 50, 50), (8, 130, 97)])
def test_green_pair(order, nx, nz):
    # dx = dz ~ 1 km at every size: CFL 4.4 km/s * 0.09 s / dx stays below the order-8 limit
    c = Config(order=order, nx=nx, nz=nz, fw_nth=3, Lx=1000.0 * (nx - 1), Lz=1000.0 * (nz - 1))
    rng = np.random.default_rng(700 + order)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 4)
    one, two = both(lambda: fd.run_forward1_green(mu, rho, src, rec, 1, config=c))
    _cmp(f"green order {order}", one, two)


@pytest.mark.parametrize("nb,fw_nth,freq_samp", [(0, 2, 5), (3, 1, 3), (0, 4, 1)])
def test_correlation_pair(orc, nb, fw_nth, freq_samp):
    c = Config(n_basis_fct=nb, fw_nth=fw_nth, freq_samp=freq_samp, order=4, nx=70, nz=52)
    rng = np.random.default_rng(710 + nb)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 3)
    G, _, _ = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    spectrum = rng.uniform(0.5, 1.5, (c.n_sample, 2 if nb else 1))
    sd = blob_distribution(rng, c, nb=nb)
    one, two = both(lambda: fd.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c))
    _cmp(f"correlation nb {nb} f {fw_nth} s {freq_samp}", one, two)


@pytest.mark.parametrize("mode,fw_nth", [(1, 4), (1, 3), (0, 2), (13, 2)])
def test_adjoint_pair(orc, mode, fw_nth):
    c = Config(fw_nth=fw_nth, order=8, nx=80, nz=64, Lx=79e3, Lz=63e3)
    rng = np.random.default_rng(720 + mode + fw_nth)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.45)
    rec = interior_points(rng, c, 3)
    G, Gout, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    spectrum = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
    sd = blob_distribution(rng, c)
    corr, C = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    stf = np.ascontiguousarray(corr[:, ::-1] / np.abs(corr).max()).astype(np.complex128)
    if mode == 0:
        # distributed source from a mode-1 run, structure kernel against the Green snapshots
        _, a1, _, _ = orc.run_noise_adjoint(mu, rho, C, stf, rec, rec, None, None, None, 1, config=c)
        one, two = both(lambda: fd.run_noise_adjoint(mu, rho, Gout, a1, rec, rec, spectrum, sd, None, 0, config=c))
    else:
        one, two = both(lambda: fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, mode, config=c))
    _cmp(f"adjoint mode {mode} f {fw_nth}", one, two)


@pytest.mark.parametrize("order,nx,nz", [(2, 50, 50), (4, 67, 45), (6, 50, 50)])
def test_product_steps_each_order(orc, order, nx, nz):
    # single steps carry the K_mu product in k_tile<R,2>; two-step launches leave it to k_kmu<R>
    c = Config(order=order, nx=nx, nz=nz, fw_nth=3, Lx=1000.0 * (nx - 1), Lz=1000.0 * (nz - 1))
    rng = np.random.default_rng(740 + order)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.45)
    rec = interior_points(rng, c, 2)
    G, _, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    spectrum = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
    corr, C = orc.run_forward2_correlation(mu, rho, G, spectrum, blob_distribution(rng, c), rec, 1, config=c)
    stf = np.ascontiguousarray(corr[:, ::-1] / np.abs(corr).max()).astype(np.complex128)
    one, two = both(lambda: fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, 1, config=c))
    assert np.abs(np.asarray(one[0])).max() > 0, "no structure kernel accumulated"
    _cmp(f"adjoint products order {order}", one, two)


@pytest.mark.parametrize("order", [4, 8])
def test_tuned_launch_shape_matches_single_steps(orc, order):
    """Mode 0: the first solve of a shape times single and two-step launches on its own early steps
    and keeps the faster; every output equals one step per launch."""
    c = Config(order=order, nx=96, nz=88, nt=150, fw_nth=2, Lx=95e3, Lz=87e3)
    rng = np.random.default_rng(730 + order)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 3)
    G, _, _ = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    spectrum = rng.uniform(0.5, 1.5, (c.n_sample, 1))
    sd = blob_distribution(rng, c)
    outs = []
    for k in (1, 0, 0):  # the second mode-0 run uses the shape tuned by the first
        _lib.set_steps_per_launch(k)
        outs.append(fd.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c))
    _lib.set_steps_per_launch(0)
    _cmp(f"tuned order {order}", outs[0], outs[1])
    _cmp(f"tuned again order {order}", outs[0], outs[2])

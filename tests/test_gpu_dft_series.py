"""Where a DFT series starts and ends (G_fft of code/run_forward1_green.m:176-183, stf_fft of
code/run_noise_adjoint.m:335-346), through the C ABI.

The device DFT contracts the sampled steps slab by slab (up to 1024 samples): the first slab of a
series starts the sums without reading them, and the slab that completes the series writes the
finished output instead of the sums. These tests hold the ends of that scheme to the FP64 oracle: a
series with no sample at all (the output is zero, whatever the output array held before), and series
whose length is an exact multiple of the slab, so that the last slab is contracted as it fills,
after two and three whole slabs.
"""
import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import Config

from conftest import rel_l2
from helpers import hetero_model

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fd2d_noise_amd import _lib
    _lib.load()


def _setup(nt, seed=71):
    c = Config(nx=40, nz=40, Lx=39 * 1200.0, Lz=39 * 1200.0, nt=nt, dt=0.09, order=4, fw_nth=1, freq_samp=5)
    rng = np.random.default_rng(seed)
    mu, rho = hetero_model(rng, c)
    src = np.array([[0.45 * c.Lx, 0.55 * c.Lz]])
    rec = np.array([[0.3 * c.Lx, 0.4 * c.Lz], [0.7 * c.Lx, 0.6 * c.Lz]])
    return c, mu, rho, src, rec


def _green_samples(c):
    return sum(1 for n in range(1, c.nt + 1) if (n + c.nt - 1) % c.freq_samp == 0)


def _adjoint_samples(c):
    return (2 * c.nt - 1) // c.freq_samp


def test_series_without_samples_is_zero():
    import torch
    from fd2d_noise_amd import device as D
    c, mu, rho, src, rec = _setup(nt=2)
    assert _green_samples(c) == 0 and _adjoint_samples(c) == 0
    mu_d, rho_d, src_d, rec_d = (D.to_device(a) for a in (mu, rho, src, rec))
    G = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
    G.fill_(7.0 + 3.0j)  # whatever the output held, a series without samples is zero
    D.run_forward1_green(mu_d, rho_d, src_d, rec_d, 0, None, None, config=c, G_fft=G)
    stf = D.to_device(np.ones((rec.shape[0], c.n_corr), dtype=np.complex128))
    K = D.empty_like_dims((c.nx, c.nz, 2), torch.float64)
    A = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
    A.fill_(-5.0)
    D.run_noise_adjoint(mu_d, rho_d, None, stf, rec_d, rec_d, None, None, None, 1, None, None, config=c, K=K,
                        stf_fft=A)
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(G)) == 0
    assert int(torch.count_nonzero(A)) == 0


@pytest.mark.parametrize("nt, slabs", [(10240, 2)])
def test_green_series_of_whole_slabs(orc, nt, slabs):
    c, mu, rho, src, rec = _setup(nt=nt)
    assert _green_samples(c) == 1024 * slabs
    G, _, seis = fd.run_forward1_green(mu, rho, src, rec, 0, config=c)
    G0, _, seis0 = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    e = rel_l2(G, G0)
    print(f"G_fft over {slabs} whole slabs: {e:.2e} (seismograms {rel_l2(seis, seis0):.2e})", flush=True)
    assert e < 1e-5


@pytest.mark.parametrize("nt, slabs", [(5121, 2), (7681, 3)])
def test_adjoint_series_of_whole_slabs(orc, nt, slabs):
    c, mu, rho, src, rec = _setup(nt=nt, seed=72)
    assert _adjoint_samples(c) == 1024 * slabs
    rng = np.random.default_rng(73)
    t = c.t_corr()
    stf = np.stack([np.exp(-((t - 40.0 * (q + 1)) / 30.0) ** 2) for q in range(rec.shape[0])]).astype(np.complex128)
    stf += 0.05 * rng.standard_normal(stf.shape)
    _, A, _, _ = fd.run_noise_adjoint(mu, rho, None, stf, rec, rec, None, None, None, 1, config=c)
    _, A0, _, _ = orc.run_noise_adjoint(mu, rho, None, stf, rec, rec, None, None, None, 1, config=c)
    e = rel_l2(A, A0)
    print(f"stf_fft over {slabs} whole slabs: {e:.2e}", flush=True)
    assert e < 1e-5

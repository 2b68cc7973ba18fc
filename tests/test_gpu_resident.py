"""Device-resident arrays (fd2d_resident_* of the C ABI) as the model and Green-spectrum handles:
mu and rho uploaded once, G_fft left on the device between the Green run and the correlation run,
results equal to the host entry points bit for bit, and the handles' own round trip and errors."""
import ctypes

import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import _lib, abi
from fd2d_noise_amd.resident import Resident

from helpers import blob_distribution, hetero_model, interior_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()


def _host(a):
    """Host descriptor of a small input (src, rec, spectrum), read in place by the device entries."""
    arr = np.asfortranarray(np.asarray(a, dtype=np.float64))
    d = abi.CArray()
    d.data = arr.ctypes.data
    d.dims[:] = abi.matlab_dims(arr.shape)
    return arr, ctypes.pointer(d)


def _check(st):
    assert st == abi.FD2D_OK, _lib.last_error()


def test_green_and_correlation_through_resident_handles():
    c = fd.Config(order=4, fw_nth=1)
    rng = np.random.default_rng(900)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 3)
    spectrum = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
    sd = blob_distribution(rng, c)
    G0, _, seis0 = fd.run_forward1_green(mu, rho, src, rec, 0, config=c)
    corr0, _ = fd.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 0, config=c)

    lib = _lib.load()
    cc = c.to_c()
    src_h, src_d = _host(src)
    rec_h, rec_d = _host(rec)
    spec_h, spec_d = _host(spectrum)
    with Resident.upload(mu) as mu_r, Resident.upload(rho) as rho_r, Resident.upload(sd) as sd_r, \
            Resident.alloc((c.nx, c.nz, c.n_sample), np.complex128) as G_r, \
            Resident.alloc((rec.shape[0], c.nt)) as seis_r, Resident.alloc((rec.shape[0], c.n_corr)) as corr_r:
        _check(lib.fd2d_run_forward1_green_device(ctypes.byref(cc), mu_r.carray(), rho_r.carray(), src_d, rec_d,
                                                  0.0, None, None, G_r.marray(), None, seis_r.marray(), None))
        # the Green spectrum goes to the correlation run without leaving the device
        _check(lib.fd2d_run_forward2_correlation_device(ctypes.byref(cc), mu_r.carray(), rho_r.carray(),
                                                        G_r.carray(), spec_d, sd_r.carray(), rec_d, 0.0, None,
                                                        None, corr_r.marray(), None, None))
        G, seis, corr = G_r.download(), seis_r.download(), corr_r.download()
    print(f"resident: G_fft {np.abs(G - G0).max():.1e}, seismograms {np.abs(seis - seis0).max():.1e}, "
          f"correlation {np.abs(corr - corr0).max():.1e}")
    assert np.array_equal(G, G0) and np.array_equal(seis, seis0) and np.array_equal(corr, corr0)


def test_resident_round_trip_and_size_check():
    rng = np.random.default_rng(901)
    z = rng.standard_normal((7, 5, 3)) + 1j * rng.standard_normal((7, 5, 3))
    with Resident.upload(z, np.complex128) as r:
        assert np.array_equal(r.download(), z)
        small = np.zeros((7, 5, 2), dtype=np.complex128, order="F")
        d = abi.MArray()
        d.data = small.ctypes.data
        d.dims[:] = (7, 5, 2)
        assert _lib.load().fd2d_resident_download(r._h, ctypes.byref(d), None) == abi.FD2D_EINVAL
        assert "bytes" in _lib.last_error()
    with Resident.alloc((4, 4), np.float32) as r:
        assert r.download().dtype == np.float32 and not r.download().any()


def test_cached_tables_and_memory_release():
    """Solves of one configuration share cached phase tables; releasing the cached memory between
    two identical correlation runs (tables, buffers) leaves the second result unchanged, and a run
    with other frequencies (its own tables) still matches its first run."""
    from fd2d_noise_amd.config import matlab_colon
    c = fd.Config(order=4, fw_nth=1)
    c2 = fd.Config(order=4, fw_nth=1, f_sample=matlab_colon(0.03, 0.002, 0.15))
    rng = np.random.default_rng(902)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 3)
    sd = blob_distribution(rng, c)
    out = []
    for cfg in (c, c2, c, c2):
        spectrum = np.exp(-((np.asarray(cfg.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
        G, _, _ = fd.run_forward1_green(mu, rho, src, rec, 0, config=cfg)
        corr, _ = fd.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 0, config=cfg)
        out.append((G, corr))
        assert _lib.load().fd2d_release_cached_memory() == abi.FD2D_OK
    for a, b in ((out[0], out[2]), (out[1], out[3])):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert out[0][0].shape[2] != out[1][0].shape[2]

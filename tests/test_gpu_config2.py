"""Parity at the single-large-solve benchmark configuration, through the C ABI.

The benchmark configuration is 2400 x 2400 nodes on 2000 km (dx = 0.834 km), FD order 8, dt = 0.06 s,
nt = 8000, forward fields stored every 4th step, the default 121 frequencies (0.02:0.0015:0.2 Hz,
input/input_interferometry.m:16-18) and a seeded heterogeneous vs (4 km/s, +-10%, 30 km correlation
length). Its full length is beyond what the FP64 oracle runs in a test, so it is pinned in two parts:

  * the full 2400 x 2400 grid, order 8, truncated to nt = 100 (100 Green steps, 199-step
    correlation and adjoint runs): every tile, the padded layout and the frequency-domain slabs at
    the real size, with receivers within reach of the waves in that time;
  * the full run length on a 256 x 256 grid with the same spacing, time step, order, fw_nth and
    frequencies: nt = 8000, i.e. 15,999-step correlation and mode-1 adjoint runs and an 8,000-step
    mode-0 adjoint run, so FP32 round-off accumulates over as many steps as in the benchmark.

The chain is that of one station's gradient (inversion/optlib_callback_fct/
eval_objective_and_gradient.m:52-98): Green run (mode 1, snapshots), correlation run (mode 1), adjoint
run mode 1 with point adjoint sources at the receivers (structure kernel against the correlation
snapshots and the DFT of the adjoint state), adjoint run mode 0 with that DFT as a distributed
source (half length, structure kernel against the Green snapshots). Each path runs its own chain;
the adjoint sources and the mode-0 source are the oracle's, so each stage's error is its own.

Tolerances (relative L2 against the FP64 oracle): correlation traces 1e-5, kernels 1e-4, G_fft,
stf_fft, snapshots and Green traces 1e-5.
"""
import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import Config
from fd2d_noise_amd.config import matlab_colon
from fd2d_noise_amd.workloads import config2_workload, gaussian_random_field

pytestmark = pytest.mark.gpu

TOL_CORR = 1e-5
TOL_KERNEL = 1e-4
TOL_FIELD = 1e-5


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fd2d_noise_amd import _lib
    _lib.load()


def rel(a, b, chunk=1 << 24):
    """||a - b|| / ||b|| over all elements, in chunks (the arrays here reach 11 GB)."""
    a = np.asarray(a).reshape(-1, order="F")
    b = np.asarray(b).reshape(-1, order="F")
    num = den = 0.0
    for q in range(0, b.size, chunk):
        d = a[q:q + chunk] - b[q:q + chunk]
        num += float(np.vdot(d, d).real)
        den += float(np.vdot(b[q:q + chunk], b[q:q + chunk]).real)
    return float(np.sqrt(num / den)) if den > 0 else float(np.sqrt(num))


def _ring(rng, centre, n, rmin, rmax):
    ang = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(rmin, rmax, n)
    return np.stack([centre[0] + r * np.cos(ang), centre[1] + r * np.sin(ang)], axis=1)


def _chain(orc, c, mu, rho, src, rec, spectrum, sd, tag):
    errs = {}
    G, Go, sg = fd.run_forward1_green(mu, rho, src, rec, 1, config=c)
    G0, Go0, sg0 = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    errs["green_seis"], errs["G_fft"], errs["G_out"] = rel(sg, sg0), rel(G, G0), rel(Go, Go0)
    print(f"{tag}: green seis {errs['green_seis']:.2e} G_fft {errs['G_fft']:.2e} G_out {errs['G_out']:.2e}",
          flush=True)
    cc, C = fd.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    del G
    cc0, C0 = orc.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 1, config=c)
    del G0
    errs["correlation"], errs["C_out"] = rel(cc, cc0), rel(C, C0)
    print(f"{tag}: correlation {errs['correlation']:.2e} C_out {errs['C_out']:.2e} "
          f"(|c|max {np.abs(cc0).max():.3e})", flush=True)
    stf = np.ascontiguousarray(cc0[:, ::-1] / np.abs(cc0).max()).astype(np.complex128)
    K1, sf, _, s1 = fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, None, None, None, 1, config=c)
    del C
    K10, sf0, _, s10 = orc.run_noise_adjoint(mu, rho, C0, stf, rec, rec, None, None, None, 1, config=c)
    del C0
    errs["adjoint1_seis"], errs["K_mu_1"], errs["stf_fft"] = rel(s1, s10), rel(K1[:, :, -1], K10[:, :, -1]), rel(sf, sf0)
    print(f"{tag}: adjoint mode 1 seis {errs['adjoint1_seis']:.2e} K_mu {errs['K_mu_1']:.2e} "
          f"stf_fft {errs['stf_fft']:.2e}", flush=True)
    del sf
    K0, _, _, _ = fd.run_noise_adjoint(mu, rho, Go, sf0, rec, None, spectrum, sd, None, 0, config=c)
    K00, _, _, _ = orc.run_noise_adjoint(mu, rho, Go0, sf0, rec, None, spectrum, sd, None, 0, config=c)
    errs["K_mu_0"] = rel(K0[:, :, -1], K00[:, :, -1])
    print(f"{tag}: adjoint mode 0 K_mu {errs['K_mu_0']:.2e} (|K|max {np.abs(K00[:, :, -1]).max():.3e})", flush=True)
    assert np.abs(cc0).max() > 0 and np.abs(K00[:, :, -1]).max() > 0 and np.abs(K10[:, :, -1]).max() > 0
    for k in ("green_seis", "G_fft", "G_out", "C_out", "stf_fft", "adjoint1_seis"):
        assert errs[k] < TOL_FIELD, (k, errs[k])
    assert errs["correlation"] < TOL_CORR
    assert errs["K_mu_1"] < TOL_KERNEL and errs["K_mu_0"] < TOL_KERNEL
    return errs


def test_config2_full_grid_truncated(orc):
    """2400 x 2400, order 8, the benchmark's model, spectrum and source, nt = 100; eight receivers
    4-20 km from the reference station (the waves travel 24 km in 100 steps)."""
    w = config2_workload(nt=100)
    c = w["config"]
    assert (c.nx, c.nz, c.order, c.fw_nth, c.n_sample, c.dt) == (2400, 2400, 8, 4, 121, 0.06)
    rec = _ring(np.random.default_rng(22), w["src"][0], 8, 4e3, 20e3)
    _chain(orc, c, w["mu"], w["rho"], w["src"], rec, w["spectrum"], w["source_dist"], "2400^2 nt=100")


def test_config2_run_length(orc):
    """The benchmark's run length (nt = 8000: 15,999-step correlation and adjoint runs) on a
    256 x 256 grid with its spacing, time step, order, fw_nth, frequencies and model statistics;
    16 receivers 25-80 km from the central station."""
    dx = 2.0e6 / 2399
    n = 256
    c = Config(Lx=dx * (n - 1), Lz=dx * (n - 1), nx=n, nz=n, dt=0.06, nt=8000, order=8, fw_nth=4,
               f_sample=matlab_colon(0.02, 0.0015, 0.2), freq_samp=5)
    c.absorb_width = 30 * c.dx
    rng = np.random.default_rng(23)
    vs = 4000.0 * (1 + 0.1 * gaussian_random_field(rng, n, n, c.dx, 30e3))
    rho = 3000.0 * np.ones((n, n), order="F")
    mu = np.asfortranarray(rho * vs ** 2)
    centre = np.array([c.Lx / 2, c.Lz / 2])
    rec = _ring(rng, centre, 16, 25e3, 80e3)
    f = np.asarray(c.f_sample)
    spectrum = np.exp(-((f - 0.1) ** 2) / (2 * 0.03 ** 2))[:, None]
    X, Z = c.x()[:, None], c.z()[None, :]
    sd = np.asfortranarray(0.1 + np.exp(-((X - 0.3 * c.Lx) ** 2 + (Z - 0.6 * c.Lz) ** 2) / (2 * (40e3) ** 2)))
    _chain(orc, c, mu, rho, centre[None, :], rec, spectrum, sd, "256^2 nt=8000")

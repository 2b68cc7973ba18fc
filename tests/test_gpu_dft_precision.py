"""Error bound of the on-the-fly DFT (stf_fft of code/run_noise_adjoint.m:335-346, the same
accumulation as G_fft of code/run_forward1_green.m:176-183), through the C ABI.

The device DFT forms each product x_j c_jk of an FP32 sample and the phase factor c_jk =
e^{-i w_k t_j} (rounded to FP32) from six bfloat16 part products on the matrix cores, sums them in
FP32 over a slab of up to 1024 samples, and adds every slab sum into FP64 accumulators with Neumaier
compensation; the (2 pi)^-1/2 dt scale is applied in FP64 at the end. Per component (real and
imaginary part separately), against the exact FP64 DFT of the same FP32 samples:

    |A_gpu - A| <= s [ 2^-22 B1 + 96 u (Bc + 2 B1) ] (1 + 2^-10)

with s = (2 pi)^-1/2 dt, u = 2^-24, B1 = sum_j |x_j| |c_jk| (the phase rounding, 2^-24, and the three
part products left out of each product, 2^-23 + 2^-32, are below 2^-22 of |x_j c_jk|), and Bc the
sum, over every 16-sample chunk of every slab, of the magnitude of the exact partial sum of that slab
before the chunk: each of a chunk's six matrix-core instructions adds 16 products to the FP32
accumulator and rounds at most 16 times, each time by at most u of the running sum, which is at most
the slab's partial sum plus twice the chunk's B1. The compensated FP64 accumulation across slabs and
the final scaling add below 4 * 2^-53 of the sum, which the (1 + 2^-10) factor covers.

The samples are recovered from the same run's snapshots (mode 13 stores u before each step's update,
so snapshot n + 1 is the DFT sample of step n), so the comparison isolates the DFT arithmetic from the
wavefield's own FP32 error. The series is 1600 samples of the displacement under a constant point
force plus a 20% oscillation: a large and growing DC component that the FP32 slab sums must carry.
The same stf_fft is also held to the FP64 oracle (relative L2 1e-5).
"""
import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import Config

from conftest import rel_l2
from helpers import hetero_model

pytestmark = pytest.mark.gpu

SLAB = 1024  # samples per DFT slab (kSpecSlab; this small grid never shortens it)
CHUNK = 16   # samples per matrix-core instruction


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fd2d_noise_amd import _lib
    _lib.load()


def test_dft_error_bound_long_series_with_dc(orc):
    c = Config(nx=40, nz=40, Lx=39 * 1200.0, Lz=39 * 1200.0, nt=4001, dt=0.09, order=4, fw_nth=1)
    N = c.n_corr
    rng = np.random.default_rng(31)
    mu, rho = hetero_model(rng, c)
    src = np.array([[0.45 * c.Lx, 0.55 * c.Lz]])
    t = c.t_corr()
    stf = (1.0 + 0.2 * np.cos(2 * np.pi * 0.05 * t))[None, :].astype(np.complex128)
    _, A_gpu, U, _ = fd.run_noise_adjoint(mu, rho, None, stf, src, None, None, None, None, 13, config=c)
    _, A_orc, _, _ = orc.run_noise_adjoint(mu, rho, None, stf, src, None, None, None, None, 13, config=c)
    e_orc = rel_l2(A_gpu, A_orc)
    # the DFT samples: u after the sampled steps n = s, 2s, ... < N (snapshot index n, 0-based)
    steps = np.arange(c.freq_samp, N, c.freq_samp)
    assert steps.size >= 1600 and U.shape[2] == N
    X = U.reshape(c.nx * c.nz, N, order="F")[:, steps].astype(np.float64)  # (P, J), exact FP32 values
    tj = (steps - c.nt) * c.dt
    E = np.exp(-1j * np.outer(tj, c.w_sample))                               # (J, K) FP64 phase factors
    scale = c.dt / np.sqrt(2 * np.pi)
    A = scale * (X @ E)                                                      # (P, K) exact FP64 DFT
    G = A_gpu.reshape(c.nx * c.nz, c.n_sample, order="F")
    u = 2.0 ** -24
    P, J = X.shape
    for part, F in (("real", np.real), ("imag", np.imag)):
        Ec = np.abs(F(E))
        B1 = np.abs(X) @ Ec
        Bc = np.zeros_like(B1)
        for s0 in range(0, J, SLAB):
            s1 = min(J, s0 + SLAB)
            run = np.zeros_like(B1)
            for q0 in range(s0, s1, CHUNK):
                Bc += np.abs(run)  # the slab's partial sum before this chunk
                q1 = min(s1, q0 + CHUNK)
                run += X[:, q0:q1] @ F(E[q0:q1])
        bound = scale * (2.0 ** -22 * B1 + 96 * u * (Bc + 2 * B1)) * (1 + 2.0 ** -10)
        err = np.abs(F(G) - F(A))
        ratio = err / np.maximum(bound, 1e-300)
        rel_b1 = err / np.maximum(scale * B1, 1e-300)
        print(f"{part}: max |err| / bound {ratio.max():.3e}, max |err| / (s B1) {rel_b1.max():.3e}, "
              f"median |err| / (s B1) {np.median(rel_b1):.3e}, max s Bc / (s B1) {(Bc / np.maximum(B1, 1e-300)).max():.2f}",
              flush=True)
        assert (err <= bound).all(), f"{part}: {int((err > bound).sum())} outputs exceed the bound"
    print(f"stf_fft against the FP64 oracle: {e_orc:.2e}; DC share of the samples "
          f"{np.abs(X.mean(axis=1)).sum() / np.abs(X).mean(axis=1).sum():.2f}", flush=True)
    assert e_orc < 1e-5

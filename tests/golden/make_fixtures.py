"""Writes tests/golden/tiny_chain.npz: inputs and outputs of one station's gradient chain on the
reference's tiny set-up (input/input_parameters.m:8-16), computed by the NumPy restatement of the
solvers (oracle/numpy_solvers.py). TEST INFRASTRUCTURE.

Chain (inversion/optlib_callback_fct/eval_objective_and_gradient.m:52-98): Green run mode 1
(fw_nth = 2), correlation run mode 1 on its G_fft, adjoint run mode 1 with the stored point adjoint
sources at the receivers (structure kernel against C, stf_fft), adjoint run mode 0 with that stf_fft as
distributed source (half length, structure kernel against the Green snapshots). Stored: the inputs,
every trace, both kernels, and G_fft / stf_fft at five frequencies (the fixture stays small).

    python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from fd2d_noise_amd import Config  # noqa: E402
from oracle import numpy_solvers as NP  # noqa: E402

K_SUB = np.array([0, 30, 60, 90, 120])


def config():
    return Config(fw_nth=2)  # the tiny set-up: 50 x 50 on 60 km, dt 0.09 s, nt 150, order 4, 121 frequencies


def inputs(c):
    rng = np.random.default_rng(2024)
    x, z = np.meshgrid(np.arange(c.nx) / (c.nx - 1), np.arange(c.nz) / (c.nz - 1), indexing="ij")
    vs = 4000.0 * (1 + 0.05 * np.sin(2 * np.pi * x) * np.cos(3 * np.pi * z) + 0.02 * rng.standard_normal((c.nx, c.nz)))
    rho = 3000.0 * np.ones((c.nx, c.nz))
    mu = rho * vs ** 2
    src = np.array([[0.45 * c.Lx, 0.55 * c.Lz]])
    rec = np.array([[0.3 * c.Lx, 0.4 * c.Lz], [0.62 * c.Lx, 0.35 * c.Lz], [0.5 * c.Lx, 0.75 * c.Lz]])
    f = np.asarray(c.f_sample)
    spectrum = np.exp(-((f - 0.1) ** 2) / 0.05 ** 2)[:, None]
    sd = 0.1 + np.exp(-((x - 0.35) ** 2 + (z - 0.6) ** 2) / 0.02)
    stf = (rng.standard_normal((rec.shape[0], c.n_corr)) * np.hanning(c.n_corr)[None, :]).astype(np.complex128)
    return dict(mu=np.asfortranarray(mu), rho=np.asfortranarray(rho), src=src, rec=rec, spectrum=spectrum,
                source_dist=np.asfortranarray(sd), stf=stf)


def chain(solvers, c, d, np_style):
    """The chain with `solvers` (the NumPy module, or a C-ABI-shaped solver object)."""
    kw = dict(config=c)
    if np_style:
        G, Gout, sg = solvers.run_forward1_green(d["mu"], d["rho"], d["src"], d["rec"], 1, None, None, **kw)
        cc, C = solvers.run_forward2_correlation(d["mu"], d["rho"], G, d["spectrum"], d["source_dist"], d["rec"], 1,
                                                 None, None, **kw)
        K1, sf, _, s1 = solvers.run_noise_adjoint(d["mu"], d["rho"], C, d["stf"], d["rec"], d["rec"], None, None, None,
                                                  1, None, None, **kw)
        K0, _, _, _ = solvers.run_noise_adjoint(d["mu"], d["rho"], Gout, sf, d["rec"], None, d["spectrum"],
                                                d["source_dist"], None, 0, None, None, **kw)
    else:
        G, Gout, sg = solvers.run_forward1_green(d["mu"], d["rho"], d["src"], d["rec"], 1, **kw)
        cc, C = solvers.run_forward2_correlation(d["mu"], d["rho"], G, d["spectrum"], d["source_dist"], d["rec"], 1,
                                                 **kw)
        K1, sf, _, s1 = solvers.run_noise_adjoint(d["mu"], d["rho"], C, d["stf"], d["rec"], d["rec"], None, None, None,
                                                  1, **kw)
        K0, _, _, _ = solvers.run_noise_adjoint(d["mu"], d["rho"], Gout, sf, d["rec"], None, d["spectrum"],
                                                d["source_dist"], None, 0, **kw)
    return dict(green_seis=sg, G_fft_sub=np.asarray(G)[:, :, K_SUB], correlation=cc, adj1_K=K1, adj1_seis=s1,
                stf_fft_sub=np.asarray(sf)[:, :, K_SUB], adj0_K=K0)


def main():
    c = config()
    d = inputs(c)
    out = chain(NP, c, d, True)
    np.savez_compressed(os.path.join(HERE, "tiny_chain.npz"), k_sub=K_SUB,
                        **{"in_" + k: v for k, v in d.items()}, **{"out_" + k: v for k, v in out.items()})
    print("wrote", os.path.join(HERE, "tiny_chain.npz"))


if __name__ == "__main__":
    main()

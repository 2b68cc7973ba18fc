"""Parity of the gfx950 HIP engine (through the C ABI) with the FP64 oracle on identical inputs.

Tolerances (relative L2 against the FP64 oracle, FP32 wavefields on the GPU):
  correlation seismograms 1e-5, kernels 1e-4 (the targets of the engine), Green seismograms,
  G_fft, stf_fft and snapshots 1e-5, Born runs 1e-4 (their reference form differences consecutive
  FP32 snapshots).
"""
import numpy as np
import pytest

import fd2d_noise_amd as fd
from fd2d_noise_amd import Config, abi

from conftest import rel_l2
from helpers import blob_distribution, hetero_model, homog_model, interior_points, smooth_field

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


def tiny(**kw):
    """The reference's tiny validation setup (input_parameters.m:8-16) with overrides."""
    return Config(**kw)


def case(rng, c):
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.4)
    rec = interior_points(rng, c, 3)
    return mu, rho, src, rec


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("mode,fw_nth", [(0, 1), (1, 1), (1, 3)])
def test_green(orc, order, mode, fw_nth):
    c = tiny(order=order, fw_nth=fw_nth)
    rng = np.random.default_rng(100 + order)
    mu, rho, src, rec = case(rng, c)
    G, Go, s = fd.run_forward1_green(mu, rho, src, rec, mode, config=c)
    G0, Go0, s0 = orc.run_forward1_green(mu, rho, src, rec, mode, config=c)
    print(f"green order={order} mode={mode} f={fw_nth}: seis {rel_l2(s, s0):.2e} G_fft {rel_l2(G, G0):.2e}")
    assert rel_l2(s, s0) < TOL_FIELD
    assert rel_l2(G, G0) < TOL_FIELD
    if mode:
        assert Go.shape == Go0.shape
        assert rel_l2(Go, Go0) < TOL_FIELD


@pytest.mark.parametrize("nb,n_src,order", [(0, 1, 4), (3, 2, 4), (0, 1, 8)])
def test_correlation(orc, nb, n_src, order):
    c = tiny(n_basis_fct=nb, order=order, fw_nth=2)
    rng = np.random.default_rng(200 + nb)
    mu, rho, src, rec = case(rng, c)
    G0, _, _ = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    spectrum = rng.uniform(0.5, 1.5, (c.n_sample, n_src))
    sd = blob_distribution(rng, c, nb=nb)
    s, C = fd.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 1, config=c)
    s0, C0 = orc.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 1, config=c)
    print(f"correlation nb={nb}: seis {rel_l2(s, s0):.2e} C {rel_l2(C, C0):.2e}")
    assert rel_l2(s, s0) < TOL_CORR
    assert rel_l2(C, C0) < TOL_FIELD


def _pipeline_inputs(orc, c, rng, n_rec=3):
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.45)
    rec = interior_points(rng, c, n_rec)
    G, Gout, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    spectrum = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
    sd = blob_distribution(rng, c)
    corr, C = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    stf = np.ascontiguousarray(corr[:, ::-1] / np.abs(corr).max()).astype(np.complex128)
    return mu, rho, src, rec, G, Gout, spectrum, sd, C, stf


@pytest.mark.parametrize("fw_nth", [1, 4])
def test_adjoint_structure_gradient_pair(orc, fw_nth):
    """The 'structure' branch of eval_objective_and_gradient.m:97-98: mode 1 against C, then mode 0
    with the distributed source stf_fft against G (half length)."""
    c = tiny(fw_nth=fw_nth)
    rng = np.random.default_rng(300 + fw_nth)
    mu, rho, src, rec, G, Gout, spectrum, sd, C, stf = _pipeline_inputs(orc, c, rng)
    K1, a1, _, s1 = fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, None, None, None, 1, config=c)
    K10, a10, _, s10 = orc.run_noise_adjoint(mu, rho, C, stf, rec, rec, None, None, None, 1, config=c)
    K2, _, _, s2 = fd.run_noise_adjoint(mu, rho, Gout, a10, rec, rec, spectrum, sd, None, 0, config=c)
    K20, _, _, s20 = orc.run_noise_adjoint(mu, rho, Gout, a10, rec, rec, spectrum, sd, None, 0, config=c)
    print(f"adjoint f={fw_nth}: K1 {rel_l2(K1[:, :, -1], K10[:, :, -1]):.2e} stf_fft {rel_l2(a1, a10):.2e} "
          f"K2 {rel_l2(K2[:, :, -1], K20[:, :, -1]):.2e} seis {rel_l2(s1, s10):.2e} {rel_l2(s2, s20):.2e}")
    assert rel_l2(K1[:, :, -1], K10[:, :, -1]) < TOL_KERNEL
    assert rel_l2(K2[:, :, -1], K20[:, :, -1]) < TOL_KERNEL
    assert rel_l2(a1, a10) < TOL_FIELD
    assert rel_l2(s1, s10) < TOL_CORR and rel_l2(s2, s20) < TOL_CORR
    assert np.all(s2[:, c.nt:] == 0)


def test_many_frequencies(orc):
    """181 frequencies (0.02:0.001:0.2): 2K = 362 phase-table columns padded to 512, so the DFT
    flushes run two column groups and the synthesis contracts over 32 chunks of spectrum rows.
    Green spectrum, correlation (synthesised source), adjoint mode 1 (stf_fft) and mode 0 (source
    from stf_fft) against the oracle."""
    from fd2d_noise_amd.config import matlab_colon
    c = tiny(f_sample=matlab_colon(0.02, 0.001, 0.2), fw_nth=2)
    assert c.n_sample == 181
    rng = np.random.default_rng(450)
    mu, rho, src, rec, G0, Gout0, spectrum, sd, C0, stf = _pipeline_inputs(orc, c, rng)
    G, Gout, _ = fd.run_forward1_green(mu, rho, src, rec, 1, config=c)
    s, C = fd.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 1, config=c)
    s0, _ = orc.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 1, config=c)
    K1, a1, _, _ = fd.run_noise_adjoint(mu, rho, C0, stf, rec, rec, None, None, None, 1, config=c)
    K10, a10, _, _ = orc.run_noise_adjoint(mu, rho, C0, stf, rec, rec, None, None, None, 1, config=c)
    K2, _, _, s2 = fd.run_noise_adjoint(mu, rho, Gout0, a10, rec, rec, spectrum, sd, None, 0, config=c)
    K20, _, _, s20 = orc.run_noise_adjoint(mu, rho, Gout0, a10, rec, rec, spectrum, sd, None, 0, config=c)
    e = {"G_fft": rel_l2(G, G0), "correlation": rel_l2(s, s0), "C": rel_l2(C, C0), "stf_fft": rel_l2(a1, a10),
         "K1": rel_l2(K1[:, :, -1], K10[:, :, -1]), "K2": rel_l2(K2[:, :, -1], K20[:, :, -1]),
         "adjoint seis": rel_l2(s2, s20)}
    print("181 frequencies: " + " ".join(f"{k} {v:.2e}" for k, v in e.items()))
    assert e["G_fft"] < TOL_FIELD and e["C"] < TOL_FIELD and e["stf_fft"] < TOL_FIELD
    assert e["correlation"] < TOL_CORR and e["adjoint seis"] < TOL_CORR
    assert e["K1"] < TOL_KERNEL and e["K2"] < TOL_KERNEL


@pytest.mark.parametrize("nb", [0, 4])
def test_adjoint_source_kernel(orc, nb):
    """Source kernel with point adjoint sources (eval_objective_and_gradient.m:91, mode 0)."""
    c = tiny(n_basis_fct=nb)
    rng = np.random.default_rng(400 + nb)
    mu, rho, src, rec, G, Gout, spectrum, sd, C, stf = _pipeline_inputs(orc, c.replace(n_basis_fct=0), rng)
    K, _, _, s = fd.run_noise_adjoint(mu, rho, None, stf, rec, rec, spectrum, None, G, 0, config=c)
    K0, _, _, s0 = orc.run_noise_adjoint(mu, rho, None, stf, rec, rec, spectrum, None, G, 0, config=c)
    nbc = max(nb, 1)
    print(f"source kernel nb={nb}: {rel_l2(K[:, :, :nbc], K0[:, :, :nbc]):.2e}")
    assert rel_l2(K[:, :, :nbc], K0[:, :, :nbc]) < TOL_KERNEL
    assert np.all(K[:, :, -1] == 0)


@pytest.mark.parametrize("mode", [10, 11, 12, 13])
def test_adjoint_hessian_modes(orc, mode):
    c = tiny(fw_nth=1)
    rng = np.random.default_rng(500 + mode)
    mu, rho, src, rec, G, Gout, spectrum, sd, C, stf = _pipeline_inputs(orc, c, rng)
    K, a, U, s = fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, mode, config=c)
    K0, a0, U0, s0 = orc.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, mode, config=c)
    print(f"adjoint mode={mode}: K {rel_l2(K, K0):.2e} seis {rel_l2(s, s0):.2e}")
    assert rel_l2(K[:, :, 0], K0[:, :, 0]) < TOL_KERNEL
    assert rel_l2(K[:, :, 1], K0[:, :, 1]) < TOL_KERNEL
    assert rel_l2(s, s0) < TOL_CORR
    if mode in (12, 13):
        assert rel_l2(a, a0) < TOL_FIELD
    else:
        assert np.all(a == 0)
    if mode in (11, 13):
        assert U.shape == U0.shape and rel_l2(U, U0) < TOL_FIELD


def test_born_runs(orc):
    c = tiny()
    rng = np.random.default_rng(600)
    mu, rho = hetero_model(rng, c)
    dmu = np.asfortranarray(1e8 * smooth_field(rng, c.nx, c.nz))
    src = interior_points(rng, c, 1, 0.45)
    rec = interior_points(rng, c, 3)
    _, Gout, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    Gb, _, sb = fd.run_forward1_green(mu, rho, src, rec, 0, dmu, Gout, config=c)
    Gb0, _, sb0 = orc.run_forward1_green(mu, rho, src, rec, 0, dmu, Gout, config=c)
    print(f"born green: seis {rel_l2(sb, sb0):.2e} G {rel_l2(Gb, Gb0):.2e}")
    assert rel_l2(sb, sb0) < TOL_KERNEL and rel_l2(Gb, Gb0) < TOL_KERNEL
    G, _, _ = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    spectrum = np.ones((c.n_sample, 1))
    sd = blob_distribution(rng, c)
    _, C = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    cb, _ = fd.run_forward2_correlation(mu, rho, Gb0, spectrum, sd, rec, 0, dmu, C, config=c)
    cb0, _ = orc.run_forward2_correlation(mu, rho, Gb0, spectrum, sd, rec, 0, dmu, C, config=c)
    print(f"born correlation: {rel_l2(cb, cb0):.2e}")
    assert rel_l2(cb, cb0) < TOL_KERNEL
    stf = np.ascontiguousarray(cb0[:, ::-1] / np.abs(cb0).max()).astype(np.complex128)
    _, _, U, _ = orc.run_noise_adjoint(mu, rho, None, stf, rec, rec, None, None, None, 11, config=c)
    Kb, _, _, ab = fd.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, 10, dmu, U, config=c)
    Kb0, _, _, ab0 = orc.run_noise_adjoint(mu, rho, C, stf, rec, rec, spectrum, None, G, 10, dmu, U, config=c)
    print(f"born adjoint: seis {rel_l2(ab, ab0):.2e} K {rel_l2(Kb, Kb0):.2e}")
    assert rel_l2(ab, ab0) < TOL_KERNEL and rel_l2(Kb, Kb0) < TOL_KERNEL


def test_config1_green_and_correlation(orc):
    """The reference-sized CPU case: 300 x 300 over 2000 km, order 4, dt = 0.5 s, nt = 2400,
    91 frequencies in 0.01-0.10 Hz, homogeneous model, 30-node taper, one reference station at the
    centre and two receivers 300-700 km away."""
    from fd2d_noise_amd.workloads import config1_workload
    w = config1_workload(seed=1)
    c = w["config"]
    G, _, sg = fd.run_forward1_green(w["mu"], w["rho"], w["src"], w["rec"], 0, config=c)
    G0, _, sg0 = orc.run_forward1_green(w["mu"], w["rho"], w["src"], w["rec"], 0, config=c)
    cc, _ = fd.run_forward2_correlation(w["mu"], w["rho"], G, w["spectrum"], w["source_dist"], w["rec"], 0, config=c)
    cc0, _ = orc.run_forward2_correlation(w["mu"], w["rho"], G0, w["spectrum"], w["source_dist"], w["rec"], 0,
                                          config=c)
    e = (rel_l2(sg, sg0), rel_l2(G, G0), rel_l2(cc, cc0))
    print(f"config1: green seis {e[0]:.2e} G_fft {e[1]:.2e} correlation {e[2]:.2e}")
    assert e[2] < TOL_CORR
    assert e[1] < TOL_FIELD


def test_errors_are_reported(orc):
    c = tiny(order=5)
    mu, rho = homog_model(tiny())
    with pytest.raises(abi.Fd2dError) as ei:
        fd.run_forward1_green(mu, rho, np.array([[3e4, 3e4]]), None, 0, config=c)
    assert ei.value.code == abi.FD2D_EUNSUPPORTED
    with pytest.raises(abi.Fd2dError) as ei:
        fd.run_forward1_green(mu[:-1], rho, np.array([[3e4, 3e4]]), None, 0, config=tiny())
    assert ei.value.code == abi.FD2D_EINVAL
    stf3 = np.zeros((50, 50, 121), np.complex128)
    with pytest.raises(abi.Fd2dError) as ei:
        fd.run_noise_adjoint(mu, rho, None, stf3, None, None, np.ones((121, 1)), np.ones((50, 50)), None, 1,
                             config=tiny())
    assert ei.value.code == abi.FD2D_EUNSUPPORTED


@pytest.mark.parametrize("order", [4, 8])
def test_edge_positions(orc, order):
    """Sources and receivers on the corners and edges of the grid, outside the domain (nearest node,
    code/run_forward1_green.m:45-60), on tile borders (x = 63, 64 nodes) and twice on the same node:
    Green, correlation and adjoint (mode 0, point adjoint sources) against the oracle."""
    d0 = tiny().Lx / (tiny().nx - 1)  # the tiny set-up's spacing, so its time step stays stable
    c = tiny(order=order, nx=130, nz=70, Lx=129 * d0, Lz=69 * d0)
    rng = np.random.default_rng(700 + order)
    mu, rho = hetero_model(rng, c)
    dx = c.Lx / (c.nx - 1)
    src = np.array([[0.0, 0.0], [63 * dx, 31 * dx], [63 * dx, 31 * dx]])
    rec = np.array([[c.Lx, c.Lz], [-5e4, 0.5 * c.Lz], [64 * dx, 32 * dx], [0.5 * c.Lx, 2 * c.Lz], [64 * dx, 32 * dx]])
    G, _, s = fd.run_forward1_green(mu, rho, src, rec, 0, config=c)
    G0, _, s0 = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    spectrum = rng.uniform(0.5, 1.5, (c.n_sample, 1))
    sd = blob_distribution(rng, c)
    corr, _ = fd.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 0, config=c)
    corr0, _ = orc.run_forward2_correlation(mu, rho, G0, spectrum, sd, rec, 0, config=c)
    stf = np.ascontiguousarray(corr0[:, ::-1] / np.abs(corr0).max()).astype(np.complex128)
    K, _, _, sa = fd.run_noise_adjoint(mu, rho, None, stf, rec, rec, spectrum, None, G0, 0, config=c)
    K0, _, _, sa0 = orc.run_noise_adjoint(mu, rho, None, stf, rec, rec, spectrum, None, G0, 0, config=c)
    e = (rel_l2(s, s0), rel_l2(G, G0), rel_l2(corr, corr0), rel_l2(sa, sa0), rel_l2(K[:, :, 0], K0[:, :, 0]))
    print(f"edge positions order={order}: green {e[0]:.2e} G_fft {e[1]:.2e} correlation {e[2]:.2e} "
          f"adjoint {e[3]:.2e} K_s {e[4]:.2e}")
    assert e[0] < TOL_FIELD and e[1] < TOL_FIELD and e[2] < TOL_CORR and e[3] < TOL_CORR and e[4] < TOL_KERNEL
    assert np.array_equal(s[2], s[4]) and np.array_equal(corr[2], corr[4])  # the same node twice


@pytest.mark.parametrize("order", [6, 8])
@pytest.mark.parametrize("mode,fw_nth,nt", [(1, 1, 150), (1, 3, 151), (0, 2, 150), (13, 1, 150), (11, 2, 150)])
def test_displacement_form(orc, order, mode, fw_nth, nt):
    """Correlation and adjoint runs at FD orders 6 and 8 step v and an FP64 displacement, with sigma
    evaluated from it (no sigma arrays) and each structure-kernel product computed in the next
    step's launch (the last step's in a launch of its own: nt = 151 gives 301 steps, a multiple of
    fw_nth = 3; the half-length mode-0 run ends on a product step at fw_nth = 2). Correlation traces
    and snapshots, adjoint traces, both kernels, stf_fft and u_out against the oracle."""
    d0 = tiny().Lx / (tiny().nx - 1)
    c = tiny(order=order, fw_nth=fw_nth, nt=nt, nx=70, nz=58, Lx=69 * d0, Lz=57 * d0)
    rng = np.random.default_rng(800 + 10 * order + mode)
    mu, rho = hetero_model(rng, c)
    src = interior_points(rng, c, 1, 0.45)
    rec = interior_points(rng, c, 3)
    G, Gout, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    spectrum = np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None]
    sd = blob_distribution(rng, c)
    s, C = fd.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    s0, C0 = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    stf = np.ascontiguousarray(s0[:, ::-1] / np.abs(s0).max()).astype(np.complex128)
    if mode == 0:
        _, a1, _, _ = orc.run_noise_adjoint(mu, rho, C0, stf, rec, rec, None, None, None, 1, config=c)
        args = (mu, rho, Gout, a1, rec, rec, spectrum, sd, None, 0)
    else:
        args = (mu, rho, C0, stf, rec, rec, spectrum, None, G, mode)
    K, a, U, sa = fd.run_noise_adjoint(*args, config=c)
    K0, a0, U0, sa0 = orc.run_noise_adjoint(*args, config=c)
    e = {"correlation": rel_l2(s, s0), "C_out": rel_l2(C, C0), "adjoint seis": rel_l2(sa, sa0),
         "K_mu": rel_l2(K[:, :, -1], K0[:, :, -1]), "K_s": rel_l2(K[:, :, 0], K0[:, :, 0])}
    if mode in (1, 13):
        e["stf_fft"] = rel_l2(a, a0)
    if mode in (11, 13):
        e["u_out"] = rel_l2(U, U0)
    print(f"displacement form order={order} mode={mode} f={fw_nth} nt={nt}: "
          + " ".join(f"{k} {v:.2e}" for k, v in e.items()))
    assert np.abs(K0[:, :, -1]).max() > 0
    assert e["correlation"] < TOL_CORR and e["adjoint seis"] < TOL_CORR
    assert e["C_out"] < TOL_FIELD and e.get("stf_fft", 0) < TOL_FIELD and e.get("u_out", 0) < TOL_FIELD
    assert e["K_mu"] < TOL_KERNEL and e["K_s"] < TOL_KERNEL

"""Pins of the oracle's three solvers, each against something the oracle does not compute itself.

Green run (code/run_forward1_green.m):
  * closed form: homogeneous 2-D SH line force F = 1e9*dx*dz with constant stf (:64) gives
    u(r,t) = int v dt = F/(2 pi mu) * acosh(t/tau), tau = r/c, before boundary arrivals;
  * the DFT at a receiver node equals the DFT of the receiver's trace (:176-198);
  * reciprocity in a heterogeneous medium (the staggered divergence is minus the transpose of the
    forward difference, div_s.m vs dx_v.m);
  * snapshots are v before the step's update (:133-137).
Adjoint run (code/run_noise_adjoint.m) against the Green run:
  * a constant 1e9 point stf gives u(n) = sum_{m<=n} v_green(m) dt.
Correlation run (code/run_forward2_correlation.m) against the adjoint run:
  * with G_fft supported at one node, the synthesised source is a point source there whose time
    function is computed here from the reference's formula (:195-216); the two runs must agree.
Adjoint run, further identities built from receiver traces:
  * distributed stf supported at one node == point stf (:255-268), including the half-length rule;
  * stf_fft at a receiver node == DFT of its trace (:335-346);
  * source kernel at a receiver node == sum_n Re(M_b(n) u(n)) from its trace and G_fft (:302-320);
  * structure kernel at a node == -fw_nth * sum D u . D u_fwd(reverse order) from the traces of the
    stencil neighbours (:324-331).
Born runs: the linearised runs match central finite differences of the full runs in mu.
"""
import numpy as np
import pytest

import oracle
from fd2d_noise_amd import Config

from helpers import blob_distribution, hetero_model, homog_model, smooth_field
from conftest import rel_l2

INV_SQRT_2PI = 1.0 / np.sqrt(2 * np.pi)


def small_cfg(**kw):
    base = dict(Lx=40e3, Lz=36e3, nx=41, nz=37, dt=0.09, nt=60, order=4, fw_nth=1,
                f_sample=np.linspace(0.1, 0.6, 12), freq_samp=5, absorb_width=4e3)
    base.update(kw)
    return Config(**base)


def pos(c, i, j):
    return np.array([[i * c.dx, j * c.dz]])


# ------------------------------------------------------------------------------------------------
# Green
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("order", [4, 8])
def test_green_closed_form(orc, order):
    c = Config(Lx=300e3, Lz=300e3, nx=151, nz=151, dt=0.2, nt=150, order=order, absorb_width=20e3,
               f_sample=np.array([0.05, 0.06]))
    mu, rho = homog_model(c)
    cvel = np.sqrt(4.8e10 / 3000.0)
    src = np.array([[150e3, 150e3]])
    rec = np.array([[190e3, 150e3], [150e3, 120e3], [174e3, 182e3]])
    _, _, seis = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    F = 1e9 * c.dx * c.dz
    r = np.hypot(rec[:, 0] - 150e3, rec[:, 1] - 150e3)
    t = np.arange(1, c.nt + 1) * c.dt
    u = np.cumsum(seis, axis=1) * c.dt
    for ir in range(3):
        tau = r[ir] / cvel
        m = (t > 1.5 * tau) & (t < 29.0)
        ana = F / (2 * np.pi * 4.8e10) * np.arccosh(t[m] / tau)
        err = np.max(np.abs(u[ir, m] - ana) / ana)
        assert err < 0.08, (ir, err)  # grid dispersion of the step-force response


def test_green_dft_matches_trace(orc):
    c = small_cfg()
    rng = np.random.default_rng(1)
    mu, rho = hetero_model(rng, c)
    rec = np.array([[13 * c.dx, 20 * c.dz], [30 * c.dx, 9 * c.dz]])
    G, _, seis = orc.run_forward1_green(mu, rho, pos(c, 20, 18), rec, 0, config=c)
    n = np.arange(1, c.nt + 1)
    samp = (n + c.nt - 1) % c.freq_samp == 0
    t = (n - 1) * c.dt
    for ir, (i, j) in enumerate([(13, 20), (30, 9)]):
        dft = np.conj(np.sum(seis[ir, samp, None] * INV_SQRT_2PI *
                             np.exp(-1j * c.w_sample[None, :] * t[samp, None]) * c.dt, axis=0))
        assert rel_l2(G[i, j, :], dft) < 1e-12


def test_green_reciprocity(orc):
    c = Config(Lx=120e3, Lz=120e3, nx=121, nz=121, dt=0.1, nt=60, order=4, absorb_width=2e3,
               f_sample=np.array([0.1, 0.2]))
    rng = np.random.default_rng(2)
    mu, rho = hetero_model(rng, c, amp=0.2)
    A = np.array([[55e3, 60e3]])
    B = np.array([[66e3, 52e3]])
    _, _, sAB = orc.run_forward1_green(mu, rho, A, B, 0, config=c)
    _, _, sBA = orc.run_forward1_green(mu, rho, B, A, 0, config=c)
    assert np.abs(sAB).max() > 0
    assert rel_l2(sAB, sBA) < 1e-10


@pytest.mark.parametrize("fw_nth", [1, 3])
def test_green_snapshots_are_pre_update_v(orc, fw_nth):
    c = small_cfg(fw_nth=fw_nth)
    mu, rho = homog_model(c)
    rec = pos(c, 25, 22)
    _, Gout, seis = orc.run_forward1_green(mu, rho, pos(c, 20, 18), rec, 1, config=c)
    steps = [n for n in range(1, c.nt + 1) if (n + c.nt - 1) % fw_nth == 0]
    assert Gout.shape == (c.nx, c.nz, len(steps)) == (c.nx, c.nz, c.n_fw_green())
    prev = np.concatenate([[0.0], seis[0, :-1]])
    for q, n in enumerate(steps):
        assert Gout[25, 22, q] == np.float32(prev[n - 1])


# ------------------------------------------------------------------------------------------------
# Adjoint point source vs Green
# ------------------------------------------------------------------------------------------------
def test_adjoint_point_source_integrates_green(orc):
    c = small_cfg()
    rng = np.random.default_rng(3)
    mu, rho = hetero_model(rng, c)
    src = pos(c, 20, 18)
    rec = np.array([[12 * c.dx, 20 * c.dz], [28 * c.dx, 11 * c.dz]])
    _, _, vg = orc.run_forward1_green(mu, rho, src, rec, 0, config=c)
    stf = np.full((1, c.n_corr), 1e9 + 0j)
    _, _, _, ua = orc.run_noise_adjoint(mu, rho, None, stf, src, rec, None, None, None, 10, config=c)
    assert rel_l2(ua[:, :c.nt], np.cumsum(vg, axis=1) * c.dt) < 1e-12


# ------------------------------------------------------------------------------------------------
# Correlation vs adjoint point source
# ------------------------------------------------------------------------------------------------
def synth_point_stf(c, coefs, nsteps):
    """q(n) = sum_k Re(coefs_k (2pi)^-1/2 exp(i w_k t_n) dw) at n % freq_samp == 0, else 0."""
    q = np.zeros(c.n_corr, dtype=np.complex128)
    t = c.t_corr()
    for n in range(1, nsteps + 1):
        if n % c.freq_samp == 0:
            q[n - 1] = np.sum(np.real(coefs * INV_SQRT_2PI * np.exp(1j * c.w_sample * t[n - 1]) * c.dw))
    return q[None, :]


@pytest.mark.parametrize("nb,n_src", [(0, 1), (3, 2)])
def test_correlation_equals_point_source_run(orc, nb, n_src):
    c = small_cfg(n_basis_fct=nb, fw_nth=2)
    rng = np.random.default_rng(4 + nb)
    mu, rho = hetero_model(rng, c)
    K = c.n_sample
    i0, j0 = 19, 17
    G = np.zeros((c.nx, c.nz, K), np.complex128, order="F")
    G[i0, j0, :] = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    spectrum = rng.uniform(0.5, 1.5, (K, n_src))
    sd = blob_distribution(rng, c, nb=nb)
    rec = np.array([[12 * c.dx, 20 * c.dz], [28 * c.dx, 11 * c.dz], [i0 * c.dx, j0 * c.dz]])
    seis_c, Cout = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    I = oracle.primitives.interpol_matrix(K, nb)
    sdv = sd[i0, j0] if nb == 0 else sd[i0, j0, :]
    dist = I @ np.atleast_1d(sdv)
    coefs = spectrum.sum(axis=1) * dist * G[i0, j0, :]
    stf = synth_point_stf(c, coefs, c.n_corr)
    _, _, uout, seis_a = orc.run_noise_adjoint(mu, rho, None, stf, pos(c, i0, j0), rec, None, None, None, 11,
                                               config=c)
    assert np.abs(seis_c).max() > 0
    assert rel_l2(seis_c, seis_a) < 1e-11
    assert Cout.shape == uout.shape == (c.nx, c.nz, c.n_fw_corr())
    assert rel_l2(Cout, uout) < 1e-6


# ------------------------------------------------------------------------------------------------
# Adjoint identities
# ------------------------------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [10, 0])
def test_adjoint_distributed_source_equals_point_source(orc, mode):
    c = small_cfg()
    rng = np.random.default_rng(5)
    mu, rho = hetero_model(rng, c)
    K = c.n_sample
    i0, j0 = 22, 15
    stf3 = np.zeros((c.nx, c.nz, K), np.complex128, order="F")
    stf3[i0, j0, :] = rng.standard_normal(K) + 1j * rng.standard_normal(K)
    spectrum = rng.uniform(0.5, 1.5, (K, 1))
    sd = blob_distribution(rng, c)
    rec = np.array([[12 * c.dx, 20 * c.dz], [i0 * c.dx, j0 * c.dz]])
    _, _, _, sd_run = orc.run_noise_adjoint(mu, rho, None, stf3, None, rec, spectrum, sd, None, mode, config=c)
    nsteps = c.nt if mode == 0 else c.n_corr  # run_noise_adjoint.m:165-167
    coefs = spectrum[:, 0] * sd[i0, j0] * np.conj(stf3[i0, j0, :])
    stf = synth_point_stf(c, coefs, nsteps)
    _, _, _, pt_run = orc.run_noise_adjoint(mu, rho, None, stf, pos(c, i0, j0), rec, None, None, None, 10, config=c)
    assert np.all(sd_run[:, nsteps:] == 0)
    assert rel_l2(sd_run[:, :nsteps], pt_run[:, :nsteps]) < 1e-11


def test_adjoint_stf_fft_matches_trace(orc):
    c = small_cfg()
    rng = np.random.default_rng(6)
    mu, rho = hetero_model(rng, c)
    stf = (rng.standard_normal((2, c.n_corr)) * 1e9).astype(np.complex128)
    src = np.array([[15 * c.dx, 15 * c.dz], [25 * c.dx, 20 * c.dz]])
    rec = np.array([[18 * c.dx, 24 * c.dz], [30 * c.dx, 12 * c.dz]])
    _, sfft, _, seis = orc.run_noise_adjoint(mu, rho, None, stf, src, rec, None, None, None, 1, config=c)
    n = np.arange(1, c.n_corr + 1)
    samp = n % c.freq_samp == 0
    t = c.t_corr()
    for ir, (i, j) in enumerate([(18, 24), (30, 12)]):
        dft = np.sum(seis[ir, samp, None] * INV_SQRT_2PI * np.exp(-1j * c.w_sample[None, :] * t[samp, None]) * c.dt,
                     axis=0)
        assert rel_l2(sfft[i, j, :], dft) < 1e-12


@pytest.mark.parametrize("nb", [0, 3])
def test_adjoint_source_kernel_from_trace(orc, nb):
    c = small_cfg(n_basis_fct=nb)
    rng = np.random.default_rng(7 + nb)
    mu, rho = hetero_model(rng, c)
    K = c.n_sample
    G = (rng.standard_normal((c.nx, c.nz, K)) + 1j * rng.standard_normal((c.nx, c.nz, K)))
    G = np.asfortranarray(G)
    spectrum = rng.uniform(0.5, 1.5, (K, 1))
    stf = (rng.standard_normal((1, c.n_corr)) * 1e9).astype(np.complex128)
    nodes = [(18, 24), (30, 12), (20, 20)]
    rec = np.array([[i * c.dx, j * c.dz] for i, j in nodes])
    Kout, _, _, seis = orc.run_noise_adjoint(mu, rho, None, stf, pos(c, 15, 15), rec, spectrum, None, G, 10,
                                             config=c)
    I = oracle.primitives.interpol_matrix(K, nb)
    t = c.t_corr()
    nbc = max(nb, 1)
    assert Kout.shape == (c.nx, c.nz, nbc + 1)
    for ir, (i, j) in enumerate(nodes):
        ks = np.zeros(nbc)
        for n in range(1, c.n_corr + 1):
            if n % c.freq_samp:
                continue
            w = spectrum[:, 0] * np.conj(G[i, j, :]) * INV_SQRT_2PI * np.exp(1j * c.w_sample * t[n - 1]) * c.dw
            ks += np.real(I.T @ w) * seis[ir, n - 1]
        assert rel_l2(Kout[i, j, :nbc], ks) < 1e-11


@pytest.mark.parametrize("order,fw_nth", [(4, 2), (8, 3), (2, 1)])
def test_adjoint_structure_kernel_from_traces(orc, order, fw_nth):
    c = small_cfg(order=order, fw_nth=fw_nth)
    rng = np.random.default_rng(8 + order)
    mu, rho = hetero_model(rng, c)
    r = order // 2
    coef = {1: [1.0], 2: [9 / 8, -1 / 24], 3: [75 / 64, -25 / 384, 3 / 640],
            4: [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168]}[r]
    i0, j0 = 19, 16
    offs = list(range(-r + 1, r + 1))
    nodes = [(i0 + a, j0) for a in offs] + [(i0, j0 + b) for b in offs if b != 0]
    rec = np.array([[i * c.dx, j * c.dz] for i, j in nodes])
    stf = (rng.standard_normal((1, c.n_corr)) * 1e9).astype(np.complex128)
    n_stored = c.n_corr // fw_nth
    n_ufwd = n_stored - 3  # fewer snapshots than stored steps: the gate of :324
    ufwd = np.asfortranarray(rng.standard_normal((c.nx, c.nz, n_ufwd)).astype(np.float32))
    Kout, _, _, seis = orc.run_noise_adjoint(mu, rho, ufwd, stf, pos(c, 25, 20), rec, None, None, None, 10,
                                             config=c)
    tr = {nd: seis[q] for q, nd in enumerate(nodes)}

    def dxu(n):
        return sum(cm * (tr[(i0 + m, j0)][n - 1] - tr[(i0 - m + 1, j0)][n - 1]) for m, cm in enumerate(coef, 1)) / c.dx

    def dzu(n):
        return sum(cm * (tr[(i0, j0 + m)][n - 1] - tr[(i0, j0 - m + 1)][n - 1]) for m, cm in enumerate(coef, 1)) / c.dz

    f64 = ufwd.astype(np.float64)
    kmu, used = 0.0, 0
    for n in range(1, c.n_corr + 1):
        if n % fw_nth or used >= n_ufwd:
            continue
        s = f64[:, :, n_ufwd - 1 - used]
        ex = sum(cm * (s[i0 + m, j0] - s[i0 - m + 1, j0]) for m, cm in enumerate(coef, 1)) / c.dx
        ez = sum(cm * (s[i0, j0 + m] - s[i0, j0 - m + 1]) for m, cm in enumerate(coef, 1)) / c.dz
        kmu -= (dxu(n) * ex + dzu(n) * ez) * fw_nth
        used += 1
    assert abs(Kout[i0, j0, -1] - kmu) < 1e-5 * abs(kmu)


# ------------------------------------------------------------------------------------------------
# Born runs
# ------------------------------------------------------------------------------------------------
def test_green_born_is_linearisation(orc):
    c = small_cfg()
    rng = np.random.default_rng(9)
    mu, rho = hetero_model(rng, c)
    dmu = np.asfortranarray(1e8 * smooth_field(rng, c.nx, c.nz))
    src = pos(c, 20, 18)
    rec = np.array([[12 * c.dx, 20 * c.dz], [28 * c.dx, 11 * c.dz]])
    _, Gout, _ = orc.run_forward1_green(mu, rho, src, rec, 1, config=c)
    Gb, _, sb = orc.run_forward1_green(mu, rho, src, rec, 0, dmu, Gout, config=c)
    h = 1e-2
    Gp, _, sp = orc.run_forward1_green(mu + h * dmu, rho, src, rec, 0, config=c)
    Gm, _, sm = orc.run_forward1_green(mu - h * dmu, rho, src, rec, 0, config=c)
    assert rel_l2(sb, (sp - sm) / (2 * h)) < 1e-4
    assert rel_l2(Gb, (Gp - Gm) / (2 * h)) < 1e-4


def test_correlation_born_is_linearisation(orc):
    c = small_cfg()
    rng = np.random.default_rng(10)
    mu, rho = hetero_model(rng, c)
    K = c.n_sample
    G = np.asfortranarray(rng.standard_normal((c.nx, c.nz, K)) + 1j * rng.standard_normal((c.nx, c.nz, K)))
    spectrum = rng.uniform(0.5, 1.5, (K, 1))
    sd = blob_distribution(rng, c)
    dmu = np.asfortranarray(1e8 * smooth_field(rng, c.nx, c.nz))
    rec = np.array([[12 * c.dx, 20 * c.dz], [28 * c.dx, 11 * c.dz]])
    _, C = orc.run_forward2_correlation(mu, rho, G, spectrum, sd, rec, 1, config=c)
    sb, _ = orc.run_forward2_correlation(mu, rho, 0 * G, spectrum, sd, rec, 0, dmu, C, config=c)
    h = 1e-2
    sp, _ = orc.run_forward2_correlation(mu + h * dmu, rho, G, spectrum, sd, rec, 0, config=c)
    sm, _ = orc.run_forward2_correlation(mu - h * dmu, rho, G, spectrum, sd, rec, 0, config=c)
    assert rel_l2(sb, (sp - sm) / (2 * h)) < 1e-3


def test_adjoint_born_is_linearisation(orc):
    c = small_cfg()
    rng = np.random.default_rng(11)
    mu, rho = hetero_model(rng, c)
    dmu = np.asfortranarray(1e8 * smooth_field(rng, c.nx, c.nz))
    stf = (rng.standard_normal((1, c.n_corr)) * 1e9).astype(np.complex128)
    src = pos(c, 20, 18)
    rec = np.array([[12 * c.dx, 20 * c.dz], [28 * c.dx, 11 * c.dz]])
    _, _, U, _ = orc.run_noise_adjoint(mu, rho, None, stf, src, rec, None, None, None, 11, config=c)
    assert U.shape[2] == c.n_corr
    _, _, _, sb = orc.run_noise_adjoint(mu, rho, None, None, src, rec, None, None, None, 10, dmu, U, config=c)
    h = 1e-2
    _, _, _, sp = orc.run_noise_adjoint(mu + h * dmu, rho, None, stf, src, rec, None, None, None, 10, config=c)
    _, _, _, sm = orc.run_noise_adjoint(mu - h * dmu, rho, None, stf, src, rec, None, None, None, 10, config=c)
    assert rel_l2(sb, (sp - sm) / (2 * h)) < 1e-3
    # u_in with a different number of slices disables the Born term (run_noise_adjoint.m:285)
    _, _, _, s0 = orc.run_noise_adjoint(mu, rho, None, None, src, rec, None, None, None, 10, dmu, U[:, :, :-1],
                                        config=c)
    assert np.all(s0 == 0)

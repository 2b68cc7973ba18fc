"""Closed-form checks of the input builders (fd2d_noise_amd/inputs.py: make_noise_source.m,
define_material_parameters.m, define_computational_domain.m) and of the general_source oracle."""
import numpy as np
import pytest

from fd2d_noise_amd import Config
from fd2d_noise_amd import inputs
from oracle import inversion as OI


def test_spectrum_closed_form():
    c = Config(f_sample=np.array([1 / 11, 1 / 7, 0.0]))
    _, spec = inputs.make_noise_source(c, "homogeneous", 0)
    # strength * exp(-(|f| - f_peak)^2 / bandwidth^2): exactly the strength at the peak
    assert spec.shape == (3, 1)
    assert spec[0, 0] == pytest.approx(0.3)
    assert spec[2, 0] == pytest.approx(0.3 * np.exp(-(1 / 11) ** 2 / 0.035 ** 2))
    _, spec2 = inputs.make_noise_source(c, "homogeneous", 0, n_noise_sources=2)
    assert spec2[1, 1] == pytest.approx(1.0)


def test_gaussian_blob_orientation_and_width():
    c = Config(Lx=2.0e5, Lz=2.0e5, nx=41, nz=41)  # 'small' set-up, dx = 5 km
    dist, _ = inputs.make_noise_source(c, "gaussian", 0)
    # blob at x = 1e5 (i = 20), z = 1.6e5 (j = 32), width 4e4, magnitude 6: exp(-r^2/width^2)
    assert dist.shape == (41, 41, 1)
    assert dist[20, 32, 0] == pytest.approx(7.0)
    assert dist[28, 32, 0] == pytest.approx(1 + 6 * np.exp(-1.0))  # 40 km along x
    assert dist[20, 24, 0] == pytest.approx(1 + 6 * np.exp(-1.0))  # 40 km along z
    assert dist[32, 20, 0] < 1.2


def test_homogeneous_ring_picture_and_errors():
    c = Config()
    dist, _ = inputs.make_noise_source(c, "homogeneous", 0)
    assert np.all(dist == 1.0)
    with pytest.raises(ValueError, match="point"):
        inputs.make_noise_source(c, "point", 0)
    big = Config(Lx=2.0e6, Lz=2.0e6, nx=101, nz=101)
    ring, _ = inputs.make_noise_source(big, "ring", 0)
    X, Z, _, _, _, _ = inputs.define_computational_domain(big.Lx, big.Lz, 101, 101)
    R = np.sqrt((X - 1e6) ** 2 + (Z - 1e6) ** 2)
    assert np.all(ring[:, :, 0][(R.T < 6.3e5) | (R.T > 7.3e5)] == 0)  # only the shell of width 1e5
    assert ring.max() > 1.0  # the tapered bright arc
    with pytest.raises(ValueError):
        inputs.make_noise_source(Config(nx=50, nz=40), "ring", 0)
    pic = np.full((c.nz, c.nx), 255.0)
    pic[0, 3] = 0.0  # image row 1 is the top: flipped to z = nz, transposed to (x, z)
    d, _ = inputs.make_noise_source(c, "picture", 0, picture=pic)
    assert d[3, c.nz - 1, 0] == 1.0 and np.count_nonzero(d) == 1


def test_material_models():
    c = Config(nx=400, nz=200)
    mu, rho = inputs.define_material_parameters(c, 1)
    assert np.all(mu == 4.8e10) and np.all(rho == 3000.0)
    mu, _ = inputs.define_material_parameters(c, 3)
    s = int(np.floor(400 / 2 * 1.25 + 0.5))  # round(250) = 250 (1-based)
    assert np.all(mu[s - 1:, :] == 5.6e10) and np.all(mu[:s - 1, :] == 4.8e10)
    mu, rho = inputs.define_material_parameters(c, 5)
    assert np.all(mu[:, 98] == 2.0e10) and np.all(mu[:, 99] == 2.0e10)  # k = 99, 100
    assert np.all(mu[:, 149] == 2.0e10 + 50 * 1e9) and np.all(mu[:, 199] == mu[:, 149])
    mu, rho = inputs.define_material_parameters(c, 6)
    assert rho[97, 122] == 5000.0 and rho[96, 122] == 3000.0
    mu, _ = inputs.define_material_parameters(c, 7)
    assert np.all(mu[:330] == 3.675e10) and np.all(mu[330:] == 2.7e10)
    sq = Config(Lx=2.0e6, Lz=2.0e6, nx=101, nz=101)
    mu, _ = inputs.define_material_parameters(sq, 2)
    assert mu[50, 50] == pytest.approx(4.8e10 - 8.0e9)
    mu, _ = inputs.define_material_parameters(sq, 999)
    # (-1)^i 1e10 exp(-(x-x_i)^2/wx^2) exp(-(z-z_i)^2/wz^2): i = 1 at x = 1.15e6 negative,
    # i = 2 at 1.45e6 positive, both at z = 1e6 (node j = 50); node i = 72 is x = 1.44e6
    x = 72 * 2e4
    expect = 4.8e10 - 1e10 * np.exp(-((x - 1.15e6) / 1.1e5) ** 2) + 1e10 * np.exp(-((x - 1.45e6) / 1.1e5) ** 2)
    assert mu[72, 50] == pytest.approx(expect, rel=1e-12)
    assert mu[72, 60] == pytest.approx(4.8e10 + (expect - 4.8e10) * np.exp(-(2e5 / 2.5e5) ** 2), rel=1e-12)
    with pytest.raises(ValueError):
        inputs.define_material_parameters(Config(), 4)  # rho(98:102, 123:127) outside 50 x 50
    with pytest.raises(ValueError, match="mu_"):
        inputs.define_material_parameters(Config(), 12345)


def test_general_source_oracle_pins():
    rng = np.random.default_rng(0)
    dist = rng.random((6, 5, 2))
    # constant spectrum s: every band map is s * sum_ns dist
    out = OI.general_source(np.full((10, 2), 0.5), dist, 3)
    for b in range(3):
        np.testing.assert_allclose(out[:, :, b], 0.5 * dist.sum(axis=2), rtol=1e-14)
    # band sizes of integration_limits(10, 3): 1..3, 4..6, 7..10; n_b-weighted bands sum to the total
    spec = rng.random((10, 2))
    out = OI.general_source(spec, dist, 3)
    total = np.einsum("kn,xzn->xz", spec, dist)
    np.testing.assert_allclose(3 * out[:, :, 0] + 3 * out[:, :, 1] + 4 * out[:, :, 2], total, rtol=1e-13)
    np.testing.assert_allclose(out[:, :, 0] * 3, np.einsum("kn,xzn->xz", spec[:3], dist), rtol=1e-13)

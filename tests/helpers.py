"""Seeded inputs shared by the oracle and parity tests."""
import numpy as np

from fd2d_noise_amd import Config


def smooth_field(rng, nx, nz, corr_cells=6.0):
    """Zero-mean smooth random field normalised to max |g| = 1 (Gaussian-filtered white noise)."""
    w = rng.standard_normal((nx, nz))
    kx = np.fft.fftfreq(nx)[:, None]
    kz = np.fft.fftfreq(nz)[None, :]
    filt = np.exp(-0.5 * (kx ** 2 + kz ** 2) * (2 * np.pi * corr_cells) ** 2)
    g = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
    g -= g.mean()
    return g / np.abs(g).max()


def hetero_model(rng, c: Config, amp=0.1):
    rho = 3000.0 * (1 + 0.05 * smooth_field(rng, c.nx, c.nz))
    vs = 4000.0 * (1 + amp * smooth_field(rng, c.nx, c.nz))
    mu = rho * vs ** 2
    return np.asfortranarray(mu), np.asfortranarray(rho)


def homog_model(c: Config):
    return 4.8e10 * np.ones((c.nx, c.nz), order="F"), 3000.0 * np.ones((c.nx, c.nz), order="F")


def gaussian_spectrum(c: Config, f0=None, sigma=None):
    f = np.asarray(c.f_sample)
    f0 = f.mean() if f0 is None else f0
    sigma = (f.max() - f.min()) / 4 if sigma is None else sigma
    return np.exp(-((f - f0) ** 2) / (2 * sigma ** 2))[:, None]


def blob_distribution(rng, c: Config, nb=0, sigma_frac=0.15, background=0.1):
    X = c.x()[:, None]
    Z = c.z()[None, :]
    n = max(nb, 1)
    out = np.zeros((c.nx, c.nz, n))
    for b in range(n):
        x0, z0 = rng.uniform(0.2, 0.8) * c.Lx, rng.uniform(0.2, 0.8) * c.Lz
        s = sigma_frac * c.Lx
        out[:, :, b] = background + np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2 * s ** 2))
    return np.asfortranarray(out if nb else out[:, :, 0])


def interior_points(rng, c: Config, n, margin_frac=0.25):
    x = rng.uniform(margin_frac, 1 - margin_frac, n) * c.Lx
    z = rng.uniform(margin_frac, 1 - margin_frac, n) * c.Lz
    return np.stack([x, z], axis=1)


def node(c: Config, p):
    """Nearest node (0-based) of a position, lowest index on ties."""
    return int(np.argmin(np.abs(c.x() - p[0]))), int(np.argmin(np.abs(c.z() - p[1])))

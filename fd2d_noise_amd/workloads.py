"""Seeded synthetic inputs of the benchmark and parity workloads.

config1: the reference-sized CPU case -- 2-D SH on 2000 x 2000 km, 300 x 300 nodes, order 4,
         dt = 0.5 s, nt = 2400, homogeneous rho = 3000 kg/m^3, mu = 4.8e10 Pa (4 km/s), a 30-node
         taper on all four sides, Gaussian noise spectrum centred at 0.05 Hz (sigma 0.02 Hz) on 91
         frequencies in 0.01-0.10 Hz, a Gaussian source blob (sigma 200 km) at a seeded location
         plus a 0.1 background, the reference station at the centre and 2 receivers 300-700 km away.
config1-stations: the config1 model with 64 reference stations seeded inside the absorbing frame,
         each correlated at the two config1 receivers (the batched form of the CPU-sized case).
config2: the single large solve -- 2400 x 2400 nodes on 2000 km (dx = 0.834 km), order 8,
         dt = 0.06 s, nt = 8000, vs = 4 km/s times (1 + 0.1 g) with g a seeded Gaussian random field
         (30 km correlation length, max |g| = 1), rho constant, forward fields stored every 4th step,
         the default 121 frequencies (0.02:0.0015:0.2 Hz), 16 receivers 400-800 km from the centre
         station and a cross-correlation traveltime misfit against the homogeneous model's traces.
config3: batched reference stations -- 1200 x 1200 nodes on 2000 km (dx = 1.67 km), order 4,
         dt = 0.12 s, nt = 4000, 128 frequencies (0.01-0.2 Hz, every 5th step), 4 source basis
         functions (seeded blobs, flat spectrum as after general_source), 64 reference stations
         seeded uniformly inside the absorbing frame and 32 receivers shared by all stations.
config4: inversion -- 600 x 600 nodes on 2000 km (dx = 3.34 km), order 4, dt = 0.25 s, nt = 3000, 91
         frequencies in 0.01-0.10 Hz, forward fields stored every step, a 20-station ring (radius
         600 km) whose stations are each other's receivers, a uniform noise source, and a true model
         with 10 seeded Gaussian vs anomalies (radius 40-120 km, +-8%) inside the ring; the initial
         model is homogeneous (4 km/s).
serving: a Poisson stream of mixed requests -- 50% correlation, 35% correlation + structure kernel,
         15% structure Hessian-vector product -- on resident models of three grid sizes (200, 283 and
         400 nodes square, log-spaced; dx = 2 km, nt = 10 nx) at FD order 4 or 8, vs = 4 km/s times
         (1 + 0.05 g) with seeded random fields g, forward fields stored every step (the HVP needs
         it), 8 candidate reference stations per model and 8 receivers per request. The full
         range: seven sizes from 200 to 4000 nodes square (SERVING_RANGE_SIZES, log-spaced), each
         request's size drawn log-uniform in [200, 4000] and sent to the nearest model.
The taper is the reference's Gaussian taper (init_absbound.m:18).
"""
from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .config import Config, matlab_colon

MU0, RHO0 = 4.8e10, 3000.0


def gaussian_random_field(rng, nx, nz, dx, corr_len):
    """Zero-mean field with Gaussian correlation of length corr_len, normalised to max |g| = 1."""
    w = rng.standard_normal((nx, nz))
    kx = 2 * np.pi * np.fft.rfftfreq(nz, d=dx)[None, :]
    kz = 2 * np.pi * np.fft.fftfreq(nx, d=dx)[:, None]
    filt = np.exp(-(kx ** 2 + kz ** 2) * corr_len ** 2 / 8.0)
    g = np.fft.irfft2(np.fft.rfft2(w) * filt, s=(nx, nz))
    g -= g.mean()
    return g / np.abs(g).max()


def _ring(rng, centre, n, rmin, rmax):
    ang = rng.uniform(0, 2 * np.pi, n)
    r = rng.uniform(rmin, rmax, n)
    return np.stack([centre[0] + r * np.cos(ang), centre[1] + r * np.sin(ang)], axis=1)


def _blob(rng, c: Config, sigma=200e3, background=0.1):
    x0 = rng.uniform(0.2, 0.8) * c.Lx
    z0 = rng.uniform(0.2, 0.8) * c.Lz
    X = c.x()[:, None]
    Z = c.z()[None, :]
    return np.asfortranarray(background + np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2 * sigma ** 2)))


def config1_workload(seed: int = 1) -> dict:
    rng = np.random.default_rng(seed)
    c = Config(Lx=2.0e6, Lz=2.0e6, nx=300, nz=300, dt=0.5, nt=2400, order=4, fw_nth=1,
               f_sample=np.linspace(0.01, 0.10, 91), freq_samp=5)
    c.absorb_width = 30 * c.dx
    f = np.asarray(c.f_sample)
    centre = np.array([c.Lx / 2, c.Lz / 2])
    return dict(
        name="config1", config=c,
        mu=MU0 * np.ones((c.nx, c.nz), order="F"), rho=RHO0 * np.ones((c.nx, c.nz), order="F"),
        src=centre[None, :], rec=_ring(rng, centre, 2, 300e3, 700e3),
        spectrum=np.exp(-((f - 0.05) ** 2) / (2 * 0.02 ** 2))[:, None],
        source_dist=_blob(rng, c))


def config1_stations_workload(seed: int = 1, n_stations: int = 64) -> dict:
    """The config1 model, source and receivers with n_stations reference stations seeded uniformly
    inside the absorbing frame (the batched form of the reference-sized case)."""
    w = config1_workload(seed)
    c = w["config"]
    rng = np.random.default_rng(seed + 100)
    a = c.absorb_width
    w["stations"] = np.stack([rng.uniform(a, c.Lx - a, n_stations), rng.uniform(a, c.Lz - a, n_stations)], axis=1)
    w["name"] = "config1-stations"
    return w


def config2_workload(seed: int = 2, nt: int = 8000) -> dict:
    rng = np.random.default_rng(seed)
    c = Config(Lx=2.0e6, Lz=2.0e6, nx=2400, nz=2400, dt=0.06, nt=nt, order=8, fw_nth=4,
               f_sample=matlab_colon(0.02, 0.0015, 0.2), freq_samp=5)
    c.absorb_width = 30 * c.dx
    g = gaussian_random_field(rng, c.nx, c.nz, c.dx, 30e3)
    vs = 4000.0 * (1 + 0.1 * g)
    rho = RHO0 * np.ones((c.nx, c.nz), order="F")
    f = np.asarray(c.f_sample)
    centre = np.array([c.Lx / 2, c.Lz / 2])
    return dict(
        name="config2", config=c,
        mu=np.asfortranarray(rho * vs ** 2), rho=rho,
        mu_obs=MU0 * np.ones((c.nx, c.nz), order="F"),
        src=centre[None, :], rec=_ring(rng, centre, 16, 400e3, 800e3),
        spectrum=np.exp(-((f - 0.1) ** 2) / (2 * 0.03 ** 2))[:, None],
        source_dist=_blob(rng, c))


def config3_workload(seed: int = 3, nt: int = 4000, n_stations: int = 64) -> dict:
    rng = np.random.default_rng(seed)
    c = Config(Lx=2.0e6, Lz=2.0e6, nx=1200, nz=1200, dt=0.12, nt=nt, order=4, fw_nth=4, n_basis_fct=4,
               f_sample=np.linspace(0.01, 0.2, 128), freq_samp=5)
    c.absorb_width = 30 * c.dx
    w = c.absorb_width
    inside = lambda n: np.stack([rng.uniform(w, c.Lx - w, n), rng.uniform(w, c.Lz - w, n)], axis=1)
    stations = inside(n_stations)
    rec = inside(32)
    basis = np.stack([_blob(rng, c, sigma=250e3, background=0.05) for _ in range(c.n_basis_fct)], axis=2)
    return dict(
        name="config3", config=c,
        mu=MU0 * np.ones((c.nx, c.nz), order="F"), rho=RHO0 * np.ones((c.nx, c.nz), order="F"),
        stations=stations, rec=rec, spectrum=np.ones((c.n_sample, 1)),
        source_dist=np.asfortranarray(basis))


def serving_models(seed: int = 5, sizes=(200, 283, 400), orders=(4, 8)) -> dict:
    """Resident models of the serving workload: name -> dict(config, mu, rho, spectrum, source_dist,
    stations)."""
    rng = np.random.default_rng(seed)
    out = {}
    for n in sizes:
        for order in orders:
            dx = 2000.0
            c = Config(Lx=dx * (n - 1), Lz=dx * (n - 1), nx=n, nz=n, dt=0.1, nt=10 * n, order=order, fw_nth=1,
                       freq_samp=5)
            c.absorb_width = 20 * c.dx
            g = gaussian_random_field(rng, n, n, c.dx, 30e3)
            vs = 4000.0 * (1 + 0.05 * g)
            rho = RHO0 * np.ones((n, n), order="F")
            f = np.asarray(c.f_sample)
            w = c.absorb_width
            st = np.stack([rng.uniform(w, c.Lx - w, 8), rng.uniform(w, c.Lz - w, 8)], axis=1)
            out[f"n{n}_o{order}"] = dict(config=c, mu=np.asfortranarray(rho * vs ** 2), rho=rho,
                                         spectrum=np.exp(-((f - 0.1) ** 2) / (2 * 0.03 ** 2))[:, None],
                                         source_dist=_blob(rng, c, sigma=0.2 * c.Lx), stations=st)
    return out


# grid sizes of the full serving range: 200 to 4000 points a side, log-spaced (200 * 20^(k/6))
SERVING_RANGE_SIZES = (200, 330, 543, 894, 1474, 2428, 4000)


def serving_stream(models: dict, rate: float, n_requests: int, seed: int = 6, mix=(0.5, 0.35, 0.15),
                   log_uniform: Optional[Tuple[float, float]] = None):
    """(arrival offset [s], kind, model name, src (1 x 2), rec (8 x 2), seed) tuples of a Poisson
    stream. The model of a request is drawn uniformly from `models`; with log_uniform = (lo, hi)
    its grid size is drawn log-uniform in [lo, hi] instead and the request goes to the model of the
    nearest size (in log distance; orders drawn uniformly)."""
    rng = np.random.default_rng(seed)
    names = sorted(models)
    t = np.cumsum(rng.exponential(1.0 / rate, n_requests))
    kinds = rng.choice(["correlation", "kernel", "hvp"], size=n_requests, p=list(mix))
    sizes = np.array(sorted({models[n]["config"].nx for n in names}), dtype=np.float64)
    out = []
    for q in range(n_requests):
        if log_uniform is None:
            name = names[rng.integers(len(names))]
        else:
            nx = np.exp(rng.uniform(np.log(log_uniform[0]), np.log(log_uniform[1])))
            size = int(sizes[np.argmin(np.abs(np.log(sizes) - np.log(nx)))])
            cands = [n for n in names if models[n]["config"].nx == size]
            name = cands[rng.integers(len(cands))]
        m = models[name]
        c = m["config"]
        w = c.absorb_width
        src = m["stations"][rng.integers(len(m["stations"]))][None, :]
        rec = np.stack([rng.uniform(w, c.Lx - w, 8), rng.uniform(w, c.Lz - w, 8)], axis=1)
        out.append((float(t[q]), str(kinds[q]), name, src, rec, int(rng.integers(1 << 30))))
    return out


def config4_workload(seed: int = 4, nt: int = 3000, n_stations: int = 20) -> dict:
    rng = np.random.default_rng(seed)
    c = Config(Lx=2.0e6, Lz=2.0e6, nx=600, nz=600, dt=0.25, nt=nt, order=4, fw_nth=1,
               f_sample=np.linspace(0.01, 0.10, 91), freq_samp=5)
    c.absorb_width = 30 * c.dx
    centre = np.array([c.Lx / 2, c.Lz / 2])
    ang = 2 * np.pi * np.arange(n_stations) / n_stations
    ring = np.stack([centre[0] + 600e3 * np.cos(ang), centre[1] + 600e3 * np.sin(ang)], axis=1)
    X = c.x()[:, None]
    Z = c.z()[None, :]
    anomaly = np.zeros((c.nx, c.nz))
    for _ in range(10):
        r = rng.uniform(0, 450e3)
        a = rng.uniform(0, 2 * np.pi)
        rad = rng.uniform(40e3, 120e3)
        amp = rng.choice([-1.0, 1.0]) * 0.08
        x0, z0 = centre[0] + r * np.cos(a), centre[1] + r * np.sin(a)
        anomaly += amp * np.exp(-((X - x0) ** 2 + (Z - z0) ** 2) / (2 * rad ** 2))
    vs_true = 4000.0 * (1 + anomaly)
    rho = RHO0 * np.ones((c.nx, c.nz), order="F")
    f = np.asarray(c.f_sample)
    return dict(name="config4", config=c, rho=rho, mu_true=np.asfortranarray(rho * vs_true ** 2),
                mu_0=MU0, array=ring, ref_stat=ring.copy(),
                spectrum=np.exp(-((f - 0.05) ** 2) / (2 * 0.02 ** 2))[:, None],
                source_dist=np.ones((c.nx, c.nz), order="F"))

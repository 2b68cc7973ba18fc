"""Inputs of the solvers: computational domain, noise sources and material models.

Host-side mirrors of
  code/propagation/define_computational_domain.m:1-24      ``define_computational_domain``
  input/make_noise_source.m:1-303                          ``make_noise_source``
  code/propagation/general_source.m:2-27                   ``general_source`` (HIP kernel)
  input/define_material_parameters.m:1-231                 ``define_material_parameters``
with the parameters the reference hard-codes in those files. The reference reads nx, nz, Lx, Lz and
the frequency axis from input_parameters() / input_interferometry(); here they come with a Config.
Arrays are MATLAB-shaped numpy arrays (nx x nz [x n], column-major). MATLAB's meshgrid X, Z are
nz x nx; where the reference adds them to an nx x nz array without a transpose (the 'ring' source,
material model 2) MATLAB needs nx == nz and adds the transposed field, which is kept here.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import numpy as np

from . import _lib, abi
from .config import Config


def define_computational_domain(Lx: float, Lz: float, nx: int, nz: int):
    """[X, Z, x, z, dx, dz] (define_computational_domain.m:48-54); X, Z are nz x nx."""
    dx = Lx / (nx - 1)
    dz = Lz / (nz - 1)
    x = np.arange(nx) * dx
    z = np.arange(nz) * dz
    X, Z = np.meshgrid(x, z)
    return X, Z, x, z, dx, dz


def _untransposed(field_nz_nx: np.ndarray, nx: int, nz: int, where: str) -> np.ndarray:
    if nx != nz:
        raise ValueError(f"{where}: the reference adds an nz x nx meshgrid field to an nx x nz array, "
                         f"which MATLAB only accepts for nx == nz")
    return np.asarray(field_nz_nx)


def general_source(noise_spectrum, noise_source_distribution, n_basis_fct: int) -> np.ndarray:
    """general_source.m:2-27 on the device (fd2d_general_source_device): nx x nz x n_basis_fct."""
    import torch
    spec = np.asfortranarray(np.atleast_2d(np.asarray(noise_spectrum, dtype=np.float64).T).T)
    K, n_src = spec.shape
    dist = np.asarray(noise_source_distribution, dtype=np.float64)
    if dist.ndim == 2:
        dist = dist[:, :, None]
    nx, nz = dist.shape[:2]
    if dist.shape[2] < n_src:
        raise ValueError("noise_source_distribution holds fewer maps than the spectrum has sources")
    P = nx * nz
    spec_d = torch.as_tensor(spec.ravel(order="F"), device="cuda")
    dist_d = torch.as_tensor(np.asfortranarray(dist).ravel(order="F"), device="cuda")
    out = torch.zeros(P * n_basis_fct, dtype=torch.float64, device="cuda")
    lib = _lib.load()
    f = lib.fd2d_general_source_device
    f.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32,
                  ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    st = f(spec_d.data_ptr(), K, n_src, dist_d.data_ptr(), P, int(n_basis_fct), out.data_ptr(),
           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())
    return out.cpu().numpy().reshape((nx, nz, n_basis_fct), order="F")


# make_noise_source.m:21-125: the user-input block of the reference
_SPECTRUM = dict(f_peak=(1 / 11, 1 / 7), bandwidth=(0.035, 0.025), strength=(0.3, 1.0))
_GAUSSIAN = {"small": dict(x=(1.0e5,), z=(1.6e5,), width=(4e4,), magnitude=(6.0,)),
             "big": dict(x=(0.6e6, 0.5e6, 1.3e6), z=(0.8e6, 1.5e6, 1.6e6), width=(1.5e5, 1.5e5, 1.8e5),
                         magnitude=(6.5, 5.0, 3.0))}
_RING = dict(x=1.0e6, z=1.0e6, radius=6.8e5, thickness=1e5, angle_cover=60.0, taper_width=20.0, taper_strength=100.0)


def make_noise_source(config: Config, source_type: str = "homogeneous", n_basis_fct: Optional[int] = None,
                      picture: Optional[np.ndarray] = None, n_noise_sources: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """[noise_source_distribution, noise_spectrum] = make_noise_source(source_type, n_basis_fct)
    (make_noise_source.m:1-230). source_type 'homogeneous', 'gaussian', 'ring' or 'picture' (the
    grey-scale image the reference reads from source.png, passed as an nz x nx array). The
    reference's 'point' set-up indexes past its one-element tables (:167) and raises here too."""
    c = config
    nb = c.n_basis_fct if n_basis_fct is None else int(n_basis_fct)
    f = np.asarray(c.f_sample, dtype=np.float64)
    nx, nz = c.nx, c.nz
    X, Z, _, _, _, _ = define_computational_domain(c.Lx, c.Lz, nx, nz)
    size = "big" if (c.Lx == 2.0e6 or c.Lz == 2.0e6) else "small"
    ns_n = int(n_noise_sources)
    if ns_n > len(_SPECTRUM["f_peak"]):
        raise ValueError("the reference specifies at most two noise spectra")
    spectrum = np.zeros((len(f), ns_n), order="F")
    for ns in range(ns_n):  # :135-137
        spectrum[:, ns] = _SPECTRUM["strength"][ns] * np.exp(-(np.abs(f) - _SPECTRUM["f_peak"][ns]) ** 2
                                                             / _SPECTRUM["bandwidth"][ns] ** 2)
    dist = np.ones((nx, nz, ns_n), order="F")  # :145
    if source_type == "gaussian":  # :149-156
        g = _GAUSSIAN[size]
        if ns_n > len(g["x"]):
            raise ValueError("more noise sources than the reference's gaussian set-up defines")
        for ns in range(ns_n):
            blob = np.exp(-((X - g["x"][ns]) ** 2 + (Z - g["z"][ns]) ** 2) / g["width"][ns] ** 2)
            dist[:, :, ns] = dist[:, :, ns] + g["magnitude"][ns] * blob.T
    elif source_type == "point":
        raise ValueError("make_noise_source 'point': the reference loops ns = 5:11 over one-element "
                         "tables (make_noise_source.m:167) and fails")
    elif source_type == "ring":  # :180-198
        r = _RING
        R = np.sqrt((X - r["x"]) ** 2 + (Z - r["z"]) ** 2)
        with np.errstate(divide="ignore", invalid="ignore"):
            angle = np.arctan(np.abs(X - r["x"]) / np.abs(Z - r["z"])) * 180 / np.pi
        angle[np.isnan(angle)] = 0.0
        shell = (R > (r["radius"] - r["thickness"] / 2)) & (R < (r["radius"] + r["thickness"] / 2))
        base = np.exp(-np.abs(R - r["radius"]) ** 2 / 9e8) * shell
        if r["angle_cover"] == 90:
            field = base
        else:
            lo = r["angle_cover"] - r["taper_width"]
            field = base + np.exp(-np.abs(R - r["radius"]) ** 2 / 9e8) * (
                5 * np.exp(-(angle - lo) ** 2 / r["taper_strength"] * ((angle > lo) & (angle < r["angle_cover"])))
                * (shell & (angle <= r["angle_cover"])))
        dist[:, :, 0] = _untransposed(field, nx, nz, "make_noise_source 'ring'")
    elif source_type == "picture":  # :201-204
        if picture is None:
            raise ValueError("make_noise_source 'picture' needs the image (the reference reads source.png)")
        A = np.asarray(picture, dtype=np.float64)
        if A.shape != (nz, nx):
            raise ValueError("the picture must be nz x nx")
        d = np.abs(A - 255.0)
        dist[:, :, 0] = np.flipud(np.abs((A - 255.0) / d.max())).T
    elif source_type != "homogeneous":
        raise ValueError(f"source_type {source_type!r}")
    if nb != 0:  # :224-227
        dist = general_source(spectrum, dist, nb)
        spectrum = np.ones((len(f), 1), order="F")
    return np.asfortranarray(dist), spectrum


def define_material_parameters(config: Config, model_type: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """[mu, rho] = define_material_parameters(nx, nz, model_type) (define_material_parameters.m:17-169).
    Model types outside the file's table load models/mu_<type>.mat in the reference; pass such
    models directly to the solvers instead."""
    nx, nz = config.nx, config.nz
    ones = np.ones((nx, nz), order="F")

    def need(rows, cols, what):
        if nx < rows or nz < cols:
            raise ValueError(f"model_type {model_type}: {what} needs nx >= {rows}, nz >= {cols} "
                             f"(MATLAB would grow the arrays)")

    def mround(v):  # MATLAB round: half away from zero
        return int(np.floor(v + 0.5)) if v >= 0 else -int(np.floor(-v + 0.5))

    if model_type == 1 or model_type in (666, 888):
        return 4.8e10 * ones, 3000.0 * ones
    if model_type == 2:  # :25-37
        rho, mu = 3000.0 * ones, 4.8e10 * ones
        X, Z, _, _, _, _ = define_computational_domain(config.Lx, config.Lz, nx, nz)
        blob = np.exp(-((X - config.Lx / 2.0) ** 2 + (Z - config.Lz / 2.0) ** 2) / (config.Lx / 10) ** 2)
        mu = mu - 8.0e9 * _untransposed(blob, nx, nz, "model_type 2")
        return np.asfortranarray(mu), rho
    if model_type == 3:  # :42-46
        rho, mu = 3000.0 * ones, (4.8e10 * ones).copy(order="F")
        mu[mround(nx / 2 * 1.25) - 1:, :] += 0.8e10
        return mu, rho
    if model_type == 4:  # :51-57
        need(102, 127, "rho(98:102,123:127)")
        rho, mu = (3000.0 * ones).copy(order="F"), (2.8e10 * ones).copy(order="F")
        h = mround(nx / 2)
        rho[:h, :] += 200.0
        mu[:h, :] += 3.5e10
        rho[97:102, 122:127] += 2000.0
        return mu, rho
    if model_type in (5, 6):  # :62-85
        need(1, 150, "mu(:,100:150)")
        rho, mu = (3000.0 * ones).copy(order="F"), (2.0e10 * ones).copy(order="F")
        for k in range(100, 151):
            mu[:, k - 1] = mu[:, k - 1] + (k - 100) * 10.0e8
        for k in range(151, nz + 1):
            mu[:, k - 1] = mu[:, 149]
        if model_type == 6:
            need(102, 127, "rho(98:102,123:127)")
            rho[97:102, 122:127] += 2000.0
        return mu, rho
    if model_type == 7:  # :90-93
        need(331, 1, "mu(1:330,:)")
        mu = ones.copy(order="F")
        mu[:330, :] = 3.675e10
        mu[330:, :] = 2.7e10
        return mu, 3000.0 * ones
    tables = {999: dict(x=(1.15e6, 1.45e6), z=(1.0e6, 1.0e6), wx=(1.1e5, 1.1e5), wz=(2.5e5, 2.5e5), amp=1.0e10),
              9999: dict(x=(2.3e5, 2.9e5), z=(2.0e5, 2.0e5), wx=(2.2e4, 2.2e4), wz=(5e4, 5e4), amp=4.0e9),
              200: dict(x=(1.275e6,), z=(0.975e6,), wx=(0.7e2,), wz=(0.7e2,), amp=3.0e31)}
    if model_type in tables:  # :104-160
        t = tables[model_type]
        rho, mu = 3000.0 * ones, 4.8e10 * ones
        X, Z, _, _, _, _ = define_computational_domain(config.Lx, config.Lz, nx, nz)
        for i in range(len(t["x"])):
            sign = (-1.0) ** (i + 1)
            mu = mu + sign * t["amp"] * np.exp(-((X - t["x"][i]) ** 2) / t["wx"][i] ** 2).T \
                * np.exp(-((Z - t["z"][i]) ** 2) / t["wz"][i] ** 2).T
        return np.asfortranarray(mu), rho
    raise ValueError(f"model_type {model_type}: the reference loads models/mu_{model_type}.mat; pass mu and rho "
                     f"arrays directly")

"""On-disk interop with the reference's MAT files (MAT v5, read and written with scipy.io).

Files and variables as the reference writes and reads them:

  array_<n_ref>_ref.mat            array (n x 2, metres), ref_stat (n_ref x 2)
                                   code/calculate_data.m:222, read at inversion/start_inversion.m:84
  data_<n_ref>_ref_<nb>.mat        c_data (n_ref*n_rec x 2nt-1), t (1 x 2nt-1)
                                   code/calculate_data.m:223, read at inversion/start_inversion.m:85
  G_fft_<i>.mat                    one variable (parsave names it x): the nx x nz x n_sample complex
                                   spectrum, tools/addons/parsave.m:1-5 / parload.m:1-13, used at
                                   inversion/optlib_callback_fct/eval_objective_and_gradient.m:44-53
  model_<it>.mat                   model: the iteration's struct with its fields sorted (orderfields)
                                   plus sigma.source / sigma.structure and config.n_basis_fct / nx / nz,
                                   inversion/optlib_callback_fct/new_iteration.m:28-46

Arrays keep MATLAB's column-major shapes (a MATLAB nx x nz array is an (nx, nz) Fortran-ordered numpy
array). ``solution.mat`` of start_inversion.m:215 is MAT v7.3 (HDF5): ``save_solution`` writes it
through ``mat73`` (the HDF5 C library), or as MAT v5 with ``version="5"``, which MATLAB reads the same
way for variables below 2 GB; ``load_solution`` reads either.

``calculate_data`` is code/calculate_data.m:180-223 on the GPU: the Green and correlation runs of all
reference stations go through one batch engine instead of a parfor loop, and the two files are
written as the reference writes them.
"""
from __future__ import annotations

import os
from typing import Dict, Optional

import numpy as np
import scipy.io as sio


def _save(path: str, variables: Dict[str, object]) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    sio.savemat(path, variables, format="5", do_compression=False, oned_as="row")


def _load(path: str) -> dict:
    d = sio.loadmat(path, squeeze_me=False, struct_as_record=False)
    return {k: v for k, v in d.items() if not k.startswith("__")}


def save_array(path: str, array: np.ndarray, ref_stat: np.ndarray) -> None:
    """save(sprintf('array_%i_ref.mat', n_ref), 'array', 'ref_stat') (calculate_data.m:222)."""
    _save(path, {"array": np.atleast_2d(np.asarray(array, dtype=np.float64)),
                 "ref_stat": np.atleast_2d(np.asarray(ref_stat, dtype=np.float64))})


def load_array(path: str) -> dict:
    """usr_par.network = load(...) (start_inversion.m:84): fields array, ref_stat."""
    d = _load(path)
    return {"array": np.asarray(d["array"], dtype=np.float64), "ref_stat": np.asarray(d["ref_stat"], dtype=np.float64)}


def save_data(path: str, c_data: np.ndarray, t: np.ndarray) -> None:
    """save(sprintf('data_%i_ref_%i.mat', n_ref, n_basis_fct), 'c_data', 't') (calculate_data.m:223);
    t is MATLAB's row vector -(nt-1)dt : dt : (nt-1)dt."""
    _save(path, {"c_data": np.atleast_2d(np.asarray(c_data, dtype=np.float64)),
                 "t": np.asarray(t, dtype=np.float64).reshape(1, -1)})


def load_data(path: str) -> dict:
    """usr_par.data = load(...) (start_inversion.m:85): fields c_data (n_ref*n_rec x N), t (1 x N)."""
    d = _load(path)
    return {"c_data": np.asarray(d["c_data"], dtype=np.float64), "t": np.asarray(d["t"], dtype=np.float64)}


def save_G_fft(path: str, G_fft: np.ndarray) -> None:
    """parsave(['G_fft_' num2str(i) '.mat'], G_fft) (eval_objective_and_gradient.m:50): the variable
    is named x (parsave.m:3)."""
    _save(path, {"x": np.asarray(G_fft, dtype=np.complex128)})


def load_G_fft(path: str) -> np.ndarray:
    """parload (parload.m:1-13): the file's only variable, whatever its name; an error otherwise."""
    d = _load(path)
    if len(d) != 1:
        raise ValueError(f"Problem with {path}!")
    return np.asarray(next(iter(d.values())))


def model_struct(model: dict, usr_par) -> dict:
    """The struct new_iteration.m:28-46 saves: the model's fields in sorted order (orderfields) plus
    sigma.source / sigma.structure and config.n_basis_fct / nx / nz (and ring when switched on)."""
    out = {k: model[k] for k in sorted(model)}
    out["sigma"] = {"source": np.asarray(usr_par.sigma_source, dtype=np.float64).reshape(1, -1),
                    "structure": np.asarray(usr_par.sigma_structure, dtype=np.float64).reshape(1, -1)}
    c = usr_par.config
    out["config"] = {"n_basis_fct": float(c.n_basis_fct), "nx": float(c.nx), "nz": float(c.nz)}
    ring = getattr(usr_par, "ring", None)
    if isinstance(ring, dict) and ring.get("switch") == "yes":
        out["ring"] = ring
    return out


def save_model(path: str, model: dict, usr_par) -> None:
    """save(sprintf('model_%i.mat', it), 'model') (new_iteration.m:46). model holds host arrays or
    scalars (m, gradient, objective, normg, normg_frac, correlation, ...)."""
    m = {k: (v.detach().cpu().numpy() if hasattr(v, "detach") else v) for k, v in model.items()}
    _save(path, {"model": model_struct(m, usr_par)})


def load_model(path: str) -> dict:
    """The model struct of a model_<it>.mat file as a dict of its fields (sub-structs as dicts)."""
    d = _load(path)["model"]

    def conv(s):
        if isinstance(s, np.ndarray) and s.dtype == object and s.size == 1:
            s = s.flat[0]
        if hasattr(s, "_fieldnames"):
            return {f: conv(getattr(s, f)) for f in s._fieldnames}
        return s
    return conv(d)


def save_solution(path: str, flag: int, mfinal: np.ndarray, usr_par_fields: dict, version: str = "7.3") -> None:
    """save('solution.mat', 'flag', 'mfinal', 'usr_par', '-v7.3') (start_inversion.m:215); version "5"
    writes MAT v5 instead."""
    v = {"flag": float(flag), "mfinal": np.asarray(mfinal, dtype=np.float64).reshape(-1, 1), "usr_par": usr_par_fields}
    if version == "7.3":
        from . import mat73
        mat73.save(path, v)
    elif version == "5":
        _save(path, v)
    else:
        raise ValueError(f"unknown MAT-file version {version!r} (7.3 or 5)")


def load_solution(path: str) -> dict:
    """flag, mfinal and usr_par of a solution file (MAT v7.3 or v5)."""
    with open(path, "rb") as f:
        head = f.read(20)
    if head.startswith(b"MATLAB 7.3"):
        from . import mat73
        return mat73.load(path)
    d = _load(path)

    def conv(s):
        if isinstance(s, np.ndarray) and s.dtype == object and s.size == 1:
            s = s.flat[0]
        if hasattr(s, "_fieldnames"):
            return {f: conv(getattr(s, f)) for f in s._fieldnames}
        return s
    return {k: conv(v) for k, v in d.items()}


def calculate_data(config, mu, rho, spectrum, source_dist, array: np.ndarray, ref_stat: np.ndarray,
                   out_dir: Optional[str] = None) -> dict:
    """code/calculate_data.m:180-223: for every reference station (a row of ref_stat) the Green run
    and the correlation run at the other stations of array, all stations batched through one engine;
    c_data stacks the traces station by station (n_ref*n_rec x 2nt-1). With out_dir, writes
    array_<n_ref>_ref.mat and data_<n_ref>_ref_<n_basis_fct>.mat there."""
    import torch
    from . import batch as B, device as D
    c = config
    array = np.atleast_2d(np.asarray(array, dtype=np.float64))
    ref_stat = np.atleast_2d(np.asarray(ref_stat, dtype=np.float64))
    n_ref = ref_stat.shape[0]
    mu_d, rho_d = D.to_device(mu), D.to_device(rho)
    sd_d = D.to_device(source_dist)
    spec_h = np.asfortranarray(np.asarray(spectrum, dtype=np.float64))
    recs = []
    for i in range(n_ref):
        keep = ~np.all(array == ref_stat[i], axis=1)  # array(find(~ismember(array, src, 'rows')), :)
        recs.append(np.asfortranarray(array[keep]))
    n_rec = recs[0].shape[0]
    if any(r.shape[0] != n_rec for r in recs):
        raise ValueError("every reference station must appear once in array")
    G = [D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128) for _ in range(n_ref)]
    B.run_batch([B.green(mu_d, rho_d, np.asfortranarray(ref_stat[i:i + 1]), None, 0, None, None, config=c,
                         G_fft=G[i]) for i in range(n_ref)])
    traces = [D.empty_like_dims((n_rec, c.n_corr), torch.float64) for _ in range(n_ref)]
    B.run_batch([B.correlation(mu_d, rho_d, G[i], spec_h, sd_d, recs[i], 0, None, None, config=c,
                               seismograms=traces[i]) for i in range(n_ref)])
    c_data = np.concatenate([D.to_host(tr) for tr in traces], axis=0)
    t = c.t_corr()
    if out_dir is not None:
        save_array(os.path.join(out_dir, f"array_{n_ref}_ref.mat"), array, ref_stat)
        save_data(os.path.join(out_dir, f"data_{n_ref}_ref_{c.n_basis_fct}.mat"), c_data, t)
    return {"c_data": c_data, "t": t.reshape(1, -1), "array": array, "ref_stat": ref_stat}

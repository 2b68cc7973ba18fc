"""The solvers' callers on the device: L-BFGS with the Powell-Wolfe rule, the objective and gradient
over reference stations, the parameter map and the regularisation.

Mirrors of
  inversion/optlib/optlib_lbfgs.m:1-210 (``optlib_lbfgs``), optlib_wolfe.m:1-160 (``optlib_wolfe``),
  optlib_init_default_parameters_lbfgs.m:1-219 (``Options``)
  inversion/optlib_callback_fct/eval_objective_and_gradient.m:2-156, eval_objective.m,
  eval_grad_objective.m (``NoiseProblem``), map_m_to_parameters.m:2-52,
  map_gradparameters_to_gradm.m:53-106, new_iteration.m:1-49,
  optlib_restore_lbfgs_information.m:50-106, usr_par_init_default_parameters_lbfgs.m:2-172 (``UsrPar``)
  inversion/addons/regularization.m:1-29, tools/addons/gaussblur2d.m:2-38.

Model vectors are FP64 torch CUDA tensors (MATLAB column vectors nx*nz*n_maps). Every inner product
and vector update runs in the HIP library: compensated FP64 sums (fd2d_dot_device), fd2d_axpby_device,
fd2d_map_*_device, fd2d_regularization_device, fd2d_make_adjoint_sources_device. The solvers run
through the *_device entry points, so wavefields, G_fft and kernels stay in HBM.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math
import string
from typing import Any, List, Optional, Sequence

import numpy as np

from . import _lib, abi
from .config import Config

_dp = ctypes.c_void_p


def _bind():
    lib = _lib.load()
    if not getattr(lib, "_inversion_bound", False):
        lib.fd2d_dot_device.argtypes = [_dp, _dp, ctypes.c_int64, ctypes.POINTER(ctypes.c_double), _dp]
        lib.fd2d_axpby_device.argtypes = [ctypes.c_double, _dp, ctypes.c_double, _dp, _dp, ctypes.c_int64, _dp]
        lib.fd2d_axpby_f32_device.argtypes = [ctypes.c_float, _dp, ctypes.c_float, _dp, _dp, ctypes.c_int64, _dp]
        lib.fd2d_gaussblur2d_device.argtypes = [_dp, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _dp, _dp,
                                                ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int32, _dp]
        d2 = ctypes.c_double * 2
        lib.fd2d_map_m_to_parameters_device.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int32, d2, d2,
                                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, _dp]
        lib.fd2d_map_gradparameters_to_gradm_device.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int32, d2, d2,
                                                                ctypes.c_double, _dp, _dp]
        lib.fd2d_regularization_device.argtypes = [ctypes.POINTER(ctypes.c_double), _dp, _dp, _dp, ctypes.c_int64,
                                                   ctypes.c_int64, _dp, ctypes.c_int64, ctypes.c_double,
                                                   ctypes.c_double, _dp]
        for f in ("fd2d_dot_device", "fd2d_axpby_device", "fd2d_axpby_f32_device", "fd2d_gaussblur2d_device", "fd2d_map_m_to_parameters_device",
                  "fd2d_map_gradparameters_to_gradm_device", "fd2d_regularization_device"):
            getattr(lib, f).restype = ctypes.c_int
        lib._inversion_bound = True
    return lib


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check(st):
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())


def _vec(x):
    assert x.is_cuda and x.is_contiguous() and x.dtype.is_floating_point and x.element_size() == 8, \
        "expected a contiguous FP64 CUDA tensor"
    return x


# ---------------------------------------------------------------------------------------------
# FP64 vector kernels
# ---------------------------------------------------------------------------------------------
def dot(x, y) -> float:
    """x'*y with compensated FP64 summation on the device."""
    _vec(x), _vec(y)
    if x.numel() != y.numel():
        raise ValueError("dot: length mismatch")
    out = ctypes.c_double()
    _check(_bind().fd2d_dot_device(x.data_ptr(), y.data_ptr(), x.numel(), ctypes.byref(out), _stream()))
    return out.value


def norm(x) -> float:
    return math.sqrt(dot(x, x))


def axpby(alpha: float, x, beta: float, y, out=None):
    """out = alpha*x + beta*y (out may alias x or y)."""
    import torch
    _vec(x), _vec(y)
    out = torch.empty_like(x) if out is None else _vec(out)
    _check(_bind().fd2d_axpby_device(float(alpha), x.data_ptr(), float(beta), y.data_ptr(), out.data_ptr(), x.numel(),
                                     _stream()))
    return out


def axpby_f32(alpha: float, x, beta: float, y, out=None):
    """out = alpha*x + beta*y on contiguous FP32 CUDA tensors (out may alias x or y)."""
    import torch
    assert x.is_cuda and x.is_contiguous() and x.dtype == torch.float32 and y.shape == x.shape
    out = torch.empty_like(x) if out is None else out
    _check(_bind().fd2d_axpby_f32_device(float(alpha), x.data_ptr(), float(beta), y.data_ptr(), out.data_ptr(),
                                         x.numel(), _stream()))
    return out


def gaussblur2d(Y, x1, x2, sigma, out=None):
    """gaussblur2d.m:2-38 on a device (n2, n1) tensor holding the MATLAB n1 x n2 array."""
    import torch
    n2, n1 = Y.shape
    _vec(Y)
    x1d = torch.as_tensor(np.asarray(x1, dtype=np.float64), device=Y.device)
    x2d = torch.as_tensor(np.asarray(x2, dtype=np.float64), device=Y.device)
    out = torch.empty_like(Y) if out is None else _vec(out)
    _check(_bind().fd2d_gaussblur2d_device(Y.data_ptr(), n1, n2, n1, x1d.data_ptr(), x2d.data_ptr(), float(sigma[0]),
                                           float(sigma[1]), out.data_ptr(), n1, _stream()))
    return out


# ---------------------------------------------------------------------------------------------
# usr_par, the parameter map and the regularisation
# ---------------------------------------------------------------------------------------------
@dataclasses.dataclass
class UsrPar:
    """The usr_par fields the objective reads, with the defaults of
    usr_par_init_default_parameters_lbfgs.m:2-172 (nested MATLAB names flattened with '_')."""
    config: Config
    ref_stat: np.ndarray                 # network.ref_stat, n_ref x 2 [m]
    array: np.ndarray                    # network.array, (n_rec + 1) x 2 [m]
    c_data: np.ndarray                   # data.c_data, (n_ref * n_rec) x N observed correlations
    t: np.ndarray                        # data.t, N times
    rho: np.ndarray                      # define_material_parameters(...) rho, nx x nz
    spectrum: np.ndarray                 # make_noise_source(...) spectrum, n_sample x n_src
    type: str = "joint"                  # 'source' | 'structure' | 'joint'
    measurement_source: str = "waveform_difference"
    measurement_structure: str = "waveform_difference"
    veldis: str = "dis"
    sigma_source: Sequence[float] = (1.0, 1.0)     # kernel.sigma.source
    sigma_structure: Sequence[float] = (1.0, 1.0)  # kernel.sigma.structure
    kernel_weighting: float = 0.5
    ref_source: float = 0.0              # initial.ref_source
    ref_structure: float = 1.0           # initial.ref_structure
    mu_0: float = 4.8e10                 # initial.mu_0
    ring: str = "no"                     # ring.switch
    apply_filter: str = "no"             # filter.apply_filter
    reg_alpha: float = 0.0               # regularization.alpha
    reg_beta: float = 0.0                # regularization.beta
    reg_weighting: Any = 1.0             # regularization.weighting: scalar or nx*nz
    m0: Any = None                       # usr_par.m0 of regularization.m (array or device tensor); None: m
    window: Sequence[float] = (100.0, 200.0)  # make_adjoint_sources.m:85-86
    model: List[dict] = dataclasses.field(default_factory=list)  # usr_par.model (new_iteration.m:29)

    def check(self):
        if self.ring != "no":
            raise abi.Fd2dError(abi.FD2D_EUNSUPPORTED, "ring.switch 'yes' needs imfilter kernels the reference "
                                                        "no longer defines (map_m_to_parameters.m:30)")
        if self.apply_filter != "no":
            raise abi.Fd2dError(abi.FD2D_EUNSUPPORTED, "filter_correlations is not part of this engine")
        if self.type not in ("source", "structure", "joint"):
            raise ValueError(f"usr_par.type {self.type!r}")


def map_m_to_parameters(m, usr: UsrPar, out=None):
    """map_m_to_parameters.m:16-49 on the device: returns an (n_maps, nz, nx) FP64 tensor."""
    import torch
    c = usr.config
    P = c.nx * c.nz
    n_maps = m.numel() // P
    if n_maps * P != m.numel():
        raise ValueError("m must hold nx*nz*n_maps values")
    out = torch.empty((n_maps, c.nz, c.nx), dtype=torch.float64, device=m.device) if out is None else out
    cc = c.to_c()
    d2 = ctypes.c_double * 2
    _check(_bind().fd2d_map_m_to_parameters_device(ctypes.byref(cc), _vec(m).data_ptr(), n_maps,
                                                   d2(*usr.sigma_source), d2(*usr.sigma_structure),
                                                   float(usr.ref_source), float(usr.ref_structure), float(usr.mu_0),
                                                   out.data_ptr(), _stream()))
    return out


def map_gradparameters_to_gradm(grad_parameters, usr: UsrPar):
    """map_gradparameters_to_gradm.m:70-103 on the device: returns the FP64 column vector."""
    import torch
    c = usr.config
    n = grad_parameters.numel()
    n_maps = n // (c.nx * c.nz)
    out = torch.empty(n, dtype=torch.float64, device=grad_parameters.device)
    cc = c.to_c()
    d2 = ctypes.c_double * 2
    _check(_bind().fd2d_map_gradparameters_to_gradm_device(ctypes.byref(cc), _vec(grad_parameters).data_ptr(), n_maps,
                                                           d2(*usr.sigma_source), d2(*usr.sigma_structure),
                                                           float(usr.mu_0), out.data_ptr(), _stream()))
    return out


def regularization(j: float, g, m, usr: UsrPar):
    """[j, g] = regularization(j, g, m, usr_par) (regularization.m:1-29); g None or 0 is the
    reference's scalar g (only j changes). g is updated in place and returned."""
    import torch
    c = usr.config
    m0 = usr.m0 if usr.m0 is not None else m
    if not torch.is_tensor(m0):
        m0 = torch.as_tensor(np.asarray(m0, dtype=np.float64), device=m.device)
    w = usr.reg_weighting
    if not torch.is_tensor(w):
        w = torch.as_tensor(np.atleast_1d(np.asarray(w, dtype=np.float64)).ravel(order="F"), device=m.device)
    gp = g.data_ptr() if torch.is_tensor(g) else None
    jj = ctypes.c_double(float(j))
    _check(_bind().fd2d_regularization_device(ctypes.byref(jj), gp, _vec(m).data_ptr(), _vec(m0).data_ptr(), m.numel(),
                                              c.nx * c.nz, _vec(w).data_ptr(), w.numel(), float(usr.reg_alpha),
                                              float(usr.reg_beta), _stream()))
    return jj.value, g


# ---------------------------------------------------------------------------------------------
# The objective and gradient over reference stations
# ---------------------------------------------------------------------------------------------
def receivers_of(array, src_row):
    """rec = array(find(~ismember(array, src, 'rows')), :) (eval_objective_and_gradient.m:40)."""
    keep = ~np.all(array == src_row[None, :], axis=1)
    return array[keep]


class NoiseProblem:
    """eval_objective_and_gradient.m:2-156 with eval_objective.m / eval_grad_objective.m.

    Per reference station: Green run, correlation run, misfits and adjoint sources of the source and
    structure measurements combined with kernel.weighting, then the adjoint runs of usr_par.type
    ('source': one mode-0 run against G_fft; 'structure' / 'joint': mode 1 against the correlation
    wavefield, then mode 0 against the Green wavefield with the transformed adjoint state). Kernels
    sum in HBM; the gradient is mapped back to m and regularised. For type 'source' the G_fft of
    each station is kept in HBM across calls, as the reference keeps G_fft_<i>.mat (:44-53)."""

    def __init__(self, usr: UsrPar, batched: bool = False):
        import torch
        from . import device as D
        usr.check()
        self.usr, self.D, self.torch = usr, D, torch
        # batched: the solves of all reference stations run together through one batch engine (the
        # reference's parfor over stations, eval_objective_and_gradient.m:54); results are identical
        self.batched = batched
        c = usr.config
        self.c = c
        self.rho = D.to_device(usr.rho)
        self.spectrum = D.to_device(usr.spectrum)
        self.t = torch.as_tensor(np.asarray(usr.t, dtype=np.float64), device="cuda")
        ref = np.atleast_2d(np.asarray(usr.ref_stat, dtype=np.float64))
        self.ref = ref
        self.n_rec = usr.array.shape[0] - 1
        self.recs = [receivers_of(np.asarray(usr.array, dtype=np.float64), ref[i]) for i in range(ref.shape[0])]
        for r in self.recs:
            if r.shape[0] != self.n_rec:
                raise ValueError("every reference station must appear once in network.array")
        nt = len(usr.t)
        if nt != c.n_corr:
            raise ValueError("data.t must be the correlation time axis (2nt-1 samples)")
        # observed correlations of station i: (N, n_rec) device tensor (MATLAB n_rec x N)
        cd = np.asarray(usr.c_data, dtype=np.float64)
        self.c_data = [D.to_device(cd[i * self.n_rec:(i + 1) * self.n_rec, :]) for i in range(ref.shape[0])]
        self.zeros_du = torch.zeros((nt, self.n_rec), dtype=torch.float64, device="cuda")
        self.g_cache = {}
        self.last_misfits = None

    def _buffers(self):
        torch, D, c = self.torch, self.D, self.c
        f64, f32, cd = torch.float64, torch.float32, torch.complex128
        nr = self.n_rec
        return dict(G_fft=D.empty_like_dims((c.nx, c.nz, c.n_sample), cd),
                    G=D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32),
                    corr=D.empty_like_dims((nr, c.n_corr), f64),
                    C=D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32),
                    stf=torch.zeros((c.n_corr, nr), dtype=cd, device="cuda"),
                    A1=D.empty_like_dims((c.nx, c.nz, c.n_sample), cd))

    def _station_program(self, i, mu, rho, spec, source_dist, only_objective):
        """The solves of reference station i (eval_objective_and_gradient.m:54-104) as a program for
        batch.run_programs: yields each stage's solves, returns (traces, j_source_n, j_structure_n,
        [kernels])."""
        from . import batch as B
        from .measurement import make_adjoint_sources_device
        torch, D, c, usr = self.torch, self.D, self.c, self.usr
        src = np.asfortranarray(self.ref[i:i + 1])
        rec = np.asfortranarray(self.recs[i])
        nr = self.n_rec
        spec_h = np.asfortranarray(D.to_host(spec)) if torch.is_tensor(spec) else spec
        corr = D.empty_like_dims((nr, c.n_corr), torch.float64)
        if usr.type == "source":
            if i not in self.g_cache:
                G_fft = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
                yield [B.green(mu, rho, src, None, 0, None, None, config=c, G_fft=G_fft)]
                self.g_cache[i] = G_fft
            G_fft = self.g_cache[i]
            yield [B.correlation(mu, rho, G_fft, spec_h, source_dist, rec, 0, None, None, config=c, seismograms=corr)]
        else:
            G_fft = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
            G = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), torch.float32)
            yield [B.green(mu, rho, src, None, 1, None, None, config=c, G_fft=G_fft, G_out=G)]
            C = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), torch.float32)
            yield [B.correlation(mu, rho, G_fft, spec_h, source_dist, rec, 1, None, None, config=c, seismograms=corr,
                                 C_out=C)]
        jn_s, a_s = make_adjoint_sources_device(corr, self.c_data[i], self.zeros_du, self.t, usr.veldis,
                                                usr.measurement_source, "1st", usr.window)
        jn_t, a_t = make_adjoint_sources_device(corr, self.c_data[i], self.zeros_du, self.t, usr.veldis,
                                                usr.measurement_structure, "1st", usr.window)
        w = float(usr.kernel_weighting)
        adstf = axpby(1.0 - w, a_s, w, a_t)  # :81-82
        jn_s_h, jn_t_h = jn_s.cpu().numpy(), jn_t.cpu().numpy()
        if only_objective:
            return corr, jn_s_h, jn_t_h, []
        nbk = max(c.n_basis_fct, 1) + 1
        stf = torch.zeros((c.n_corr, nr), dtype=torch.complex128, device="cuda")
        torch.view_as_real(stf)[..., 0].copy_(adstf)  # complex(adstf)
        K1 = D.empty_like_dims((c.nx, c.nz, nbk), torch.float64)
        if usr.type == "source":
            yield [B.adjoint(mu, rho, None, stf, rec, None, spec_h, None, G_fft, 0, None, None, config=c, K=K1)]
            return corr, jn_s_h, jn_t_h, [K1]
        A1 = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
        yield [B.adjoint(mu, rho, C, stf, rec, None, spec_h if usr.type == "joint" else None, None,
                         G_fft if usr.type == "joint" else None, 1, None, None, config=c, K=K1, stf_fft=A1)]
        del C
        K2 = D.empty_like_dims((c.nx, c.nz, nbk), torch.float64)
        yield [B.adjoint(mu, rho, G, A1, rec, None, spec_h, source_dist, None, 0, None, None, config=c, K=K2)]
        return corr, jn_s_h, jn_t_h, [K1, K2]

    def _eval_batched(self, m, only_objective):
        from . import batch as B
        torch, c, usr = self.torch, self.c, self.usr
        Pm = map_m_to_parameters(m, usr)
        n_maps = Pm.shape[0]
        nb_k = max(c.n_basis_fct, 1) + 1
        source_dist, mu = Pm[:n_maps - 1], Pm[n_maps - 1]
        n_ref, nr, N = self.ref.shape[0], self.n_rec, c.n_corr
        # a station starts when its peak fits in 70% of the available device memory beside the
        # stations running: G_fft and the transformed adjoint state, the forward snapshots G and
        # C, two kernels, traces and adjoint sources, and the library buffers of its largest solve
        P, K = c.nx * c.nz, c.n_sample
        lib = max(B.solve_bytes(abi.FD2D_GREEN, c, 1, 1, 0),
                  B.solve_bytes(abi.FD2D_CORRELATION, c, 1, 0, nr, True, 1, True),
                  B.solve_bytes(abi.FD2D_ADJOINT, c, 1, nr, nr, True, 1, False, (N, nr, 1), fft_out=True),
                  B.solve_bytes(abi.FD2D_ADJOINT, c, 0, nr, nr, False, 1, True, (c.nx, c.nz, K)))
        per = 2 * 16 * P * K + 4 * P * (c.n_fw_green() + c.n_fw_corr()) + 2 * 8 * P * nb_k + 4 * 8 * nr * N + lib
        out = B.run_programs([self._station_program(i, mu, self.rho, self.spectrum, source_dist, only_objective)
                              for i in range(n_ref)], costs=[per] * n_ref, budget=int(0.7 * B.available_bytes()))
        c_all = torch.zeros((N, n_ref * nr), dtype=torch.float64, device="cuda")
        grad = torch.zeros((nb_k, c.nz, c.nx), dtype=torch.float64, device="cuda")
        w = float(usr.kernel_weighting)
        j = j_source = j_structure = 0.0
        misfits = []
        for i, (corr, jn_s_h, jn_t_h, Ks) in enumerate(out):  # station order, as the sequential loop
            c_all[:, i * nr:(i + 1) * nr].copy_(corr)
            j_n = (1.0 - w) * jn_s_h + w * jn_t_h
            misfits.append(j_n)
            for K in Ks:
                axpby(1.0, grad.view(-1), 1.0, K.view(-1), out=grad.view(-1))
            j = j + float(np.sum(j_n))
            j_source = j_source + float(np.sum(jn_s_h))
            j_structure = j_structure + float(np.sum(jn_t_h))
        self.last_misfits = {"total": j, "source": j_source, "structure": j_structure, "per_station": misfits}
        if only_objective:
            j, _ = regularization(j, None, m, usr)
            return j, 0, c_all
        if grad.shape[0] != n_maps:
            raise ValueError(f"m holds {n_maps} maps but the kernels {grad.shape[0]} (n_basis_fct + 1)")
        g = map_gradparameters_to_gradm(grad, usr)
        j, g = regularization(j, g, m, usr)
        return j, g, c_all

    def eval_objective_and_gradient(self, m, name: str = "", only_objective: bool = False):
        if self.batched:
            return self._eval_batched(m, only_objective)
        from .measurement import make_adjoint_sources_device
        torch, D, c, usr = self.torch, self.D, self.c, self.usr
        Pm = map_m_to_parameters(m, usr)
        n_maps = Pm.shape[0]
        nb_k = max(c.n_basis_fct, 1) + 1
        source_dist = Pm[:n_maps - 1]
        mu = Pm[n_maps - 1]
        rho, spec = self.rho, self.spectrum
        n_ref = self.ref.shape[0]
        nr, N = self.n_rec, c.n_corr
        c_all = torch.zeros((N, n_ref * nr), dtype=torch.float64, device="cuda")
        grad = torch.zeros((nb_k, c.nz, c.nx), dtype=torch.float64, device="cuda")
        K1 = torch.zeros_like(grad)
        K2 = torch.zeros_like(grad)
        b = self._buffers()
        w = float(usr.kernel_weighting)
        j = j_source = j_structure = 0.0
        misfits = []
        for i in range(n_ref):
            src = D.to_device(self.ref[i:i + 1])
            rec = D.to_device(self.recs[i])
            if usr.type == "source":
                if i not in self.g_cache:
                    G_fft = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
                    D.run_forward1_green(mu, rho, src, None, 0, None, None, config=c, G_fft=G_fft)
                    self.g_cache[i] = G_fft
                G_fft = self.g_cache[i]
                D.run_forward2_correlation(mu, rho, G_fft, spec, source_dist, rec, 0, None, None, config=c,
                                           seismograms=b["corr"])
            else:
                G_fft = b["G_fft"]
                D.run_forward1_green(mu, rho, src, None, 1, None, None, config=c, G_fft=G_fft, G_out=b["G"])
                D.run_forward2_correlation(mu, rho, G_fft, spec, source_dist, rec, 1, None, None, config=c,
                                           seismograms=b["corr"], C_out=b["C"])
            c_all[:, i * nr:(i + 1) * nr].copy_(b["corr"])
            jn_s, a_s = make_adjoint_sources_device(b["corr"], self.c_data[i], self.zeros_du, self.t, usr.veldis,
                                                    usr.measurement_source, "1st", usr.window)
            jn_t, a_t = make_adjoint_sources_device(b["corr"], self.c_data[i], self.zeros_du, self.t, usr.veldis,
                                                    usr.measurement_structure, "1st", usr.window)
            # eval_objective_and_gradient.m:81-82
            adstf = axpby(1.0 - w, a_s, w, a_t)
            jn_s_h, jn_t_h = jn_s.cpu().numpy(), jn_t.cpu().numpy()
            j_n = (1.0 - w) * jn_s_h + w * jn_t_h
            misfits.append(j_n)
            if not only_objective:
                stf = b["stf"]
                torch.view_as_real(stf)[..., 0].copy_(adstf)  # complex(adstf)
                if usr.type == "source":
                    D.run_noise_adjoint(mu, rho, None, stf, rec, None, spec, None, G_fft, 0, None, None, config=c,
                                        K=K1)
                    axpby(1.0, grad.view(-1), 1.0, K1.view(-1), out=grad.view(-1))
                else:
                    D.run_noise_adjoint(mu, rho, b["C"], stf, rec, None, spec if usr.type == "joint" else None,
                                        None, G_fft if usr.type == "joint" else None, 1, None, None, config=c, K=K1,
                                        stf_fft=b["A1"])
                    D.run_noise_adjoint(mu, rho, b["G"], b["A1"], rec, None, spec, source_dist, None, 0, None, None,
                                        config=c, K=K2)
                    axpby(1.0, grad.view(-1), 1.0, K1.view(-1), out=grad.view(-1))
                    axpby(1.0, grad.view(-1), 1.0, K2.view(-1), out=grad.view(-1))
            j = j + float(np.sum(j_n))
            j_source = j_source + float(np.sum(jn_s_h))
            j_structure = j_structure + float(np.sum(jn_t_h))
        self.last_misfits = {"total": j, "source": j_source, "structure": j_structure, "per_station": misfits}
        if only_objective:
            j, _ = regularization(j, None, m, usr)
            return j, 0, c_all
        if grad.shape[0] != n_maps:
            raise ValueError(f"m holds {n_maps} maps but the kernels {grad.shape[0]} (n_basis_fct + 1)")
        g = map_gradparameters_to_gradm(grad, usr)
        j, g = regularization(j, g, m, usr)
        return j, g, c_all

    def eval_hessian_vector_product(self, m, dm_parameters, name: str = ""):
        """eval_hessian_vector_product_2.m:2-197: H_m for the model m and the parameter
        perturbation dm_parameters (physical units, nx x nz x n_maps as an (n_maps, nz, nx) device
        tensor, or only its source / structure maps for usr.type 'source' / 'structure'). Per
        station: Green and Born Green (mode 5), correlation, Born correlation and the correlation
        of the perturbed sources, first- and second-order adjoint sources, then the adjoint runs
        YELLOW (mode 13), GRAY (11), BLACK (12, Born) and RED (10, Born). The Born inputs are the
        saved wavefields of every step, so the configuration needs fw_nth = 1."""
        from .measurement import make_adjoint_sources_device
        torch, D, c, usr = self.torch, self.D, self.c, self.usr
        Pm, mu, sd_ref = self._hvp_model(m)
        rho, spec = self.rho, self.spectrum
        sd_pert, dmu, has_dmu, has_dsd = self._hvp_perturbation(Pm, dm_parameters)
        f64, f32, cd = torch.float64, torch.float32, torch.complex128
        nr, N = self.n_rec, c.n_corr
        w = float(usr.kernel_weighting)
        nbk = max(c.n_basis_fct, 1) + 1
        H = torch.zeros((nbk, c.nz, c.nx), dtype=f64, device="cuda")
        K = torch.zeros_like(H)

        def green(mode, dmu_, G_in):
            G_fft = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
            G_out = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32)
            D.run_forward1_green(mu, rho, src, None, mode, dmu_, G_in, config=c, G_fft=G_fft, G_out=G_out)
            return G_fft, G_out

        def corr(G_fft, sd, dmu_, C_in):
            s_ = D.empty_like_dims((nr, N), f64)
            C_ = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
            D.run_forward2_correlation(mu, rho, G_fft, spec, sd, rec, 1, dmu_, C_in, config=c, seismograms=s_,
                                       C_out=C_)
            return s_, C_

        def adjoint(u_fwd, stf, sd, G_fft, mode, dmu_, u_in):
            K.zero_()
            sf = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd) if mode in (12, 13) else None
            uo = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32) if mode in (11, 13) else None
            D.run_noise_adjoint(mu, rho, u_fwd, stf, rec, None, spec, sd, G_fft, mode, dmu_, u_in, config=c, K=K,
                                stf_fft=sf, u_out=uo)
            return sf, uo

        def add(Kpart=None):
            # H += K, with the kernel maps of the other parameter class zeroed (:120, :157-159)
            if Kpart == "structure":
                K[:nbk - 1].zero_()
            elif Kpart == "source":
                K[nbk - 1].zero_()
            axpby(1.0, H.view(-1), 1.0, K.view(-1), out=H.view(-1))

        for i in range(self.ref.shape[0]):
            src = D.to_device(self.ref[i:i + 1])
            rec = D.to_device(self.recs[i])
            G_fft, G = green(1, None, None)  # :57
            dG_fft = dG = None
            if has_dmu:
                dG_fft, dG = green(5, dmu, G)  # :61
            c_ref, C = corr(G_fft, sd_ref, None, None)  # :71
            c_pert = torch.zeros_like(c_ref)
            dC = None
            if has_dmu:  # :75
                c1, dC = corr(dG_fft, sd_ref, dmu, C)
                axpby(1.0, c_pert.view(-1), 1.0, c1.view(-1), out=c_pert.view(-1))
            if has_dsd:  # :84
                c2, dC2 = corr(G_fft, sd_pert, None, None)
                axpby(1.0, c_pert.view(-1), 1.0, c2.view(-1), out=c_pert.view(-1))
                dC = dC2 if dC is None else axpby_f32(1.0, dC, 1.0, dC2)  # single arrays (:91)
            stf = {}
            for order in ("1st", "2nd"):  # :99-105
                _, a_s = make_adjoint_sources_device(c_ref, self.c_data[i], c_pert, self.t, usr.veldis,
                                                     usr.measurement_source, order, usr.window)
                _, a_t = make_adjoint_sources_device(c_ref, self.c_data[i], c_pert, self.t, usr.veldis,
                                                     usr.measurement_structure, order, usr.window)
                st = torch.zeros((N, nr), dtype=cd, device="cuda")
                torch.view_as_real(st)[..., 0].copy_(axpby(1.0 - w, a_s, w, a_t))
                stf[order] = st
            u_dag = uT = p = None
            if usr.type != "source":  # YELLOW :113-121
                uT, u_dag = adjoint(dC, stf["1st"], None, dG_fft, 13, None, None)
                add("structure" if usr.type == "structure" else None)
            if usr.type != "source" and has_dmu:  # GRAY :129-133
                _, p = adjoint(dG, uT, sd_ref, None, 11, None, None)
                add()
            if usr.type == "joint" and has_dsd:  # :141-146
                _, pa = adjoint(G, uT, sd_pert, None, 11, None, None)
                add()
                p = pa if p is None else axpby_f32(1.0, p, 1.0, pa)
            # BLACK :154-160
            kT, _ = adjoint(C, stf["2nd"], None, G_fft, 12, dmu if u_dag is not None else None, u_dag)
            add("source" if usr.type == "source" else ("structure" if usr.type == "structure" else None))
            if usr.type != "source":  # RED :168-170
                adjoint(G, kT, sd_ref, None, 10, dmu if p is not None else None, p)
                add()
        return self._hvp_finish(H)

    def _hvp_model(self, m):
        if self.c.fw_nth != 1:
            raise ValueError("the Hessian-vector product needs fw_nth = 1 (eval_hessian_vector_product_2.m:75,154)")
        Pm = map_m_to_parameters(m, self.usr)
        n_maps = Pm.shape[0]
        return Pm, Pm[n_maps - 1], Pm[:n_maps - 1]

    def _hvp_perturbation(self, Pm, dm_parameters):
        """dm_parameters as (source maps, structure map, has_dmu, has_dsd): the full parameter
        array, or only the source / structure maps for usr.type 'source' / 'structure'
        (eval_hessian_vector_product_2.m:19-31)."""
        torch, usr = self.torch, self.usr
        n_maps = Pm.shape[0]
        dmp = dm_parameters if torch.is_tensor(dm_parameters) else torch.as_tensor(
            np.ascontiguousarray(np.asarray(dm_parameters, dtype=np.float64).reshape(-1, order="F")), device="cuda")
        if dmp.numel() != Pm.numel():
            dm = torch.zeros_like(Pm)
            if usr.type == "source":
                dm[:n_maps - 1].view(-1).copy_(dmp.reshape(-1))
            elif usr.type == "structure":
                dm[n_maps - 1].view(-1).copy_(dmp.reshape(-1))
        else:
            dm = dmp.reshape(Pm.shape).contiguous()
        sd_pert, dmu = dm[:n_maps - 1], dm[n_maps - 1]
        return sd_pert, dmu, bool(torch.count_nonzero(dmu).item()), bool(torch.count_nonzero(sd_pert).item())

    def _hvp_finish(self, H):
        Hm = map_gradparameters_to_gradm(H, self.usr)  # map_Hparameters_to_Hm.m:19-38
        # :191-196: alpha / beta * weighting on the interior (the reference adds them without dm)
        if self.usr.reg_alpha or self.usr.reg_beta:
            add_v = self._hvp_reg_term(Hm.numel())
            axpby(1.0, Hm, 1.0, add_v, out=Hm)
        return Hm

    def _hvp_station_program(self, i, mu, sd_ref, probes):
        """The solves of reference station i for the Hessian-vector products of several
        perturbations (eval_hessian_vector_product_2.m:57-170) as a program for batch.run_programs.
        The Green and correlation runs of the model (:57, :71) do not depend on the perturbation
        and run once for all of them; every perturbation then runs its own Born Green (:61) and
        Born correlation (:75-91), its adjoint sources (:99-105) and the adjoint runs YELLOW (:113),
        GRAY (:129, :141), BLACK (:154) and RED (:168), each stage's solves of all perturbations
        in one engine stage. Arrays are dropped as soon as the last run that reads them is done.
        Returns, per perturbation, the kernels in the order the sequential loop adds them, each as
        (K, part) with part the parameter class whose maps are zeroed before the sum."""
        from . import batch as B
        from .measurement import make_adjoint_sources_device
        torch, D, c, usr = self.torch, self.D, self.c, self.usr
        f64, f32, cd = torch.float64, torch.float32, torch.complex128
        src = np.asfortranarray(self.ref[i:i + 1])
        rec = np.asfortranarray(self.recs[i])
        nr, N = self.n_rec, c.n_corr
        rho = self.rho
        spec_h = np.asfortranarray(D.to_host(self.spectrum))
        w = float(usr.kernel_weighting)
        nbk = max(c.n_basis_fct, 1) + 1
        Q = len(probes)

        def spectra():
            return D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)

        def snaps(n):
            return D.empty_like_dims((c.nx, c.nz, n), f32)

        def kern():
            return D.empty_like_dims((c.nx, c.nz, nbk), f64)

        G_fft, G = spectra(), snaps(c.n_fw_green())
        yield [B.green(mu, rho, src, None, 1, None, None, config=c, G_fft=G_fft, G_out=G)]  # :57
        c_ref, C = D.empty_like_dims((nr, N), f64), snaps(c.n_fw_corr())
        stage = [B.correlation(mu, rho, G_fft, spec_h, sd_ref, rec, 1, None, None, config=c, seismograms=c_ref,
                               C_out=C)]  # :71
        dG_fft, dG = [None] * Q, [None] * Q
        for q, (_, dmu, has_dmu, _) in enumerate(probes):
            if has_dmu:  # :61
                dG_fft[q], dG[q] = spectra(), snaps(c.n_fw_green())
                stage.append(B.green(mu, rho, src, None, 5, dmu, G, config=c, G_fft=dG_fft[q], G_out=dG[q]))
        yield stage
        c1, dC1, c2, dC2 = [None] * Q, [None] * Q, [None] * Q, [None] * Q
        stage = []
        for q, (sd_pert, dmu, has_dmu, has_dsd) in enumerate(probes):
            if has_dmu:  # :75
                c1[q], dC1[q] = D.empty_like_dims((nr, N), f64), snaps(c.n_fw_corr())
                stage.append(B.correlation(mu, rho, dG_fft[q], spec_h, sd_ref, rec, 1, dmu, C, config=c,
                                           seismograms=c1[q], C_out=dC1[q]))
            if has_dsd:  # :84
                c2[q], dC2[q] = D.empty_like_dims((nr, N), f64), snaps(c.n_fw_corr())
                stage.append(B.correlation(mu, rho, G_fft, spec_h, sd_pert, rec, 1, None, None, config=c,
                                           seismograms=c2[q], C_out=dC2[q]))
        if stage:
            yield stage
        dC, stf1, stf2 = [None] * Q, [None] * Q, [None] * Q
        for q in range(Q):
            c_pert = torch.zeros_like(c_ref)
            if c1[q] is not None:
                axpby(1.0, c_pert.view(-1), 1.0, c1[q].view(-1), out=c_pert.view(-1))
            if c2[q] is not None:
                axpby(1.0, c_pert.view(-1), 1.0, c2[q].view(-1), out=c_pert.view(-1))
            dC[q] = dC1[q] if dC2[q] is None else (dC2[q] if dC1[q] is None else axpby_f32(1.0, dC1[q], 1.0, dC2[q]))
            for order in ("1st", "2nd"):  # :99-105
                _, a_s = make_adjoint_sources_device(c_ref, self.c_data[i], c_pert, self.t, usr.veldis,
                                                     usr.measurement_source, order, usr.window)
                _, a_t = make_adjoint_sources_device(c_ref, self.c_data[i], c_pert, self.t, usr.veldis,
                                                     usr.measurement_structure, order, usr.window)
                st = torch.zeros((N, nr), dtype=cd, device="cuda")
                torch.view_as_real(st)[..., 0].copy_(axpby(1.0 - w, a_s, w, a_t))
                (stf1 if order == "1st" else stf2)[q] = st
        del c1, dC1, c2, dC2
        out = [[] for _ in range(Q)]
        uT, u_dag = [None] * Q, [None] * Q
        if usr.type != "source":  # YELLOW :113-121
            stage = []
            for q in range(Q):
                uT[q], u_dag[q], KY = spectra(), snaps(c.n_fw_corr()), kern()
                stage.append(B.adjoint(mu, rho, dC[q], stf1[q], rec, None, spec_h, None, dG_fft[q], 13, None, None,
                                       config=c, K=KY, stf_fft=uT[q], u_out=u_dag[q]))
                out[q].append((KY, "structure" if usr.type == "structure" else None))
            yield stage
        del dC, stf1, dG_fft
        p, kT = [None] * Q, [None] * Q
        stage = []
        for q, (sd_pert, dmu, has_dmu, has_dsd) in enumerate(probes):
            if usr.type != "source" and has_dmu:  # GRAY :129-133
                p[q], KG = snaps(c.n_fw_corr()), kern()
                stage.append(B.adjoint(mu, rho, dG[q], uT[q], rec, None, spec_h, sd_ref, None, 11, None, None,
                                       config=c, K=KG, u_out=p[q]))
                out[q].append((KG, None))
            if usr.type == "joint" and has_dsd:  # :141-146
                pa, KA = snaps(c.n_fw_corr()), kern()
                stage.append(B.adjoint(mu, rho, G, uT[q], rec, None, spec_h, sd_pert, None, 11, None, None,
                                       config=c, K=KA, u_out=pa))
                out[q].append((KA, None))
                p[q] = pa if p[q] is None else (p[q], pa)
            # BLACK :154-160
            kT[q], KB = spectra(), kern()
            stage.append(B.adjoint(mu, rho, C, stf2[q], rec, None, spec_h, None, G_fft, 12,
                                   dmu if u_dag[q] is not None else None, u_dag[q], config=c, K=KB, stf_fft=kT[q]))
            out[q].append((KB, "source" if usr.type == "source" else ("structure" if usr.type == "structure" else None)))
        yield stage
        del dG, uT, u_dag, stf2, C, G_fft
        if usr.type != "source":  # RED :168-170
            stage = []
            for q, (_, dmu, _, _) in enumerate(probes):
                if isinstance(p[q], tuple):
                    p[q] = axpby_f32(1.0, p[q][0], 1.0, p[q][1])
                KR = kern()
                stage.append(B.adjoint(mu, rho, G, kT[q], rec, None, spec_h, sd_ref, None, 10,
                                       dmu if p[q] is not None else None, p[q], config=c, K=KR))
                out[q].append((KR, None))
            yield stage
        return out

    def eval_hessian_vector_products(self, m, dms, memory_fraction: float = 0.7):
        """H dm for every perturbation in dms at the model m, the solves of all reference stations
        and perturbations run together through the batch engine (eval_hessian_vector_product_2.m
        per perturbation, its station loop :44-183 as the reference's parfor). Each result is
        identical to eval_hessian_vector_product(m, dm). A program (a station with a group of
        perturbations) starts when its arrays fit in memory_fraction of the available device memory
        beside the programs running; when one station with all perturbations does not fit, the
        perturbations are split into groups of equal size, each of which repeats the station's model
        Green and correlation runs."""
        from . import batch as B
        torch, D, c, usr = self.torch, self.D, self.c, self.usr
        Pm, mu, sd_ref = self._hvp_model(m)
        probes = [self._hvp_perturbation(Pm, dm) for dm in dms]
        Q, n_ref = len(probes), self.ref.shape[0]
        if Q == 0:
            return []
        P, K, nr, N = c.nx * c.nz, c.n_sample, self.n_rec, c.n_corr
        spec_b, g_snap, c_snap = 16 * P * K, 4 * P * c.n_fw_green(), 4 * P * c.n_fw_corr()
        lib = max(B.solve_bytes(abi.FD2D_GREEN, c, 1, 1, 0, dmu=True),
                  B.solve_bytes(abi.FD2D_CORRELATION, c, 1, 0, nr, True, 1, True, dmu=True),
                  B.solve_bytes(abi.FD2D_ADJOINT, c, 13, nr, nr, True, 1, True, (N, nr, 1), dmu=True, fft_out=True))
        # a station: G_fft, G, C, c_ref and the library buffers of one solve. A perturbation at its
        # peak (YELLOW, or GRAY with BLACK): two spectra (dG_fft or kT, uT), dG, two correlation
        # wavefields (dC and u_dag, or u_dag and p) plus the second Born correlation wavefield when
        # it perturbs the sources, its traces and adjoint sources, four kernels and the library
        # buffers of two solves
        any_dsd = any(pr[3] for pr in probes)
        shared = spec_b + g_snap + c_snap + 8 * nr * N + lib
        per = (2 * spec_b + g_snap + (3 if any_dsd else 2) * c_snap + 2 * 16 * N * nr + 8 * 8 * nr * N
               + 4 * 8 * P * (c.n_basis_fct + 1) + 2 * lib)
        budget = int(memory_fraction * B.available_bytes())
        q_per = int(max(1, min(Q, (budget - shared) // per)))
        n_groups = -(-Q // q_per)
        q_per = -(-Q // n_groups)  # groups of equal size
        groups = [(i, q0) for i in range(n_ref) for q0 in range(0, Q, q_per)]
        costs = [shared + len(probes[q0:q0 + q_per]) * per for _, q0 in groups]
        self.last_hvp_plan = {"perturbations_per_group": q_per, "groups_per_station": n_groups,
                              "programs_in_flight": max(1, budget // max(costs)),
                              "bytes_station": shared, "bytes_perturbation": per, "budget": budget}
        nbk = max(c.n_basis_fct, 1) + 1
        H = [torch.zeros((nbk, c.nz, c.nx), dtype=torch.float64, device="cuda") for _ in range(Q)]
        res = B.run_programs([self._hvp_station_program(i, mu, sd_ref, probes[q0:q0 + q_per]) for i, q0 in groups],
                             costs=costs, budget=budget)
        for (i, q0), kernels in zip(groups, res):  # station order, as the sequential loop
            for dq, Ks in enumerate(kernels):
                for Kq, part in Ks:
                    if part == "structure":
                        Kq[:nbk - 1].zero_()
                    elif part == "source":
                        Kq[nbk - 1].zero_()
                    h = H[q0 + dq]
                    axpby(1.0, h.view(-1), 1.0, Kq.view(-1), out=h.view(-1))
        del res
        return [self._hvp_finish(h) for h in H]

    def _hvp_reg_term(self, n):
        """alpha * weighting (source maps) and beta * weighting (structure map) on the nodes where
        init_absbound == 1, as FP64 device vector (eval_hessian_vector_product_2.m:191-196)."""
        import torch
        c, usr = self.c, self.usr
        ax, az = np.ones(c.nx), np.ones(c.nz)
        w_ = c.absorb_width

        def lo(X):
            return (X > w_) + np.exp(-(X - w_) ** 2 / (2 * w_) ** 2) * (X <= w_)

        def hi(X, L_):
            return (X < (L_ - w_)) + np.exp(-(X - (L_ - w_)) ** 2 / (2 * w_) ** 2) * (X >= (L_ - w_))
        X, Z = c.x(), c.z()
        if c.absorb_left == 1:
            ax = ax * lo(X)
        if c.absorb_right == 1:
            ax = ax * hi(X, c.Lx)
        if c.absorb_bottom == 1:
            az = az * lo(Z)
        if c.absorb_top == 1:
            az = az * hi(Z, c.Lz)
        inner = (np.outer(ax, az) == 1.0).astype(np.float64).ravel(order="F")
        ns = c.nx * c.nz
        wv = np.atleast_1d(np.asarray(usr.reg_weighting, dtype=np.float64)).ravel(order="F")
        v = np.zeros(n)
        v[:n - ns] = np.tile(inner, (n - ns) // ns) * usr.reg_alpha * np.tile(wv, (n - ns) // wv.size)
        v[n - ns:] = inner * usr.reg_beta * (wv if wv.size == ns else wv[0])
        return torch.as_tensor(v, device="cuda")

    def eval_objective(self, m, name: str = ""):
        j, _, c = self.eval_objective_and_gradient(m, name, only_objective=True)
        return j, c

    def eval_grad_objective(self, m, name: str = ""):
        _, g, c = self.eval_objective_and_gradient(m, name)
        return g, c


# ---------------------------------------------------------------------------------------------
# L-BFGS with the Powell-Wolfe step size rule
# ---------------------------------------------------------------------------------------------
@dataclasses.dataclass
class Options:
    """optlib_init_default_parameters_lbfgs.m:3-217."""
    init_step_length: float = 1.0
    wolfe_delta: float = 1e-3
    wolfe_theta: float = 0.6
    wolfe_try_to_increase_step_length: bool = True
    sufficient_decrease_angle: float = 0.1
    max_memory: int = 5
    verbose: bool = False
    output_file: Optional[str] = None     # the reference appends to 'iterations_lbfgs.tab'
    max_iterations: int = 100
    tolerance: float = 1e-3
    successive_change: float = 1.0
    successive_iterations: int = 5


class _Log:
    def __init__(self, path):
        self.f = open(path, "a+") if path else None

    def __call__(self, text):
        if self.f:
            self.f.write(text)

    def close(self):
        if self.f:
            self.f.close()


def _random_string(rng, n=8):
    """optlib_generate_random_string: the name a model is evaluated under."""
    letters = string.ascii_letters
    return "".join(letters[k] for k in rng.integers(0, len(letters), n))


def new_iteration(it, model, usr_par):
    """new_iteration.m:1-49: usr_par.model(it+1) = model (replaced or appended). usr_par is any
    object with a ``model`` list (UsrPar for the noise problem)."""
    model = dict(model)
    model["normg_frac"] = model["normg"] / usr_par.model[0]["normg"] if it > 0 else 1.0
    if it < len(usr_par.model):
        usr_par.model[it] = model
    else:
        usr_par.model.append(model)
    return usr_par


def optlib_wolfe(xj, s, stg, f, delta, theta, sig0, try_larger_steps, verbose, usr_par, problem, rng=None):
    """[sig, model] = optlib_wolfe(xj, s, stg, f, del, theta, sig0, try_larger_steps, verbose, usr_par)
    (optlib_wolfe.m:1-160): xn = xj - sig*s satisfying Armijo f - fn >= del*sig*stg and Wolfe
    gn'*s <= theta*stg, by halving, doubling and at most 5 bisections."""
    rng = np.random.default_rng() if rng is None else rng
    sig = sig0
    xn = axpby(1.0, xj, -sig, s)
    name = _random_string(rng)
    fn, gn, cn = problem.eval_objective_and_gradient(xn, name)
    while f - fn < delta * sig * stg:  # :33-48
        sig = 0.5 * sig
        xn = axpby(1.0, xj, -sig, s)
        name = _random_string(rng)
        fn, cn = problem.eval_objective(xn, name)
    if sig != sig0:  # :54-56
        gn, cn = problem.eval_grad_objective(xn, name)
    wolfe_ok = dot(gn, s) <= theta * stg
    armijo = (sig, xn, gn, fn, cn, name)
    if (not wolfe_ok) or try_larger_steps:  # :76
        if sig == sig0:
            xnn = axpby(1.0, xj, -2 * sig, s)
            name_nn = _random_string(rng)
            fnn, gnn, cnn = problem.eval_objective_and_gradient(xnn, name_nn)
            while dot(gn, s) > theta * stg and f - fnn >= 2 * delta * sig * stg:  # :85-99
                sig = 2 * sig
                xn, fn, gn, cn, name = xnn, fnn, gnn, cnn, name_nn
                xnn = axpby(1.0, xj, -2 * sig, s)
                name_nn = _random_string(rng)
                fnn, gnn, cnn = problem.eval_objective_and_gradient(xnn, name_nn)
        sigp = 2 * sig
        it_bisec, bisec_max = 1, 5
        while dot(gn, s) > theta * stg:  # :105-150
            if it_bisec > bisec_max:
                if sig > np.finfo(np.float64).eps:
                    sig, xn, gn, fn, cn, name = armijo
                    break
                raise RuntimeError("Seems we're stuck in bisection...")
            sigb = 0.5 * (sig + sigp)
            xb = axpby(1.0, xj, -sigb, s)
            name_b = _random_string(rng)
            fnn, gnn, cnn = problem.eval_objective_and_gradient(xb, name_b)
            if f - fnn >= delta * sigb * stg:
                sig, xn, fn, gn, cn, name = sigb, xb, fnn, gnn, cnn, name_b
            else:
                sigp = sigb
            it_bisec += 1
    return sig, {"m": xn, "gradient": gn, "objective": fn, "correlation": cn, "name": name}


def optlib_lbfgs(m0, options: Options, usr_par: UsrPar, problem=None, rng=None, on_iteration=None):
    """[flag, mfinal, usr_par] = optlib_lbfgs(m0, options, usr_par) (optlib_lbfgs.m:1-210).

    ``problem`` supplies eval_objective_and_gradient / eval_objective / eval_grad_objective
    (default: NoiseProblem(usr_par)). Models left in usr_par.model by an earlier run warm-start the
    curvature pairs (optlib_lbfgs.m:28-57). flag 0: ||g|| <= tolerance*||g0||; 1: max_iterations;
    2: misfit change over successive_iterations below successive_change. ``on_iteration(it,
    model)``, when given, runs after each new_iteration (it = 0 for the initial model), where the
    reference's user edits new_iteration.m for per-iteration work."""
    import torch
    rng = np.random.default_rng() if rng is None else rng
    problem = NoiseProblem(usr_par) if problem is None else problem
    log = _Log(options.output_file)
    it = 0
    m0 = _vec(m0)
    name = _random_string(rng)
    j0, g0, c0 = problem.eval_objective_and_gradient(m0, name)
    model = {"m": m0, "name": name, "objective": j0, "gradient": g0, "correlation": c0}
    normg0 = norm(g0)
    model["normg"] = normg0
    usr_par = new_iteration(it, model, usr_par)
    if on_iteration is not None:
        on_iteration(it, model)
    log(f"it={it}   j={j0:e}   ||g||={normg0:e}  \n")

    # optlib_restore_lbfgs_information.m:69-103
    nprev = len(usr_par.model) - 1
    prev = [usr_par.model[len(usr_par.model) - 1 - k] for k in range(min(nprev, 5), 0, -1)]
    n = m0.numel()
    dev = m0.device
    if prev:
        lh = len(prev)
        lmax = max(options.max_memory, lh)
        P = torch.zeros((lmax, n), dtype=torch.float64, device=dev)
        D = torch.zeros((lmax, n), dtype=torch.float64, device=dev)
        rho = np.zeros(lmax)
        for i in range(1, lh):
            axpby(1.0, prev[i - 1]["m"], -1.0, prev[i]["m"], out=P[i - 1])
            axpby(1.0, prev[i - 1]["gradient"], -1.0, prev[i]["gradient"], out=D[i - 1])
        axpby(1.0, prev[lh - 1]["m"], -1.0, model["m"], out=P[lh - 1])
        axpby(1.0, prev[lh - 1]["gradient"], -1.0, model["gradient"], out=D[lh - 1])
        for i in range(lh):
            rho[i] = 1.0 / dot(D[i], P[i])
        gak = dot(D[lh - 1], P[lh - 1]) / norm(D[lh - 1]) ** 2
        l = ln = min(lmax, lh)
    else:
        lmax = options.max_memory
        P = torch.zeros((lmax, n), dtype=torch.float64, device=dev)
        D = torch.zeros((lmax, n), dtype=torch.float64, device=dev)
        rho = np.zeros(lmax)
        l, ln, gak = 0, 1, 1.0
    ga = np.zeros(lmax)
    q = torch.empty(n, dtype=torch.float64, device=dev)
    r = torch.empty(n, dtype=torch.float64, device=dev)

    while model["normg"] > options.tolerance * normg0 and it < options.max_iterations:  # :76-195
        if it >= options.successive_iterations:
            if abs(usr_par.model[it]["objective"] - usr_par.model[it - options.successive_iterations]["objective"]) \
                    < options.successive_change:
                break
        it += 1
        g = model["gradient"]
        q.copy_(g)
        for jj in range(1, l + 1):  # :90-94
            i = (ln - jj - 1) % lmax
            ga[i] = rho[i] * dot(P[i], q)
            axpby(1.0, q, -ga[i], D[i], out=q)
        axpby(gak, q, 0.0, q, out=r)  # :97
        for jj in range(l, 0, -1):  # :101-105
            i = (ln - jj - 1) % lmax
            be = rho[i] * dot(D[i], r)
            axpby(1.0, r, ga[i] - be, P[i], out=r)
        s = r.clone()
        step = "LBFGS"
        stg = dot(s, g)
        if stg / (model["normg"] * norm(s)) < options.sufficient_decrease_angle:  # :110-116
            log("\tWarning: Bad angle of search direction. Use steepest descent instead.\n")
            s = g
            stg = dot(s, g)
            step = "Grad"
        sigma = options.init_step_length if it == 1 else 1.0
        sigma, model_new = optlib_wolfe(model["m"], s, stg, model["objective"], options.wolfe_delta,
                                        options.wolfe_theta, sigma, options.wolfe_try_to_increase_step_length,
                                        options.verbose, usr_par, problem, rng)
        gn = model_new["gradient"]
        model_new["normg"] = norm(gn)
        log(f"it={it}   j={model_new['objective']:e}   ||g||={model_new['normg']:e}   sigma={sigma:6.5f} "
            f"||s||={norm(s):e}  step={step}\n")
        d = axpby(1.0, g, -1.0, gn)  # :173-175
        p = axpby(1.0, model["m"], -1.0, model_new["m"])
        dtp = dot(d, p)
        if dtp >= 1e-8 * norm(d) * norm(p):  # :176-189
            rho[ln - 1] = 1.0 / dtp
            D[ln - 1].copy_(d)
            P[ln - 1].copy_(p)
            l = min(l + 1, lmax)
            ln = ln % lmax + 1
            gak = dtp / dot(d, d)
        else:
            log("\tWarning: Bad angle of search direction. Update of BFGS curvature information skipped.\n")
        model = model_new
        usr_par = new_iteration(it, model, usr_par)
        if on_iteration is not None:
            on_iteration(it, model)

    if model["normg"] <= options.tolerance * normg0:
        log(f"Successful termination with ||g||<{options.tolerance:e}*min(1,||g0||):\n")
        flag = 0
    elif it >= options.max_iterations:
        log("Maximum number of iterations reached.\n")
        flag = 1
    else:
        log(f"Improvement of misfit over {options.successive_iterations} iterations smaller than "
            f"{options.successive_change:f}.\n")
        flag = 2
    log.close()
    return flag, model["m"], usr_par

"""FP64 numpy restatement of the solvers' callers in inversion/ and tools/addons/ (TEST
INFRASTRUCTURE ONLY: the checker of fd2d_noise_amd/inversion.py and the caller kernels of
fd2d_noise_amd/csrc/caller.hip).

Restated from:
  tools/addons/gaussblur2d.m:2-38                    gaussblur2d (matrix form of the two loops)
  inversion/optlib_callback_fct/map_m_to_parameters.m:2-52, map_gradparameters_to_gradm.m:53-106
  inversion/addons/regularization.m:1-29
  inversion/optlib/optlib_lbfgs.m:1-210, optlib_wolfe.m:1-160, optlib_init_default_parameters_lbfgs.m
  inversion/optlib_callback_fct/eval_objective_and_gradient.m:2-156 (composition over stations)
Pinned by tests/test_oracle_inversion.py (closed forms, the dense BFGS update, the Wolfe
conditions, finite differences); nothing here is timed or shipped.
"""
from __future__ import annotations

import dataclasses
from typing import Callable, List, Optional

import numpy as np

from . import measurement
from .oracle import primitives

SQRT14 = np.sqrt(14.0)


# ---------------------------------------------------------------------------------------------
# gaussblur2d and the parameter map
# ---------------------------------------------------------------------------------------------
def _weights(x, sigma):
    d = np.abs(np.asarray(x, dtype=np.float64)[:, None] - np.asarray(x, dtype=np.float64)[None, :])
    return np.where(d < SQRT14 * sigma, np.exp(-(d / sigma) ** 2 / 2.0), 0.0)


def gaussblur2d(Y, x1, x2, sigma):
    """gaussblur2d.m:2-38: Mtmp = W1 Y, GY = Mtmp W2', divided by w1' * w2 where w1(j), w2(j)
    accumulate the weights of column j (lines 18, 29)."""
    W1 = _weights(x1, sigma[0])
    W2 = _weights(x2, sigma[1])
    GY = W1 @ np.asarray(Y, dtype=np.float64) @ W2.T
    return GY / np.outer(W1.sum(axis=0), W2.sum(axis=0))


def map_frame(config):
    """[ix, iz] = find(absbound == 1, 1, 'first') (map_m_to_parameters.m:18-19), 0-based."""
    ab = primitives.absbound(config)
    k = np.flatnonzero(ab.ravel(order="F") == 1.0)
    if k.size == 0:
        raise ValueError("no node with absbound == 1")
    return int(k[0] % config.nx), int(k[0] // config.nx)


def _blur_maps(config, P, sigma_source, sigma_structure, out):
    ix, iz = map_frame(config)
    x, z = config.x(), config.z()
    sx = slice(ix, config.nx - ix)
    sz = slice(iz, config.nz - iz)
    n_maps = P.shape[2]
    for q in range(n_maps):
        sg = sigma_source if q < n_maps - 1 else sigma_structure
        out[sx, sz, q] = gaussblur2d(P[sx, sz, q], x[sx], z[sz], sg)
    return out


def map_m_to_parameters(m, config, sigma_source=(1.0, 1.0), sigma_structure=(1.0, 1.0), ref_source=0.0,
                        ref_structure=1.0, mu_0=4.8e10):
    """map_m_to_parameters.m:16-49 with ring.switch 'no'; returns nx x nz x n_maps."""
    P = np.asarray(m, dtype=np.float64).reshape((config.nx, config.nz, -1), order="F").copy(order="F")
    _blur_maps(config, P.copy(order="F"), sigma_source, sigma_structure, P)
    P[:, :, :-1] = ref_source + P[:, :, :-1]
    P[:, :, -1] = mu_0 * (ref_structure + P[:, :, -1])
    return P


def map_gradparameters_to_gradm(grad_parameters, config, sigma_source=(1.0, 1.0), sigma_structure=(1.0, 1.0),
                                mu_0=4.8e10):
    """map_gradparameters_to_gradm.m:70-103 with ring.switch 'no'; returns the column vector."""
    G = np.asarray(grad_parameters, dtype=np.float64).reshape((config.nx, config.nz, -1), order="F")
    temp = np.asfortranarray(0.0 * G)
    _blur_maps(config, G, sigma_source, sigma_structure, temp)
    temp[:, :, -1] = mu_0 * temp[:, :, -1]
    return temp.ravel(order="F")


def regularization(j, g, m, m0, nx, nz, alpha=0.0, beta=0.0, weighting=1.0):
    """regularization.m:1-29; g None stands for the reference's scalar g (objective only)."""
    n = m.size
    ns = nx * nz
    w = np.atleast_1d(np.asarray(weighting, dtype=np.float64))
    w_src = np.tile(w, (n - ns) // w.size)
    w_str = w if w.size == ns else np.full(ns, w[0])
    d_src = m[:n - ns] - m0[:n - ns]
    d_str = m[n - ns:] - m0[n - ns:]
    j = j + alpha / 2 * np.sum(w_src * d_src ** 2)
    j = j + beta / 2 * np.sum(w_str * d_str ** 2)
    if g is not None:
        g = g.copy()
        g[:n - ns] += alpha * w_src * d_src
        g[n - ns:] += beta * w_str * d_str
    return j, g


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
    max_iterations: int = 100
    tolerance: float = 1e-3
    successive_change: float = 1.0
    successive_iterations: int = 5


@dataclasses.dataclass
class Problem:
    """The callbacks optlib_lbfgs / optlib_wolfe call: f(m) -> (j, g, c), j-only and g-only forms."""
    eval_objective_and_gradient: Callable
    eval_objective: Optional[Callable] = None
    eval_grad_objective: Optional[Callable] = None

    def fgc(self, m):
        return self.eval_objective_and_gradient(m)

    def fc(self, m):
        if self.eval_objective is not None:
            return self.eval_objective(m)
        j, _, c = self.eval_objective_and_gradient(m)
        return j, c

    def gc(self, m):
        if self.eval_grad_objective is not None:
            return self.eval_grad_objective(m)
        _, g, c = self.eval_objective_and_gradient(m)
        return g, c


def wolfe(xj, s, stg, f, delta, theta, sig0, try_larger_steps, problem: Problem, log: Optional[list] = None):
    """optlib_wolfe.m:19-159; returns (sigma, model dict)."""
    log = [] if log is None else log
    sig = sig0
    xn = xj - sig * s
    fn, gn, cn = problem.fgc(xn)
    while f - fn < delta * sig * stg:  # :33-48
        sig = 0.5 * sig
        xn = xj - sig * s
        fn, cn = problem.fc(xn)
    if sig != sig0:  # :54-56
        gn, cn = problem.gc(xn)
    wolfe_ok = gn @ s <= theta * stg
    armijo = (sig, xn, gn, fn, cn)
    if (not wolfe_ok) or try_larger_steps:  # :76
        if sig == sig0:
            xnn = xj - 2 * sig * s
            fnn, gnn, cnn = problem.fgc(xnn)
            while (gn @ s > theta * stg) and (f - fnn >= 2 * delta * sig * stg):  # :85-99
                sig = 2 * sig
                xn, fn, gn, cn = xnn, fnn, gnn, cnn
                xnn = xj - 2 * sig * s
                fnn, gnn, cnn = problem.fgc(xnn)
        sigp = 2 * sig
        it_bisec, bisec_max = 1, 5
        while gn @ s > theta * stg:  # :105-150
            if it_bisec > bisec_max:
                if sig > np.finfo(np.float64).eps:
                    sig, xn, gn, fn, cn = armijo
                    log.append("bisection limit: Armijo step kept")
                    break
                raise RuntimeError("Seems we're stuck in bisection...")
            sigb = 0.5 * (sig + sigp)
            xb = xj - sigb * s
            fnn, gnn, cnn = problem.fgc(xb)
            if f - fnn >= delta * sigb * stg:
                sig, xn, fn, gn, cn = sigb, xb, fnn, gnn, cnn
            else:
                sigp = sigb
            it_bisec += 1
    return sig, {"m": xn, "gradient": gn, "objective": fn, "correlation": cn}


def two_loop(g, P, D, rho, gak, l, ln, lmax):
    """s = H g by the two-loop recursion of optlib_lbfgs.m:88-106 (1-based ring indices ln)."""
    q = g.copy()
    ga = np.zeros(lmax)
    for jj in range(1, l + 1):
        i = (ln - jj - 1) % lmax
        ga[i] = rho[i] * (P[:, i] @ q)
        q = q - ga[i] * D[:, i]
    r = gak * q
    for jj in range(l, 0, -1):
        i = (ln - jj - 1) % lmax
        be = rho[i] * (D[:, i] @ r)
        r = r + (ga[i] - be) * P[:, i]
    return r


def new_iteration(it, model, history):
    """usr_par.model(it+1) = model (new_iteration.m:29): entry it is replaced or appended."""
    if it < len(history):
        history[it] = dict(model)
    else:
        history.append(dict(model))


def lbfgs(m0, options: Options, problem: Problem, history: Optional[List[dict]] = None, record=None):
    """optlib_lbfgs.m:1-210. history: the usr_par.model list new_iteration.m fills; entries left
    by an earlier run warm-start the curvature pairs (optlib_lbfgs.m:28-57).
    Returns (flag, m_final, history); record(event, data) sees every iteration."""
    history = [] if history is None else history
    it = 0
    j0, g0, c0 = problem.fgc(m0)
    model = {"m": m0, "objective": j0, "gradient": g0, "correlation": c0}
    normg0 = np.linalg.norm(g0)
    model["normg"] = normg0
    new_iteration(it, model, history)
    # optlib_restore_lbfgs_information.m:69-103: up to 5 stored models before the last one
    nprev = len(history) - 1
    prev = [history[len(history) - 1 - k] for k in range(min(nprev, 5), 0, -1)]
    n = m0.size
    if prev:
        lh = len(prev)
        lmax = max(options.max_memory, lh)
        P = np.zeros((n, lmax))
        D = np.zeros((n, lmax))
        rho = np.zeros(lmax)
        for i in range(1, lh):
            P[:, i - 1] = prev[i - 1]["m"] - prev[i]["m"]
            D[:, i - 1] = prev[i - 1]["gradient"] - prev[i]["gradient"]
        P[:, lh - 1] = prev[lh - 1]["m"] - model["m"]
        D[:, lh - 1] = prev[lh - 1]["gradient"] - model["gradient"]
        for i in range(lh):
            rho[i] = 1.0 / (D[:, i] @ P[:, i])
        gak = (D[:, lh - 1] @ P[:, lh - 1]) / np.linalg.norm(D[:, lh - 1]) ** 2
        l = ln = min(lmax, lh)
    else:
        lmax = options.max_memory
        P = np.zeros((n, lmax))
        D = np.zeros((n, lmax))
        rho = np.zeros(lmax)
        l, ln, gak = 0, 1, 1.0
    while model["normg"] > options.tolerance * normg0 and it < options.max_iterations:
        if it >= options.successive_iterations:
            if abs(history[it]["objective"] - history[it - options.successive_iterations]["objective"]) \
                    < options.successive_change:
                break
        it += 1
        g = model["gradient"]
        s = two_loop(g, P, D, rho, gak, l, ln, lmax)
        step = "LBFGS"
        stg = s @ g
        if stg / (model["normg"] * np.linalg.norm(s)) < options.sufficient_decrease_angle:
            s = g
            stg = s @ g
            step = "Grad"
        sigma = options.init_step_length if it == 1 else 1.0
        sigma, new = wolfe(model["m"], s, stg, model["objective"], options.wolfe_delta, options.wolfe_theta, sigma,
                           options.wolfe_try_to_increase_step_length, problem)
        new["normg"] = np.linalg.norm(new["gradient"])
        if record:
            record("iteration", {"it": it, "objective": new["objective"], "normg": new["normg"], "sigma": sigma,
                                 "step": step})
        d = g - new["gradient"]
        p = model["m"] - new["m"]
        dtp = d @ p
        if dtp >= 1e-8 * np.linalg.norm(d) * np.linalg.norm(p):
            rho[ln - 1] = 1.0 / dtp
            D[:, ln - 1] = d
            P[:, ln - 1] = p
            l = min(l + 1, lmax)
            ln = ln % lmax + 1
            gak = dtp / (d @ d)
        model = new
        new_iteration(it, model, history)
    if model["normg"] <= options.tolerance * normg0:
        flag = 0
    elif it >= options.max_iterations:
        flag = 1
    else:
        flag = 2
    return flag, model["m"], history


# ---------------------------------------------------------------------------------------------
# eval_objective_and_gradient composition
# ---------------------------------------------------------------------------------------------
def receivers_of(array, src_row):
    """rec = array(find(~ismember(array, src, 'rows')), :) (eval_objective_and_gradient.m:40)."""
    keep = ~np.all(array == src_row[None, :], axis=1)
    return array[keep]


def eval_objective_and_gradient(m, usr, solvers, only_objective=False, g_cache=None):
    """eval_objective_and_gradient.m:2-156 on the oracle solvers, FP64 host arrays.

    ``usr`` is an fd2d_noise_amd.inversion.UsrPar (the usr_par fields the function reads)."""
    c = usr.config
    P = map_m_to_parameters(m, c, usr.sigma_source, usr.sigma_structure, usr.ref_source, usr.ref_structure,
                            usr.mu_0)
    source_dist = np.asfortranarray(P[:, :, :-1])
    spectrum = usr.spectrum
    mu = np.asfortranarray(P[:, :, -1])
    rho = usr.rho
    t = usr.t
    nt = len(t)
    ref = np.atleast_2d(usr.ref_stat)
    n_ref = ref.shape[0]
    n_rec = usr.array.shape[0] - 1
    c_it = np.zeros((n_ref, n_rec, nt))
    grad = np.zeros_like(P)
    j = j_src = j_str = 0.0
    w = usr.kernel_weighting
    for i in range(n_ref):
        src = ref[i:i + 1]
        rec = receivers_of(usr.array, ref[i])
        if usr.type == "source":
            if g_cache is not None and i in g_cache:
                G_fft = g_cache[i]
            else:
                G_fft, _, _ = solvers.run_forward1_green(mu, rho, src, None, 0, None, None, config=c)
                if g_cache is not None:
                    g_cache[i] = G_fft
            cc, _ = solvers.run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, 0, None, None,
                                                     config=c)
        else:
            G_fft, G, _ = solvers.run_forward1_green(mu, rho, src, None, 1, None, None, config=c)
            cc, C = solvers.run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, 1, None, None,
                                                     config=c)
        c_it[i] = cc
        c_data = usr.c_data[i * n_rec:(i + 1) * n_rec, :]
        jn_s, a_s = measurement.make_adjoint_sources(cc, c_data, 0 * c_data, t, usr.veldis, usr.measurement_source,
                                                     "1st", usr.window)
        jn_t, a_t = measurement.make_adjoint_sources(cc, c_data, 0 * c_data, t, usr.veldis,
                                                     usr.measurement_structure, "1st", usr.window)
        j_n = (1.0 - w) * jn_s + w * jn_t
        adstf = (1.0 - w) * a_s + w * a_t
        if not only_objective:
            stf = np.asfortranarray(adstf.astype(np.complex128))
            if usr.type == "source":
                K1, _, _, _ = solvers.run_noise_adjoint(mu, rho, None, stf, rec, None, spectrum, None, G_fft, 0, None,
                                                        None, config=c)
                K2 = 0.0 * K1
            elif usr.type == "structure":
                K1, A1, _, _ = solvers.run_noise_adjoint(mu, rho, C, stf, rec, None, None, None, None, 1, None, None,
                                                         config=c)
                K2, _, _, _ = solvers.run_noise_adjoint(mu, rho, G, A1, rec, None, spectrum, source_dist, None, 0,
                                                        None, None, config=c)
            else:
                K1, A1, _, _ = solvers.run_noise_adjoint(mu, rho, C, stf, rec, None, spectrum, None, G_fft, 1, None,
                                                         None, config=c)
                K2, _, _, _ = solvers.run_noise_adjoint(mu, rho, G, A1, rec, None, spectrum, source_dist, None, 0,
                                                        None, None, config=c)
            grad = grad + K1.reshape(grad.shape, order="F") + K2.reshape(grad.shape, order="F")
        j = j + np.sum(j_n)
        j_src = j_src + np.sum(jn_s)
        j_str = j_str + np.sum(jn_t)
    c_all = c_it.reshape(n_ref * n_rec, nt)
    m0 = np.asarray(usr.m0, dtype=np.float64) if usr.m0 is not None else m
    if only_objective:
        j, _ = regularization(j, None, m, m0, c.nx, c.nz, usr.reg_alpha, usr.reg_beta, usr.reg_weighting)
        return j, 0, c_all
    g = map_gradparameters_to_gradm(grad, c, usr.sigma_source, usr.sigma_structure, usr.mu_0)
    j, g = regularization(j, g, m, m0, c.nx, c.nz, usr.reg_alpha, usr.reg_beta, usr.reg_weighting)
    return j, g, c_all


def general_source(noise_spectrum, noise_source_distribution, n_basis_fct):
    """general_source.m:10-24: band sums of spectrum(k, ns) * distribution(:, :, ns) / n_i over the
    bands of integration_limits.m:4-12."""
    spec = np.atleast_2d(np.asarray(noise_spectrum, dtype=np.float64).T).T
    K, n_src = spec.shape
    dist = np.asarray(noise_source_distribution, dtype=np.float64)
    if dist.ndim == 2:
        dist = dist[:, :, None]
    out = np.zeros(dist.shape[:2] + (n_basis_fct,), order="F")
    step = K / n_basis_fct
    for ib in range(1, n_basis_fct + 1):
        lo, hi = int(np.floor((ib - 1) * step + 1)), int(np.floor(ib * step))
        ni = hi - lo + 1
        for k in range(lo, hi + 1):
            for ns in range(n_src):
                out[:, :, ib - 1] += spec[k - 1, ns] * dist[:, :, ns] / ni
    return out


def eval_hessian_vector_product(m, dm_parameters, usr, solvers):
    """eval_hessian_vector_product_2.m:2-197 on the oracle solvers: H_m for the model m and the
    parameter perturbation dm_parameters (nx x nz x n_maps, or only its source / structure maps
    for usr.type 'source' / 'structure', :19-31). Needs fw_nth = 1 (the Born inputs are the saved
    wavefields of every step)."""
    c = usr.config
    P = map_m_to_parameters(m, c, usr.sigma_source, usr.sigma_structure, usr.ref_source, usr.ref_structure,
                            usr.mu_0)
    mu = np.asfortranarray(P[:, :, -1])
    rho = usr.rho
    sd_ref = np.asfortranarray(P[:, :, :-1])
    spectrum = usr.spectrum
    dmp = np.asarray(dm_parameters, dtype=np.float64)
    if dmp.size != P.size:  # :19-31
        dm = np.zeros_like(P)
        if usr.type == "source":
            dm[:, :, :-1] = dmp.reshape(P[:, :, :-1].shape, order="F")
        elif usr.type == "structure":
            dm[:, :, -1] = dmp.reshape(P.shape[:2], order="F")
    else:
        dm = dmp.reshape(P.shape, order="F")
    sd_pert = np.asfortranarray(dm[:, :, :-1])
    dmu = np.asfortranarray(dm[:, :, -1])
    has_dmu = np.any(dmu != 0)
    has_dsd = np.any(sd_pert != 0)
    t = usr.t
    ref = np.atleast_2d(usr.ref_stat)
    n_rec = usr.array.shape[0] - 1
    w = usr.kernel_weighting
    H = np.zeros_like(P)
    for i in range(ref.shape[0]):
        src = ref[i:i + 1]
        rec = receivers_of(usr.array, ref[i])
        c_data = usr.c_data[i * n_rec:(i + 1) * n_rec, :]
        G_fft, G, _ = solvers.run_forward1_green(mu, rho, src, None, 1, None, None, config=c)  # :57
        dG_fft = dG = None
        if has_dmu:  # :60-65
            dG_fft, dG, _ = solvers.run_forward1_green(mu, rho, src, None, 5, dmu, G, config=c)
        c_ref, C = solvers.run_forward2_correlation(mu, rho, G_fft, spectrum, sd_ref, rec, 1, None, None, config=c)
        c_pert = np.zeros_like(c_ref)
        dC = None
        if has_dmu:  # :74-80
            c1, dC = solvers.run_forward2_correlation(mu, rho, dG_fft, spectrum, sd_ref, rec, 1, dmu, C, config=c)
            c_pert = c_pert + c1
        if has_dsd:  # :83-89
            c2, dC2 = solvers.run_forward2_correlation(mu, rho, G_fft, spectrum, sd_pert, rec, 1, None, None, config=c)
            c_pert = c_pert + c2
            dC = dC2 if dC is None else dC + dC2
        a = {}
        for order in ("1st", "2nd"):  # :99-105
            _, a_s = measurement.make_adjoint_sources(c_ref, c_data, c_pert, t, usr.veldis, usr.measurement_source,
                                                      order, usr.window)
            _, a_t = measurement.make_adjoint_sources(c_ref, c_data, c_pert, t, usr.veldis,
                                                      usr.measurement_structure, order, usr.window)
            a[order] = np.asfortranarray(((1.0 - w) * a_s + w * a_t).astype(np.complex128))
        K1 = K2 = K4 = 0.0
        u_dag = uT = p = None
        if usr.type != "source":  # YELLOW :113-121
            K1, uT, u_dag, _ = solvers.run_noise_adjoint(mu, rho, dC, a["1st"], rec, None, spectrum, None, dG_fft, 13,
                                                         None, None, config=c)
            if usr.type == "structure":
                K1[:, :, :-1] = 0.0
        if usr.type != "source" and has_dmu:  # GRAY :129-133
            K2, _, p, _ = solvers.run_noise_adjoint(mu, rho, dG, uT, rec, None, spectrum, sd_ref, None, 11, None, None,
                                                    config=c)
        if usr.type == "joint" and has_dsd:  # :141-146
            K2a, _, pa, _ = solvers.run_noise_adjoint(mu, rho, G, uT, rec, None, spectrum, sd_pert, None, 11, None, None,
                                                      config=c)
            K2 = K2 + K2a
            p = pa if p is None else p + pa
        # BLACK :154-160
        K3, kT, _, _ = solvers.run_noise_adjoint(mu, rho, C, a["2nd"], rec, None, spectrum, None, G_fft, 12,
                                                 dmu if u_dag is not None else None, u_dag, config=c)
        if usr.type == "source":
            K3[:, :, -1] = 0.0
        elif usr.type == "structure":
            K3[:, :, :-1] = 0.0
        if usr.type != "source":  # RED :168-170
            K4, _, _, _ = solvers.run_noise_adjoint(mu, rho, G, kT, rec, None, spectrum, sd_ref, None, 10,
                                                    dmu if p is not None else None, p, config=c)
        H = H + K1 + K2 + K3 + K4
    Hm = map_gradparameters_to_gradm(H, c, usr.sigma_source, usr.sigma_structure, usr.mu_0)  # :187
    # :191-196: the regularisation term adds alpha / beta * weighting on the interior (not times dm)
    inner = (primitives.absbound(c).ravel(order="F") == 1.0).astype(np.float64)
    ns = c.nx * c.nz
    wv = np.atleast_1d(np.asarray(usr.reg_weighting, dtype=np.float64))
    w_src = np.tile(wv, (Hm.size - ns) // wv.size)
    Hm[:-ns] = Hm[:-ns] + np.tile(inner, (Hm.size - ns) // ns) * usr.reg_alpha * w_src
    Hm[-ns:] = Hm[-ns:] + inner * usr.reg_beta * (wv if wv.size == ns else wv[0])
    return Hm

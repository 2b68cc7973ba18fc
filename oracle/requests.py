"""FP64 oracle of the serving requests (TEST INFRASTRUCTURE ONLY: the checker of the served results).

Each function restates one request program on the oracle solvers, with host arrays in MATLAB
layout, and returns its outputs in MATLAB layout:

  * correlation: run_forward1_green (mode 0) then run_forward2_correlation (mode 0) for one
    reference station (inversion/optlib_callback_fct/eval_objective_and_gradient.m:57-72 without
    the stored wavefields);
  * kernel: the same with stored wavefields, adjoint sources of the misfit, run_noise_adjoint mode 1
    against the correlation wavefield and mode 0 against the Green wavefield with the transformed
    adjoint state; K = K1 + K2 (eval_objective_and_gradient.m:57-98 for one station);
  * hvp: the structure Hessian-vector product of one station
    (eval_hessian_vector_product_2.m:57-170 with type 'structure' and dm on mu only).
"""
import numpy as np

from . import measurement


def _stf(corr, observed, du, t, veldis, meas, order, window):
    _, a = measurement.make_adjoint_sources(corr, observed, du, t, veldis, meas, order, window)
    return np.asfortranarray(a.astype(np.complex128))


def correlation(orc, c, mu, rho, spectrum, source_dist, src, rec):
    G, _, _ = orc.run_forward1_green(mu, rho, src, None, 0, None, None, config=c)
    corr, _ = orc.run_forward2_correlation(mu, rho, G, spectrum, source_dist, rec, 0, None, None, config=c)
    return {"correlations": corr}


def kernel(orc, c, mu, rho, spectrum, source_dist, src, rec, observed, measurement_="waveform_difference",
           veldis="dis", window=(100.0, 200.0)):
    t = c.t_corr()
    G, Gout, _ = orc.run_forward1_green(mu, rho, src, None, 1, None, None, config=c)           # :57
    corr, C = orc.run_forward2_correlation(mu, rho, G, spectrum, source_dist, rec, 1, None, None, config=c)  # :71
    stf = _stf(corr, observed, None, t, veldis, measurement_, "1st", window)                    # :78-82
    K1, A1, _, _ = orc.run_noise_adjoint(mu, rho, C, stf, rec, None, None, None, None, 1, None, None, config=c)  # :88
    K2, _, _, _ = orc.run_noise_adjoint(mu, rho, Gout, A1, rec, None, spectrum, source_dist, None, 0, None, None,
                                        config=c)                                              # :93
    return {"correlations": corr, "K": K1 + K2}


def hvp(orc, c, mu, rho, spectrum, source_dist, src, rec, observed, dmu, measurement_="waveform_difference",
        veldis="dis", window=(100.0, 200.0)):
    t = c.t_corr()
    G, Gout, _ = orc.run_forward1_green(mu, rho, src, None, 1, None, None, config=c)           # :57
    dG, dGout, _ = orc.run_forward1_green(mu, rho, src, None, 5, dmu, Gout, config=c)          # :61
    c_ref, C = orc.run_forward2_correlation(mu, rho, G, spectrum, source_dist, rec, 1, None, None, config=c)  # :71
    c_pert, dC = orc.run_forward2_correlation(mu, rho, dG, spectrum, source_dist, rec, 1, dmu, C, config=c)  # :75
    stf1 = _stf(c_ref, observed, c_pert, t, veldis, measurement_, "1st", window)               # :99-105
    stf2 = _stf(c_ref, observed, c_pert, t, veldis, measurement_, "2nd", window)
    KY, uT, u_dag, _ = orc.run_noise_adjoint(mu, rho, dC, stf1, rec, None, spectrum, None, dG, 13, None, None,
                                             config=c)                                         # YELLOW :113
    KG, _, p, _ = orc.run_noise_adjoint(mu, rho, dGout, uT, rec, None, spectrum, source_dist, None, 11, None, None,
                                        config=c)                                              # GRAY :129
    KB, kT, _, _ = orc.run_noise_adjoint(mu, rho, C, stf2, rec, None, spectrum, None, G, 12, dmu, u_dag,
                                         config=c)                                             # BLACK :154
    KR, _, _, _ = orc.run_noise_adjoint(mu, rho, Gout, kT, rec, None, spectrum, source_dist, None, 10, dmu, p,
                                        config=c)                                              # RED :168
    KY[:, :, :-1] = 0.0  # type 'structure': only the structure map of YELLOW and BLACK (:120, :157-159)
    KB[:, :, :-1] = 0.0
    return {"correlations": c_ref, "H": KY + KG + KB + KR}

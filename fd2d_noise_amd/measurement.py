"""Misfits and adjoint source time functions of the caller side (tools/measurement/), FP64 on the host.

These turn correlation traces into the point adjoint sources that run_noise_adjoint consumes
(inversion/optlib_callback_fct/eval_objective_and_gradient.m:76-82). Restated from:
  make_adjoint_sources.m:60-175 (windows [100, 200] s, causal + acausal branch, 'dis'/'vel'),
  cc_time_shift.m:21-59, waveform_difference.m:23-48, addons/get_window.m:2-39 ('cos_taper'),
  addons/cross_correlation_td.m:7-17 (xcorr over lags -(N-1)..N-1).
The lag correlation is evaluated with an FFT; its argmax is the brute-force xcorr's.
"""
from __future__ import annotations

import numpy as np


def get_window(t, t_min, t_max, kind="cos_taper"):
    t = np.asarray(t, dtype=np.float64)
    if t_min > t_max:
        t_min, t_max = t_max, t_min
    win = np.ones_like(t)
    if kind == "box":
        return win * (t > t_min) * (t < t_max)
    if kind == "hann":
        win = win * (t > t_min) * (t < t_max)
        nz = np.flatnonzero(win)
        if nz.size:
            i1, i2 = nz[0], nz[-1]
            n = i2 - i1 + 1
            win[i1:i2 + 1] = 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(1, n + 1) / (n + 1))
        return win
    width = abs(t_max - t_min) * 0.1
    win = win * (t > t_min) * (t < t_max)
    win = win + (0.5 + 0.5 * np.cos(np.pi * (t_max - t) / width)) * (t >= t_max) * (t < t_max + width)
    win = win + (0.5 + 0.5 * np.cos(np.pi * (t_min - t) / width)) * (t > t_min - width) * (t <= t_min)
    return win


def xcorr(x, y):
    """MATLAB xcorr(x, y): R(m) = sum_n x(n+m) y(n), m = -(N-1)..N-1."""
    n = len(x)
    L = 1 << int(np.ceil(np.log2(2 * n - 1)))
    r = np.fft.irfft(np.fft.rfft(x, L) * np.conj(np.fft.rfft(y, L)), L)
    return np.concatenate([r[L - (n - 1):], r[:n]])


def cc_time_shift(u, u0, win, t):
    dt = abs(t[1] - t[0])
    if np.all(u0 == 0):
        T = 1.0
    else:
        cc = xcorr(win * u0, win * u)  # cross_correlation_td(win.*u, win.*u_0) = xcorr(g, f)
        t_cc = (np.arange(len(cc)) - (len(t) - 1)) * dt
        T = t_cc[int(np.argmax(cc))]
        if abs(T) > 3.5:
            T = 0.0
    misfit = T ** 2 / 2.0
    v = np.zeros_like(u)
    v[:-1] = np.diff(u) / dt
    adstf = T * (win ** 2 * v) / (np.sum((win * v) ** 2) * dt)
    return misfit, adstf[::-1].copy()


def waveform_difference(u, u0, du, win, t, deriv_order="1st"):
    dt = abs(t[1] - t[0])
    if deriv_order == "1st":
        misfit = 0.5 * np.sum(win ** 2 * (u - u0) ** 2) * dt
        adstf = win ** 2 * (u - u0) * dt
    else:
        misfit = None
        adstf = win ** 2 * du * dt
    return misfit, adstf[::-1].copy()


def make_adjoint_sources(u, u0, du, t, veldis="dis", measurement="cc_time_shift", deriv_order="1st",
                         window=(100.0, 200.0)):
    """Misfit per receiver and adjoint stf (n_rec x nt) as make_adjoint_sources.m:31-175."""
    u = np.atleast_2d(np.asarray(u, dtype=np.float64))
    u0 = np.atleast_2d(np.asarray(u0, dtype=np.float64))
    du = np.zeros_like(u) if du is None else np.atleast_2d(np.asarray(du, dtype=np.float64))
    t = np.asarray(t, dtype=np.float64)
    nr, nt = u.shape
    dt = t[1] - t[0]
    misfit = np.zeros(nr)
    adstf = np.zeros((nr, nt))
    for n in range(nr):
        if measurement == "waveform_difference":
            left, right = t[0], t[-1]
        else:
            left, right = window
            if left < 0:
                left = t[np.flatnonzero(t == 0)[0] + 1]
            right = min(right, t[-1])
            if left > t[-1]:
                raise ValueError("time series too short?")
        win = get_window(t, left, right, "cos_taper")
        if measurement == "waveform_difference":
            m, a = waveform_difference(u[n], u0[n], du[n], win, t, deriv_order)
            misfit[n] = m if m is not None else 0.0
            adstf[n] = a
        elif measurement == "cc_time_shift":
            if deriv_order == "2nd":
                raise ValueError("no Hessian vector products for traveltime measurements")
            m1, a1 = cc_time_shift(u[n], u0[n], win, t)
            win2 = get_window(t, -right, -left, "cos_taper")
            m2, a2 = cc_time_shift(u[n], u0[n], win2, t)
            misfit[n] = m1 + m2
            adstf[n] = a1 + a2
        else:
            raise ValueError(f"measurement {measurement!r} is not restated")
        if veldis == "vel":
            adstf[n, :nt - 1] = -np.diff(adstf[n]) / dt
    return misfit, adstf

"""FP64 numpy restatement of tools/measurement/ (TEST INFRASTRUCTURE ONLY: the checker of the
device implementation fd2d_make_adjoint_sources in fd2d_noise_amd/csrc/caller.hip).

Restated from make_adjoint_sources.m:31-175 (windows, causal + acausal branches, 'dis'/'vel'),
waveform_difference.m:23-48, cc_time_shift.m:21-59, amp_diff.m:23-67, log_amp_ratio.m:23-93,
addons/get_window.m:2-39 ('cos_taper', 'hann', 'box') and addons/cross_correlation_td.m:7-17.
MATLAB's xcorr(x, y) is R(m) = sum_n x(n+m) y(n), m = -(N-1)..N-1 (numpy.correlate 'full');
hann(L) is the symmetric Hann window 0.5(1 - cos(2 pi n/(L-1))); trapz(y) uses unit spacing.
The Signal Processing Toolbox (MATLAB r2015a) is not available here; xcorr and hann are restated
from their textbook definitions, pinned by the tests in tests/test_oracle_measurement.py.
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
            win[i1:i2 + 1] = 1.0 if n == 1 else 0.5 * (1 - np.cos(2 * np.pi * np.arange(n) / (n - 1)))
        return win
    width = abs(t_max - t_min) * 0.1
    win = win * (t > t_min) * (t < t_max)
    win = win + (0.5 + 0.5 * np.cos(np.pi * (t_max - t) / width)) * (t >= t_max) * (t < t_max + width)
    win = win + (0.5 + 0.5 * np.cos(np.pi * (t_min - t) / width)) * (t > t_min - width) * (t <= t_min)
    return win


def xcorr(x, y):
    """MATLAB xcorr(x, y): R(m) = sum_n x(n+m) y(n), m = -(N-1)..N-1 (direct sum)."""
    return np.correlate(np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), "full")


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


def amp_diff(u, u0, du, win, t, deriv_order="1st"):
    dt = abs(t[1] - t[0])
    eps = np.finfo(np.float64).eps
    energy_0 = np.sum((win * u0) ** 2) * dt
    meas = (np.sum((win * u) ** 2) * dt - energy_0) / (energy_0 + eps)
    if deriv_order == "1st":
        if np.all(u0 == 0):
            misfit = 1.0
            adstf = 2 * u[::-1] / (np.sum(u ** 2) * dt)
        else:
            misfit = meas ** 2 / 2
            adstf = 2 * meas * win ** 2 * u / np.sum((win * u0) ** 2)
    else:
        misfit = None
        adstf = (2 * meas * win ** 2 * du / energy_0 * dt
                 + 4 * win ** 2 * u / energy_0 ** 2 * np.sum(win ** 2 * u * du) * dt ** 2)
    return misfit, adstf[::-1].copy()


def trapz(y):
    y = np.asarray(y, dtype=np.float64)
    return float(np.sum(y) - 0.5 * (y[0] + y[-1])) if len(y) > 1 else 0.0


def log_amp_ratio(u, u0, du, win_caus, t, deriv_order="1st"):
    dt = abs(t[1] - t[0])
    eps = np.finfo(np.float64).eps
    index = int(np.argmax(win_caus))
    if t[index] < 0:
        win_acaus = win_caus
        win_caus = win_acaus[::-1]
    else:
        win_acaus = win_caus[::-1]
    e_caus = trapz((win_caus * u) ** 2) * dt
    e_acaus = trapz((win_acaus * u) ** 2) * dt + eps
    e0_caus = trapz((win_caus * u0) ** 2) * dt
    e0_acaus = trapz((win_acaus * u0) ** 2) * dt + eps
    A = np.log(e_caus / e_acaus)
    A0 = np.log(e0_caus / e0_acaus)
    if deriv_order == "1st":
        misfit = 0.5 * (A - A0) ** 2
        de_caus = 2 * win_caus ** 2 * u * dt
        de_acaus = 2 * win_acaus ** 2 * u * dt
        if np.all(u0 == 0):
            adstf = de_caus / e_caus - de_acaus / e_acaus
        else:
            adstf = (A - A0) * (de_caus / e_caus - de_acaus / e_acaus)
    else:
        misfit = None
        de_caus = trapz(2 * win_caus ** 2 * u * du) * dt
        de_acaus = trapz(2 * win_acaus ** 2 * u * du) * dt
        adstf = ((A - A0) * (2 / e_caus * win_caus ** 2 - 2 / e_acaus * win_acaus ** 2) * du * dt
                 + (A - A0) * (-2 / e_caus ** 2 * win_caus ** 2 * u * de_caus
                               + 2 / e_acaus ** 2 * win_acaus ** 2 * u * de_acaus) * dt
                 + (2 / e_caus * win_caus ** 2 * u - 2 / e_acaus * win_acaus ** 2 * u)
                 * (de_caus / e_caus - de_acaus / e_acaus) * dt)
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
    dt = abs(t[1] - t[0])
    if veldis == "vel":  # :40-52
        v = np.zeros_like(u)
        v0 = np.zeros_like(u0)
        v[:, :nt - 1] = np.diff(u, axis=1) / dt
        v0[:, :nt - 1] = np.diff(u0, axis=1) / dt
        u, u0 = v, v0
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
        elif measurement in ("cc_time_shift", "amplitude_difference"):
            if measurement == "cc_time_shift" and deriv_order == "2nd":
                raise ValueError("no Hessian vector products for traveltime measurements")
            f = cc_time_shift if measurement == "cc_time_shift" else None
            if f is not None:
                m1, a1 = cc_time_shift(u[n], u0[n], win, t)
            else:
                m1, a1 = amp_diff(u[n], u0[n], du[n], win, t, deriv_order)
            win2 = get_window(t, -right, -left, "cos_taper")
            if f is not None:
                m2, a2 = cc_time_shift(u[n], u0[n], win2, t)
            else:
                m2, a2 = amp_diff(u[n], u0[n], du[n], win2, t, deriv_order)
            if deriv_order == "1st":
                misfit[n] = m1 + m2
            adstf[n] = a1 + a2
        elif measurement == "log_amplitude_ratio":
            win = get_window(t, left, right, "hann")
            m, a = log_amp_ratio(u[n], u0[n], du[n], win, t, deriv_order)
            misfit[n] = m if m is not None else 0.0
            adstf[n] = a
        else:
            raise ValueError(f"measurement {measurement!r}")
        if veldis == "vel":  # :166-168
            adstf[n, :nt - 1] = -np.diff(adstf[n]) / dt
    return misfit, adstf

"""Pins of the FP64 measurement oracle (oracle/measurement.py) on the CPU.

The reference holds no golden vectors for tools/measurement/; the oracle is pinned by what the
mathematics fixes: the lag convention stated in cross_correlation_td.m:5 (cc_i = sum_j f(j) g(j+i)),
known integer delays, the textbook symmetric Hann window and trapezoid rule, and finite differences
of each smooth misfit, whose gradient the (time-reversed) adjoint source must be.
"""
import numpy as np
import pytest

from oracle import measurement as M

NT = 601
T_AXIS = np.arange(NT) - 300.0  # correlation time axis, dt = 1 s


def _wavelet(shift=0.0, amp=1.0):
    a = np.abs(T_AXIS) - 150.0 - shift
    return amp * np.exp(-(a / 12.0) ** 2) * np.cos(0.2 * a)


def _traces(seed=0):
    rng = np.random.default_rng(seed)
    u = _wavelet(1.3, 1.1) + 0.01 * rng.standard_normal(NT)
    u0 = _wavelet(0.0)
    du = rng.standard_normal(NT) * np.exp(-((np.abs(T_AXIS) - 150.0) / 30.0) ** 2)
    return u, u0, du


def test_xcorr_lag_convention_brute_force():
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(7), rng.standard_normal(7)
    n = len(x)
    brute = [sum(x[i + m] * y[i] for i in range(n) if 0 <= i + m < n) for m in range(-(n - 1), n)]
    np.testing.assert_allclose(M.xcorr(x, y), brute, rtol=0, atol=1e-14)
    # a delta at 5 correlated with a delta at 2 peaks at lag +3 (index 3 + n - 1)
    e5, e2 = np.eye(n)[5], np.eye(n)[2]
    assert int(np.argmax(M.xcorr(e5, e2))) == 3 + n - 1


@pytest.mark.parametrize("delay", [-3, -1, 2, 3])
def test_cc_time_shift_recovers_integer_delay(delay):
    """u = u_0 delayed by d samples: cc_i = sum_j u(j) u_0(j+i) peaks at i = -d, so T = -d dt."""
    u0 = _wavelet()
    u = np.roll(u0, delay)
    win = M.get_window(T_AXIS, 100.0, 200.0)
    misfit, _ = M.cc_time_shift(u, u0, win, T_AXIS)
    assert misfit == pytest.approx((delay * 1.0) ** 2 / 2)
    # both branches of a symmetric delay: the acausal branch moves the other way in t
    uc = u0.copy()
    uc[T_AXIS > 0] = np.roll(u0, delay)[T_AXIS > 0]
    uc[T_AXIS < 0] = np.roll(u0, -delay)[T_AXIS < 0]
    m, _ = M.make_adjoint_sources(uc, u0, None, T_AXIS, "dis", "cc_time_shift")
    assert m[0] == pytest.approx(delay ** 2)


def test_cc_time_shift_threshold_and_data_independent():
    u0 = _wavelet()
    win = M.get_window(T_AXIS, 100.0, 200.0)
    m, _ = M.cc_time_shift(np.roll(u0, 4), u0, win, T_AXIS)  # |T| = 4 s > 3.5 s -> T = 0
    assert m == 0.0
    m, a = M.cc_time_shift(u0, np.zeros(NT), win, T_AXIS)  # u_0 == 0 -> T = 1
    assert m == 0.5 and np.any(a != 0)


def test_cc_time_shift_adjoint_source_predicts_shift_rate():
    """Advancing u by eps (analytic shift) changes T by +eps to first order; the adjoint source
    (un-reversed) contracted with that change and dt must give T * eps."""
    u0 = _wavelet()
    u = _wavelet(2.0)  # delayed ~2 s on the causal side
    win = M.get_window(T_AXIS, 100.0, 200.0)
    T = np.sqrt(2 * M.cc_time_shift(u, u0, win, T_AXIS)[0])
    _, a = M.cc_time_shift(u, u0, win, T_AXIS)
    eps = 1e-4
    du = _wavelet(2.0 - eps) - u
    pred = np.sum(a[::-1] * du)  # dt = 1
    sign = np.sign(M.xcorr(win * u0, win * u).argmax() - (NT - 1))
    assert pred == pytest.approx(sign * T * eps, rel=2e-2)


def test_hann_is_symmetric_textbook_window():
    w = M.get_window(T_AXIS, 100.0, 110.0, "hann")
    idx = np.flatnonzero((T_AXIS > 100) & (T_AXIS < 110))
    L = len(idx)
    np.testing.assert_allclose(w[idx], 0.5 - 0.5 * np.cos(2 * np.pi * np.arange(L) / (L - 1)), atol=1e-15)
    assert w[idx[0]] == 0.0 and w[idx[-1]] == pytest.approx(0.0, abs=1e-15)
    np.testing.assert_allclose(w[idx], w[idx][::-1], atol=1e-15)
    assert np.all(w[:idx[0]] == 0) and np.all(w[idx[-1] + 1:] == 0)


def test_cos_taper_window_values():
    w = M.get_window(T_AXIS, 100.0, 200.0)
    assert np.all(w[(T_AXIS > 100) & (T_AXIS < 200)] == 1.0)
    assert w[T_AXIS == 100][0] == 1.0 and w[T_AXIS == 200][0] == 1.0  # cos(0) on both edges
    assert w[T_AXIS == 105 - 10][0] == pytest.approx(0.5 + 0.5 * np.cos(np.pi * 0.5))
    assert np.all(w[T_AXIS >= 210] == 0) and np.all(w[T_AXIS <= 90] == 0)


def test_trapz_unit_spacing():
    y = np.array([1.0, 4.0, 2.0, 7.0])
    assert M.trapz(y) == pytest.approx(0.5 * (1 + 4) + 0.5 * (4 + 2) + 0.5 * (2 + 7))


@pytest.mark.parametrize("measurement", ["waveform_difference", "amplitude_difference", "log_amplitude_ratio"])
@pytest.mark.parametrize("veldis", ["dis", "vel"])
def test_adjoint_source_is_misfit_gradient(measurement, veldis):
    """Central differences of the misfit along du against the time-reversed adjoint source.
    'vel' returns minus the gradient: the reference differentiates the reversed source with
    -diff(adstf)/dt (make_adjoint_sources.m:166-168), which the oracle keeps."""
    u, u0, du = _traces()
    h = 1e-6
    mp, _ = M.make_adjoint_sources(u + h * du, u0, None, T_AXIS, veldis, measurement)
    mm, _ = M.make_adjoint_sources(u - h * du, u0, None, T_AXIS, veldis, measurement)
    _, a = M.make_adjoint_sources(u, u0, None, T_AXIS, veldis, measurement)
    fd = (mp[0] - mm[0]) / (2 * h)
    pred = np.sum(a[0][::-1] * du)
    assert pred == pytest.approx(fd if veldis == "dis" else -fd, rel=1e-6)


@pytest.mark.parametrize("measurement", ["waveform_difference", "amplitude_difference", "log_amplitude_ratio"])
def test_second_order_is_directional_derivative(measurement):
    """'2nd' adjoint sources are the derivative of the '1st' ones along du."""
    u, u0, du = _traces(3)
    h = 1e-5
    _, ap = M.make_adjoint_sources(u + h * du, u0, None, T_AXIS, "dis", measurement)
    _, am = M.make_adjoint_sources(u - h * du, u0, None, T_AXIS, "dis", measurement)
    _, a2 = M.make_adjoint_sources(u, u0, du, T_AXIS, "dis", measurement, "2nd")
    fd = (ap - am) / (2 * h)
    assert np.linalg.norm(a2 - fd) <= 1e-6 * np.linalg.norm(fd)


def test_data_independent_branches():
    u, _, _ = _traces()
    z = np.zeros(NT)
    m, a = M.make_adjoint_sources(u, z, None, T_AXIS, "dis", "amplitude_difference")
    assert m[0] == 2.0  # 1 per branch
    c = 2 * u / np.sum(u ** 2)
    np.testing.assert_allclose(a[0], 2 * c, rtol=1e-14)  # flipped twice per branch: unreversed
    # log_amp_ratio.m:40-65: with u_0 == 0 the data energy is 0, A0 = log(0) and the misfit is
    # infinite, while the adjoint source drops the (A - A0) factor and stays finite
    with np.errstate(divide="ignore"):
        m, a = M.make_adjoint_sources(u, z, None, T_AXIS, "dis", "log_amplitude_ratio")
    assert m[0] == np.inf and np.all(np.isfinite(a)) and np.any(a != 0)


def test_window_clipping_and_errors():
    t = np.arange(301) - 150.0  # t_end = 150 < 200: right edge clipped to t_end
    u = np.cos(0.1 * t)
    m, a = M.make_adjoint_sources(u, 0.5 * u, None, t, "dis", "amplitude_difference")
    assert np.isfinite(m[0]) and np.all(np.isfinite(a))
    with pytest.raises(ValueError, match="too short"):
        M.make_adjoint_sources(u, u, None, np.arange(50) - 25.0, "dis", "cc_time_shift")
    with pytest.raises(ValueError, match="Hessian"):
        M.make_adjoint_sources(u, u, u, t, "dis", "cc_time_shift", "2nd")

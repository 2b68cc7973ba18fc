"""fd2d_make_adjoint_sources_noalloc (no device allocation, static scratch) against the entry point
it varies, fd2d_make_adjoint_sources_queued, through the C ABI: the same kernels, so the outputs must
be identical bit for bit, on consecutive calls (the second with a smaller input, over scratch the
first left behind), for every measurement, 'dis' and 'vel', '1st' and '2nd' order; no device memory
is taken; sizes beyond the static scratch are rejected."""
import ctypes

import numpy as np
import pytest

from fd2d_noise_amd import _lib, abi

pytestmark = pytest.mark.gpu

DT = 0.5


@pytest.fixture(scope="module")
def lib():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = _lib.load()
    for name in ("fd2d_make_adjoint_sources_queued", "fd2d_make_adjoint_sources_noalloc"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_void_p] * 5 + [ctypes.c_int64, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_int32, ctypes.c_double, ctypes.c_double, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p]
        f.restype = ctypes.c_int
    return L


def _traces(nr, nt, seed):
    """Correlation-like traces on the symmetric time axis -(nt-1)/2 .. (nt-1)/2 * DT: wave packets
    at +-(120..180) s and a time-shifted observed copy."""
    rng = np.random.default_rng(seed)
    t = (np.arange(nt) - (nt - 1) // 2) * DT
    u = np.zeros((nr, nt))
    u0 = np.zeros((nr, nt))
    for r in range(nr):
        tc = rng.uniform(120, 180)
        sh = rng.uniform(-2, 2)
        for sgn, amp in ((1, 1.0), (-1, 0.7)):
            u[r] += amp * np.exp(-((t - sgn * tc) / 8) ** 2) * np.cos(0.3 * (t - sgn * tc))
            u0[r] += amp * 1.1 * np.exp(-((t - sgn * (tc + sh)) / 8) ** 2) * np.cos(0.3 * (t - sgn * (tc + sh)))
    du = 0.01 * rng.standard_normal((nr, nt))
    return t, u, u0, du


def _call(lib, name, torch, t, u, u0, du, veldis, meas, order):
    nr, nt = u.shape
    dev = lambda a: torch.as_tensor(np.ascontiguousarray(a.T), device="cuda")  # (nt, nr): MATLAB nr x nt
    ud, u0d, dud, td = dev(u), dev(u0), dev(du), torch.as_tensor(t, device="cuda")
    misfit = torch.full((nr,), np.nan, dtype=torch.float64, device="cuda")
    adstf = torch.full((nt, nr), np.nan, dtype=torch.float64, device="cuda")
    th = np.ascontiguousarray(t)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    st = getattr(lib, name)(ud.data_ptr(), u0d.data_ptr(), dud.data_ptr() if order == 2 else None, td.data_ptr(),
                            th.ctypes.data, nr, nt, veldis, meas, order, 100.0, 200.0, misfit.data_ptr(),
                            adstf.data_ptr(), stream)
    if st != abi.FD2D_OK:
        raise abi.Fd2dError(st, _lib.last_error())
    torch.cuda.synchronize()
    return misfit, adstf


CASES = [(v, m, o) for v in (0, 1) for m in range(4) for o in (1, 2) if not (m == 1 and o == 2)]


def test_noalloc_matches_queued_on_consecutive_calls(lib):
    import torch
    sizes = [(12, 2401, 5), (5, 901, 6)]  # the second call is smaller than the first
    for veldis, meas, order in CASES:
        for nr, nt, seed in sizes:
            t, u, u0, du = _traces(nr, nt, seed)
            mq, aq = _call(lib, "fd2d_make_adjoint_sources_queued", torch, t, u, u0, du, veldis, meas, order)
            mn, an = _call(lib, "fd2d_make_adjoint_sources_noalloc", torch, t, u, u0, du, veldis, meas, order)
            assert torch.isfinite(mq).all() and torch.isfinite(aq).all()
            same = torch.equal(mq, mn) and torch.equal(aq, an)
            print(f"veldis={veldis} measurement={meas} order={order} nr={nr} nt={nt}: "
                  f"{'identical' if same else 'DIFFERENT'}", flush=True)
            assert same


def test_noalloc_takes_no_device_memory(lib):
    import torch
    t, u, u0, du = _traces(8, 1601, 7)
    _call(lib, "fd2d_make_adjoint_sources_noalloc", torch, t, u, u0, du, 0, 1, 1)  # first call on the device
    ud = torch.as_tensor(np.ascontiguousarray(u.T), device="cuda")
    u0d = torch.as_tensor(np.ascontiguousarray(u0.T), device="cuda")
    td = torch.as_tensor(t, device="cuda")
    misfit = torch.zeros(8, dtype=torch.float64, device="cuda")
    adstf = torch.zeros((1601, 8), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for meas in (0, 1, 2, 3):
        st = lib.fd2d_make_adjoint_sources_noalloc(ud.data_ptr(), u0d.data_ptr(), None, td.data_ptr(),
                                                   np.ascontiguousarray(t).ctypes.data, 8, 1601, 0, meas, 1, 100.0,
                                                   200.0, misfit.data_ptr(), adstf.data_ptr(), stream)
        assert st == abi.FD2D_OK
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_rejects_sizes_beyond_its_scratch(lib):
    import torch
    nr, nt = 300, 2401  # 3 * nr * nt = 2.16e6 doubles > 2^21
    t, u, u0, du = _traces(2, nt, 8)
    u = np.repeat(u, 150, axis=0)
    u0 = np.repeat(u0, 150, axis=0)
    with pytest.raises(abi.Fd2dError) as e:
        _call(lib, "fd2d_make_adjoint_sources_noalloc", torch, t, u, u0, du, 0, 0, 1)
    assert e.value.code == abi.FD2D_EINVAL and "nr * nt <= 699050" in str(e.value)
    # the queued entry point allocates and takes the same call
    _call(lib, "fd2d_make_adjoint_sources_queued", torch, t, u, u0, du, 0, 0, 1)

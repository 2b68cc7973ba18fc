"""The C ABI library loads, exports every entry point include/fd2d_noise.h declares, and its
pure-host logic (output sizes, the config struct layout) matches the reference's rules."""
import ctypes
import os
import re

import numpy as np

from fd2d_noise_amd import Config, _lib, abi
from fd2d_noise_amd.config import CConfig

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(fd2d_\w+)\s*\(", text, flags=re.M))
    return names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = declared_symbols()
    assert {"fd2d_run_forward1_green", "fd2d_run_forward2_correlation", "fd2d_run_noise_adjoint"} <= names
    for n in names:
        assert hasattr(lib, n), n
    assert set(_lib.EXPORTS) == names
    assert "gfx950" in _lib.version()


def test_config_struct_layout():
    # offsets of the C struct on x86-64 (natural alignment)
    assert CConfig.Lx.offset == 0 and CConfig.nx.offset == 16 and CConfig.dt.offset == 24
    assert CConfig.nt.offset == 32 and CConfig.n_sample.offset == 48 and CConfig.w_sample.offset == 56
    assert CConfig.dw.offset == 64 and CConfig.absorb_width.offset == 72 and CConfig.absorb_bottom.offset == 92
    assert ctypes.sizeof(CConfig) == 96


def query(c, solver, mode, nr):
    dims = (ctypes.c_int64 * 3 * 4)()
    cc = c.to_c()
    n = _lib.load().fd2d_query_output_dims(ctypes.byref(cc), solver, float(mode), nr,
                                            ctypes.cast(dims, ctypes.POINTER(ctypes.c_int64 * 3)))
    return n, [tuple(d) for d in dims]


def test_query_output_dims_follow_reference_rules():
    c = Config(nt=150, fw_nth=4)  # tiny setup with fw_nth = 4
    n, d = query(c, abi.FD2D_GREEN, 1, 3)
    # run_forward1_green.m:98-103: count of n in nt..2nt-1 with mod(n, fw_nth) == 0
    n_fw = sum(1 for q in range(150, 300) if q % 4 == 0)
    assert n == 3 and d[0] == (50, 50, 121) and d[1] == (50, 50, n_fw) and d[2] == (3, 150, 1)
    n, d = query(c, abi.FD2D_GREEN, 0, 3)
    assert d[1] == (0, 0, 0)
    n, d = query(c, abi.FD2D_CORRELATION, 1, 2)  # floor((2nt-1)/fw_nth), run_forward2_correlation.m:94
    assert n == 2 and d[0] == (2, 299, 1) and d[1] == (50, 50, 299 // 4)
    n, d = query(c.replace(n_basis_fct=4), abi.FD2D_ADJOINT, 13, 2)
    assert n == 4 and d[0] == (50, 50, 5) and d[1] == (50, 50, 121) and d[2] == (50, 50, 299 // 4)
    n, d = query(c, abi.FD2D_ADJOINT, 0, 2)
    assert d[0] == (50, 50, 2) and d[1] == (0, 0, 0) and d[2] == (0, 0, 0)
    n, _ = query(c.replace(order=5), abi.FD2D_GREEN, 0, 1)
    assert n == -abi.FD2D_EUNSUPPORTED
    assert "order 5" in _lib.last_error()


def test_matlab_dims():
    assert abi.matlab_dims((5,)) == (5, 1, 1)
    assert abi.matlab_dims((5, 2)) == (5, 2, 1)
    assert abi.matlab_dims(np.zeros((2, 3, 4)).shape) == (2, 3, 4)


def _solve_struct(c, solver, mode=0.0, **arrays):
    """An fd2d_solve whose descriptors carry sizes only (the byte estimate reads no data)."""
    from fd2d_noise_amd.batch import CSolve
    keep = [c.to_c()]
    s = CSolve()
    s.solver = solver
    s.mode = mode
    s.cfg = ctypes.cast(ctypes.pointer(keep[0]), ctypes.c_void_p)
    for name, dims in arrays.items():
        d = abi.CArray()
        d.data = 1  # non-NULL marks the array as given; never dereferenced
        d.dims[:] = dims
        keep.append(d)
        setattr(s, name, ctypes.pointer(d))
    return s, keep


def _bytes(s):
    from fd2d_noise_amd import batch
    b = ctypes.c_int64(0)
    st = batch._bind().fd2d_solve_device_bytes(ctypes.byref(s), ctypes.byref(b))
    return st, b.value


def test_solve_device_bytes():
    small, big = Config(nx=100, nz=100), Config(nx=2400, nz=2400)
    for c in (small, big):
        P = c.nx * c.nz
        s, _ = _solve_struct(c, abi.FD2D_GREEN, src=(1, 2, 1), rec=(3, 2, 1))
        st, b = _bytes(s)
        # two ping-pong sets of v, sxy, szy and the two coefficients: 8 floats per node of the padded
        # layout (64 x 32 tiles, one tile column and 32 rows of ghosts)
        E = 64 * (-(-c.nx // 64) + 1) * (32 * -(-c.nz // 32) + 32)
        assert st == abi.FD2D_OK and 8 * 4 * E <= b <= 8 * 4 * E + 64 * 1024
        s, _ = _solve_struct(c, abi.FD2D_CORRELATION, rec=(3, 2, 1), G_fft=(c.nx, c.nz, c.n_sample),
                             spectrum=(c.n_sample, 1, 1), source_dist=(c.nx, c.nz, 1))
        st, bc = _bytes(s)
        # on top of the Green run's buffers: the FP32 spectrum rows (2K padded to 256) and the slab
        # of synthesised slices (up to 512), 4 bytes per node each
        rows = 256 * -(-2 * c.n_sample // 256)
        assert st == abi.FD2D_OK and bc > b + 4 * P * (rows + min(512, c.n_corr // c.freq_samp))
    s, _ = _solve_struct(small, 7)
    assert _bytes(s)[0] == abi.FD2D_EINVAL and "solver" in _lib.last_error()


def test_resident_arguments_are_checked_before_the_device():
    """fd2d_resident_*: element sizes other than 4/8/16, negative dims, NULL handles and host arrays
    of the wrong size are FD2D_EINVAL; freeing NULL is a no-op (no device needed for any of it)."""
    lib = _lib.load()
    h = ctypes.c_void_p()
    dims = (ctypes.c_int64 * 3)(2, 3, 1)
    assert lib.fd2d_resident_alloc(ctypes.byref(dims), 3, ctypes.byref(h)) == abi.FD2D_EINVAL
    assert "elem_bytes" in _lib.last_error()
    neg = (ctypes.c_int64 * 3)(2, -3, 1)
    assert lib.fd2d_resident_alloc(ctypes.byref(neg), 8, ctypes.byref(h)) == abi.FD2D_EINVAL
    d = abi.CArray()
    d.data = None
    d.dims[:] = (4, 1, 1)
    assert lib.fd2d_resident_upload(ctypes.byref(d), 8, ctypes.byref(h)) == abi.FD2D_EINVAL
    assert "NULL" in _lib.last_error()
    assert lib.fd2d_resident_view(None, None, None) == abi.FD2D_EINVAL
    assert lib.fd2d_resident_download(None, None, None) == abi.FD2D_EINVAL
    assert lib.fd2d_resident_free(None) == abi.FD2D_OK

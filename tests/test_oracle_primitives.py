"""Pins of the oracle's discretisation primitives (code/propagation/*.m) by closed forms.

* dx_v / dz_v / div_s: a staggered first-derivative stencil of order M is exact, at the staggered
  midpoint, for polynomials of degree <= M, and not for degree M+1. The set nodes are the zero
  frames of dx_v.m:22,28 and div_s.m:85,95 (and their generalisation for orders 6 and 8).
* absbound: the Gaussian taper exp(-(d)^2/(2w)^2) of init_absbound.m:18, separable, 1 inside.
* interpol_matrix: known-answer tables worked by hand from run_forward2_correlation.m:127-165 and
  integration_limits.m:2-14.
"""
import numpy as np
import pytest

import oracle
from fd2d_noise_amd import Config

P = oracle.primitives


def _poly(coef, x):
    return sum(c * x ** d for d, c in enumerate(coef))


def _dpoly(coef, x):
    return np.zeros_like(x) + sum(d * c * x ** (d - 1) for d, c in enumerate(coef) if d > 0)


@pytest.mark.parametrize("order", [2, 4, 6, 8])
@pytest.mark.parametrize("axis", ["x", "z"])
def test_forward_derivative_polynomial_exactness(order, axis):
    rng = np.random.default_rng(order)
    nx, nz, h = 23, 19, 0.25
    r = order // 2
    for deg in range(order + 2):
        coef = rng.standard_normal(deg + 1)
        coef[-1] = 1.0
        if axis == "x":
            x = (np.arange(nx) - (nx - 1) / 2) * h
            v = np.repeat(_poly(coef, x)[:, None], nz, axis=1)
            out = P.dx_v(v, h, order)[:, 0]
            exact = _dpoly(coef, x[:-1] + h / 2)
        else:
            z = (np.arange(nz) - (nz - 1) / 2) * h
            v = np.repeat(_poly(coef, z)[None, :], nx, axis=0)
            out = P.dz_v(v, h, order)[0, :]
            exact = _dpoly(coef, z[:-1] + h / 2)
        n_out = len(out)
        lo, hi = r - 1, n_out - r  # set entries: 0-based lo..hi (order 4: MATLAB 2..n-2)
        set_idx = np.arange(lo, hi + 1)
        frame = np.setdiff1d(np.arange(n_out), set_idx)
        assert np.all(out[frame] == 0.0)
        err = np.max(np.abs(out[set_idx] - exact[set_idx])) / max(1.0, np.max(np.abs(exact)))
        if deg <= order:
            assert err < 1e-12, (deg, err)
        else:
            assert err > 1e-9, (deg, err)  # not exact beyond the order: the order is what it claims


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_div_s_polynomial_exactness(order):
    rng = np.random.default_rng(10 + order)
    nx, nz, h = 21, 19, 0.25
    r = order // 2
    for deg in range(order + 2):
        cx = rng.standard_normal(deg + 1)
        cz = rng.standard_normal(deg + 1)
        cx[-1] = cz[-1] = 1.0
        x = (np.arange(nx) - (nx - 1) / 2) * h
        z = (np.arange(nz) - (nz - 1) / 2) * h
        xs = x[:-1] + h / 2  # sxy(i) sits half a cell right of node i
        zs = z[:-1] + h / 2
        sxy = np.repeat(_poly(cx, xs)[:, None], nz, axis=1)
        szy = np.repeat(_poly(cz, zs)[None, :], nx, axis=0)
        out = P.div_s(sxy, szy, h, h, order)
        ix = np.arange(r, nx - r)  # 0-based set rows: order 4 -> MATLAB 3..nx-2
        iz = np.arange(r, nz - r)
        mx = np.zeros(nx, bool)
        mx[ix] = True
        mz = np.zeros(nz, bool)
        mz[iz] = True
        exact = (np.where(mx, _dpoly(cx, x), 0.0)[:, None] + np.where(mz, _dpoly(cz, z), 0.0)[None, :])
        err = np.max(np.abs(out - exact)) / max(1.0, np.max(np.abs(exact)))
        if deg <= order:
            assert err < 1e-12, (deg, err)
        else:
            assert err > 1e-9, (deg, err)


def test_order4_coefficients_match_reference_text():
    # dx_v.m:29 writes 9*(v(i+1)-v(i))/(8*dx) - (v(i+2)-v(i-1))/(24*dx)
    v = np.zeros((8, 3))
    v[4, :] = 1.0
    out = P.dx_v(v, 1.0, 4)[:, 0]
    # i (1-based) = 2..6: v(i+1)=1 at i=4 (9/8), v(i)=1 at i=5 (-9/8), v(i+2)=1 at i=3 (-1/24),
    # v(i-1)=1 at i=6 (+1/24)
    assert np.allclose(out, [0, 0, -1 / 24, 9 / 8, -9 / 8, 1 / 24, 0])


def test_absbound_closed_form():
    c = Config(Lx=1.0e5, Lz=0.8e5, nx=101, nz=81, absorb_width=1.0e4)
    ab = P.absbound(c)
    x, z, w = c.x(), c.z(), c.absorb_width
    # x = 0: left taper exp(-w^2/(2w)^2) = exp(-1/4); centre column z untouched
    jc = c.nz // 2
    assert abs(ab[0, jc] - np.exp(-0.25)) < 1e-14
    assert abs(ab[-1, jc] - np.exp(-0.25)) < 1e-14
    assert abs(ab[c.nx // 2, 0] - np.exp(-0.25)) < 1e-14
    # interior exactly 1
    inner = (x[:, None] > w) & (x[:, None] < c.Lx - w) & (z[None, :] > w) & (z[None, :] < c.Lz - w)
    assert np.all(ab[inner] == 1.0)
    # separable: ab(i,j) = ab(i,jc) * ab(ic,j)
    ic = c.nx // 2
    assert np.allclose(ab, ab[:, jc:jc + 1] * ab[ic:ic + 1, :], rtol=1e-15, atol=0)
    # a point 0.3w inside the left frame
    d = w - x[3]
    assert abs(ab[3, jc] - np.exp(-d * d / (2 * w) ** 2)) < 1e-14
    # switches: only the top side
    c2 = c.replace(absorb_left=0, absorb_right=0, absorb_bottom=0, absorb_top=1)
    ab2 = P.absbound(c2)
    assert np.all(ab2[:, 0] == 1.0) and abs(ab2[ic, -1] - np.exp(-0.25)) < 1e-14


def test_interpol_matrix_known_answers():
    I2 = P.interpol_matrix(10, 2)
    exp2 = np.array([[0.7, 0], [0.9, 0], [0.9, 0.1], [0.7, 0.3], [0.5, 0.5],
                     [0.3, 0.7], [0.1, 0.9], [0, 0.9], [0, 0.7], [0, 0.5]])
    assert np.allclose(I2, exp2, atol=1e-15)
    I3 = P.interpol_matrix(10, 3)
    a, b = 5 / 6, 1 / 6
    exp3 = np.array([[a, 0, 0], [a, b, 0], [0.5, 0.5, 0],
                     [b, a, 0], [0, a, b], [0, 0.5, 0.5],
                     [0, 0.25, 0.75], [0, 0, 1], [0, 0, 0.75], [0, 0, 0.5]])
    assert np.allclose(I3, exp3, atol=1e-15)
    assert np.all(P.interpol_matrix(7, 0) == 1.0)
    with pytest.raises(Exception):
        P.interpol_matrix(10, 1)  # the reference indexes a second band that does not exist

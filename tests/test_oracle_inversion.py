"""Pins of the FP64 caller oracle (oracle/inversion.py) on the CPU: closed forms of the Gaussian
blur and the parameter map, the regularisation's value and gradient, the two-loop recursion against
the dense BFGS inverse-Hessian update, the Powell-Wolfe conditions, L-BFGS convergence and exit
flags, and a finite-difference check of the composed objective and gradient."""
import numpy as np
import pytest

from fd2d_noise_amd import Config
from oracle import inversion as I

from helpers import hetero_model


def test_gaussblur_preserves_constants_and_orientation():
    x1 = np.arange(30) * 2.0
    x2 = np.arange(40) * 3.0
    Y = np.full((30, 40), 2.5)
    np.testing.assert_allclose(I.gaussblur2d(Y, x1, x2, (4.0, 7.0)), Y, rtol=1e-14)
    # a delta in the interior spreads as exp(-d^2/(2 sigma^2)) along each axis with its own sigma
    D = np.zeros((30, 40))
    D[15, 20] = 1.0
    G = I.gaussblur2d(D, x1, x2, (4.0, 7.0))
    r1 = G[15:19, 20] / G[15, 20]
    r2 = G[15, 20:23] / G[15, 20]
    np.testing.assert_allclose(r1, np.exp(-(np.arange(4) * 2.0 / 4.0) ** 2 / 2), rtol=1e-13)
    np.testing.assert_allclose(r2, np.exp(-(np.arange(3) * 3.0 / 7.0) ** 2 / 2), rtol=1e-13)
    # the truncation at sqrt(14) sigma: 14.97 along x1 (nodes up to 7 apart), 26.19 along x2 (8 apart)
    assert G[15 + 7, 20] > 0 and G[15 + 8, 20] == 0 and G[15, 28] > 0 and G[15, 29] == 0
    # the interior value is 1 / (w1 * w2) with w the truncated Gaussian sums
    w1 = np.sum(np.exp(-(np.arange(-7, 8) * 2.0 / 4.0) ** 2 / 2))
    w2 = np.sum(np.exp(-(np.arange(-8, 9) * 3.0 / 7.0) ** 2 / 2))
    assert G[15, 20] == pytest.approx(1.0 / (w1 * w2), rel=1e-13)


def test_gaussblur_tiny_sigma_is_identity():
    rng = np.random.default_rng(0)
    Y = rng.standard_normal((9, 7))
    np.testing.assert_array_equal(I.gaussblur2d(Y, np.arange(9.0), np.arange(7.0), (0.2, 0.2)), Y)


def test_map_frame_first_unit_taper_node():
    c = Config()  # width 5000 m, dx = 60 km / 49: first X > width at node 5
    assert I.map_frame(c) == (5, 5)
    c2 = Config(absorb_left=0, absorb_bottom=0)
    assert I.map_frame(c2) == (0, 0)


def test_maps_with_identity_blur():
    c = Config(nx=20, nz=16)
    rng = np.random.default_rng(1)
    m = rng.standard_normal(c.nx * c.nz * 3)
    P = I.map_m_to_parameters(m, c, (1.0, 1.0), (1.0, 1.0), ref_source=0.3, ref_structure=1.0, mu_0=2.0)
    M = m.reshape((c.nx, c.nz, 3), order="F")
    np.testing.assert_allclose(P[:, :, :2], 0.3 + M[:, :, :2], rtol=1e-15)
    np.testing.assert_allclose(P[:, :, 2], 2.0 * (1.0 + M[:, :, 2]), rtol=1e-15)
    g = rng.standard_normal((c.nx, c.nz, 3))
    gm = I.map_gradparameters_to_gradm(g, c, (1.0, 1.0), (1.0, 1.0), mu_0=2.0).reshape(g.shape, order="F")
    ix, iz = I.map_frame(c)
    inner = np.zeros_like(g, dtype=bool)
    inner[ix:c.nx - ix, iz:c.nz - iz, :] = True
    np.testing.assert_array_equal(gm[~inner], 0.0)
    np.testing.assert_allclose(gm[ix:c.nx - ix, iz:c.nz - iz, :2], g[ix:c.nx - ix, iz:c.nz - iz, :2], rtol=1e-15)
    np.testing.assert_allclose(gm[ix:c.nx - ix, iz:c.nz - iz, 2], 2.0 * g[ix:c.nx - ix, iz:c.nz - iz, 2], rtol=1e-15)


def test_map_blurs_only_inside_frame():
    c = Config(nx=24, nz=24)
    rng = np.random.default_rng(2)
    m = rng.standard_normal(c.nx * c.nz * 2)
    P = I.map_m_to_parameters(m, c, (3000.0, 3000.0), (2000.0, 4000.0), 0.0, 1.0, 1.0)
    M = m.reshape((c.nx, c.nz, 2), order="F")
    ix, iz = I.map_frame(c)
    np.testing.assert_array_equal(P[:ix, :, 0], M[:ix, :, 0])
    np.testing.assert_array_equal(P[:, :iz, 1], 1.0 + M[:, :iz, 1])
    assert np.std(P[ix:-ix, iz:-iz, 0]) < 0.7 * np.std(M[ix:-ix, iz:-iz, 0])  # smoothed


@pytest.mark.parametrize("weighting", ["scalar", "field"])
def test_regularization_value_and_gradient(weighting):
    nx, nz = 5, 4
    rng = np.random.default_rng(3)
    m = rng.standard_normal(nx * nz * 3)
    m0 = rng.standard_normal(nx * nz * 3)
    w = 2.0 if weighting == "scalar" else rng.uniform(1, 4, nx * nz)
    alpha, beta = 0.7, 1.9
    j, g = I.regularization(1.5, np.zeros_like(m), m, m0, nx, nz, alpha, beta, w)
    wv = np.full(nx * nz, w) if weighting == "scalar" else w
    d = m - m0
    expect = 1.5 + alpha / 2 * np.sum(np.tile(wv, 2) * d[:40] ** 2) + beta / 2 * np.sum(wv * d[40:] ** 2)
    assert j == pytest.approx(expect, rel=1e-14)
    h = 0.25  # central differences are exact for a quadratic
    for k in (0, 17, 33, 45, 59):
        e = np.zeros_like(m)
        e[k] = h
        jp, _ = I.regularization(0.0, None, m + e, m0, nx, nz, alpha, beta, w)
        jm, _ = I.regularization(0.0, None, m - e, m0, nx, nz, alpha, beta, w)
        assert (jp - jm) / (2 * h) == pytest.approx(g[k], rel=1e-10, abs=1e-13)


def test_two_loop_equals_dense_bfgs_inverse():
    """Nocedal & Wright (7.19)/(6.17): H_{k+1} = (I - rho s y') H_k (I - rho y s') + rho s s',
    H_0 = gak I, over the stored pairs oldest first."""
    rng = np.random.default_rng(4)
    n, lmax = 12, 5
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)
    P = np.zeros((n, lmax))
    D = np.zeros((n, lmax))
    rho = np.zeros(lmax)
    ln, l = 1, 0
    pairs = []
    for k in range(7):  # wraps the ring buffer
        p = rng.standard_normal(n)
        d = A @ p
        rho[ln - 1] = 1 / (d @ p)
        P[:, ln - 1], D[:, ln - 1] = p, d
        pairs.append((p, d))
        l = min(l + 1, lmax)
        ln = ln % lmax + 1
    gak = 0.37
    g = rng.standard_normal(n)
    H = gak * np.eye(n)
    for p, d in pairs[-l:]:
        r = 1 / (d @ p)
        V = np.eye(n) - r * np.outer(d, p)
        H = V.T @ H @ V + r * np.outer(p, p)
    np.testing.assert_allclose(I.two_loop(g, P, D, rho, gak, l, ln, lmax), H @ g, rtol=1e-11)


def _quadratic(n=20, seed=5, cond=50.0):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.geomspace(1, cond, n)) @ Q.T
    b = rng.standard_normal(n)
    return A, b


def _quad_problem(A, b, calls=None):
    def fgc(m):
        if calls is not None:
            calls.append("fg")
        return 0.5 * m @ A @ m - b @ m, A @ m - b, None
    return I.Problem(fgc)


@pytest.mark.parametrize("sig0", [1e-3, 1.0, 40.0])
@pytest.mark.parametrize("larger", [True, False])
def test_wolfe_step_satisfies_conditions(sig0, larger):
    A, b = _quadratic()
    x = np.zeros(len(b))
    f, g, _ = _quad_problem(A, b).fgc(x)
    s = g.copy()  # xn = x - sig*s
    stg = s @ g
    sig, new = I.wolfe(x, s, stg, f, 1e-3, 0.6, sig0, larger, _quad_problem(A, b))
    assert f - new["objective"] >= 1e-3 * sig * stg  # Armijo
    assert new["gradient"] @ s <= 0.6 * stg  # Wolfe
    np.testing.assert_allclose(new["m"], x - sig * s)
    np.testing.assert_allclose(new["gradient"], A @ new["m"] - b)


def test_lbfgs_quadratic_and_rosenbrock_converge():
    A, b = _quadratic()
    flag, m, hist = I.lbfgs(np.zeros(len(b)), I.Options(tolerance=1e-8, successive_change=0.0),
                            _quad_problem(A, b))
    assert flag == 0
    np.testing.assert_allclose(m, np.linalg.solve(A, b), rtol=1e-6)
    assert all(h2["objective"] <= h1["objective"] for h1, h2 in zip(hist, hist[1:]))

    def rosen(x):
        f = np.sum(100 * (x[1:] - x[:-1] ** 2) ** 2 + (1 - x[:-1]) ** 2)
        g = np.zeros_like(x)
        g[:-1] = -400 * x[:-1] * (x[1:] - x[:-1] ** 2) - 2 * (1 - x[:-1])
        g[1:] += 200 * (x[1:] - x[:-1] ** 2)
        return f, g, None
    flag, m, _ = I.lbfgs(np.array([-1.2, 1.0, -0.5, 0.8]), I.Options(tolerance=1e-9, successive_change=0.0,
                                                                     max_iterations=500), I.Problem(rosen))
    assert flag == 0
    np.testing.assert_allclose(m, 1.0, atol=1e-6)


def test_lbfgs_exit_flags():
    A, b = _quadratic()
    flag, _, hist = I.lbfgs(np.zeros(len(b)), I.Options(max_iterations=2, tolerance=1e-12, successive_change=0.0),
                            _quad_problem(A, b))
    assert flag == 1 and len(hist) == 3
    flag, _, hist = I.lbfgs(np.zeros(len(b)), I.Options(tolerance=1e-12, successive_change=1e30,
                                                        successive_iterations=3), _quad_problem(A, b))
    assert flag == 2 and len(hist) == 4


def test_lbfgs_first_step_is_gradient_with_init_length():
    """No curvature pairs yet: s = g; at x = 0 the exact line-search step b'b / b'Ab satisfies Armijo
    and Wolfe at once, so with try_to_increase off the step is taken after one evaluation."""
    A, b = _quadratic()
    sig0 = (b @ b) / (b @ A @ b)
    calls = []
    _, _, hist = I.lbfgs(np.zeros(len(b)), I.Options(max_iterations=1, init_step_length=sig0,
                                                     wolfe_try_to_increase_step_length=False, tolerance=1e-12),
                         _quad_problem(A, b, calls))
    np.testing.assert_allclose(hist[1]["m"], sig0 * b, rtol=1e-14)
    assert calls == ["fg", "fg"]


def _noise_usr(typ, c=None):
    from fd2d_noise_amd.inversion import UsrPar
    from helpers import gaussian_spectrum
    c = Config() if c is None else c
    X = c.x()[:, None]
    Z = c.z()[None, :]
    blob = np.exp(-((X - 3.2e4) ** 2 + (Z - 2.8e4) ** 2) / (2 * 6e3 ** 2))
    ref = np.array([[2.0e4, 3.0e4]])
    arr = np.array([[2.0e4, 3.0e4], [4.0e4, 3.5e4], [3.5e4, 2.0e4]])
    usr = UsrPar(config=c, ref_stat=ref, array=arr, c_data=np.zeros((2, c.n_corr)), t=c.t_corr(),
                 rho=3000.0 * np.ones((c.nx, c.nz), order="F"), spectrum=gaussian_spectrum(c), type=typ)
    m_true = np.concatenate([(1 + 0.5 * blob).ravel(order="F"), (0.05 * blob).ravel(order="F")])
    return usr, m_true, blob


@pytest.mark.parametrize("typ", ["source", "structure", "joint"])
def test_objective_gradient_finite_differences(orc, typ):
    """eval_objective_and_gradient: g'dm against central differences of j along source- and
    structure-map perturbations (the optlib_check_derivative_clean.m:6-35 test). The discrete
    adjoint is not exact (FP32 snapshots, the structure kernel's fw_nth factor), hence 2e-3."""
    usr, m_true, blob = _noise_usr(typ)
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = I.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    j, g, _ = I.eval_objective_and_gradient(m, usr, orc)
    assert j > 0
    rng = np.random.default_rng(0)
    for part in ("source", "structure"):
        dm = np.zeros(2 * P)
        sl = slice(0, P) if part == "source" else slice(P, 2 * P)
        dm[sl] = (blob * (0.5 + rng.random((c.nx, c.nz)))).ravel(order="F")
        h = 1e-3
        jp, _, _ = I.eval_objective_and_gradient(m + h * dm, usr, orc, only_objective=True)
        jm, _, _ = I.eval_objective_and_gradient(m - h * dm, usr, orc, only_objective=True)
        fd = (jp - jm) / (2 * h)
        if (typ, part) in (("source", "structure"), ("structure", "source")):
            assert g @ dm == 0.0  # that kernel is not requested by the type
        else:
            assert g @ dm == pytest.approx(fd, rel=1e-4 if part == "source" else 2e-3)


@pytest.mark.parametrize("typ,measurement", [("source", "waveform_difference"), ("structure", "waveform_difference"),
                                             ("joint", "waveform_difference"),
                                             ("structure", "amplitude_difference")])
def test_hessian_vector_product_finite_differences(orc, typ, measurement):
    """eval_hessian_vector_product_2: H dm against central differences of the gradient along dm
    (the optlib_check_hessian.m test). dm lies in parameter space (d mu in Pa); with the identity
    blur the model-space direction is (d source, d mu / mu_0). For the source problem the misfit is
    quadratic in the source maps and the product is exact; the structure terms carry the
    inexactness of the discrete adjoint."""
    usr, m_true, blob = _noise_usr(typ)
    usr.measurement_source = usr.measurement_structure = measurement
    usr.window = (2.0, 8.0)
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = I.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    rng = np.random.default_rng(1)
    dm = np.zeros(2 * P)
    if typ != "structure":
        dm[:P] = (blob * (0.5 + rng.random((c.nx, c.nz)))).ravel(order="F")
    if typ != "source":
        dm[P:] = 0.03 * (blob * (0.5 + rng.random((c.nx, c.nz)))).ravel(order="F")
    dmp = dm.copy()
    dmp[P:] *= usr.mu_0
    Hm = I.eval_hessian_vector_product(m, dmp, usr, orc)
    h = 1e-3
    _, gp, _ = I.eval_objective_and_gradient(m + h * dm, usr, orc)
    _, gm, _ = I.eval_objective_and_gradient(m - h * dm, usr, orc)
    fd = (gp - gm) / (2 * h)
    tol = {"source": (1e-9, 0), "structure": (0, 2e-3), "joint": (1e-6, 1e-2)}[typ]
    for sl, t in ((slice(0, P), tol[0]), (slice(P, 2 * P), tol[1])):
        if t == 0:
            assert np.all(Hm[sl] == 0) and np.all(fd[sl] == 0)
        else:
            assert np.linalg.norm(Hm[sl] - fd[sl]) <= t * np.linalg.norm(fd[sl])


def test_hessian_regularisation_term(orc):
    """eval_hessian_vector_product_2.m:191-196 adds alpha / beta * weighting on the interior nodes,
    independent of dm (pinned as the reference states it)."""
    usr, m_true, _ = _noise_usr("source")
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = I.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    dm = np.zeros(2 * P)
    dm[:P] = 1e-3
    H0 = I.eval_hessian_vector_product(m, dm, usr, orc)
    usr.reg_alpha, usr.reg_beta = 0.25, 0.5
    H1 = I.eval_hessian_vector_product(m, dm, usr, orc)
    inner = (orc_absbound(c).ravel(order="F") == 1.0)
    np.testing.assert_allclose((H1 - H0)[:P], 0.25 * inner, atol=1e-15)
    np.testing.assert_allclose((H1 - H0)[P:], 0.5 * inner, atol=1e-15)


def orc_absbound(c):
    from oracle import primitives
    return primitives.absbound(c)

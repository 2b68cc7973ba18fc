"""Parity of the caller-side HIP kernels and drivers (through the C ABI) with the FP64 oracle.

FP64 kernels (adjoint sources, dot, axpby, gaussblur2d, parameter maps, regularisation): relative
1e-10 (compensated sums against numpy's pairwise sums). L-BFGS on a device quadratic: the same
iterates as the oracle's L-BFGS to 1e-9. The composed objective and gradient (FP32 wavefields in
the solves): misfit 1e-4, gradient 1e-3 relative L2.
"""
import math
import types

import numpy as np
import pytest

from fd2d_noise_amd import abi

from conftest import rel_l2

pytestmark = pytest.mark.gpu

NT = 601
T_AXIS = np.arange(NT) - 300.0


@pytest.fixture(scope="module", autouse=True)
def gpu():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from fd2d_noise_amd import _lib
    _lib.load()


def _traces(nr=3, seed=0):
    rng = np.random.default_rng(seed)

    def wavelet(shift, amp):
        a = np.abs(T_AXIS) - 150.0 - shift
        return amp * np.exp(-(a / 12.0) ** 2) * np.cos(0.2 * a)
    u = np.stack([wavelet(1.3 * (k + 1), 1.0 + 0.1 * k) + 0.01 * rng.standard_normal(NT) for k in range(nr)])
    u0 = np.stack([wavelet(0.0, 1.0) for _ in range(nr)])
    du = rng.standard_normal((nr, NT)) * np.exp(-((np.abs(T_AXIS) - 150.0) / 30.0) ** 2)
    return u, u0, du


CASES = [(m, v, o) for m in ("waveform_difference", "cc_time_shift", "amplitude_difference", "log_amplitude_ratio")
         for v in ("dis", "vel") for o in ("1st", "2nd") if not (m == "cc_time_shift" and o == "2nd")]


@pytest.mark.parametrize("measurement,veldis,order", CASES)
def test_adjoint_sources_device(measurement, veldis, order):
    import torch
    from fd2d_noise_amd import device as D
    from fd2d_noise_amd.measurement import make_adjoint_sources_device
    from oracle import measurement as M
    u, u0, du = _traces()
    m_ref, a_ref = M.make_adjoint_sources(u, u0, du, T_AXIS, veldis, measurement, order)
    td = torch.as_tensor(T_AXIS, device="cuda")
    m, a = make_adjoint_sources_device(D.to_device(u), D.to_device(u0), D.to_device(du), td, veldis, measurement,
                                       order)
    m, a = m.cpu().numpy(), D.to_host(a)
    print(f"{measurement} {veldis} {order}: misfit {rel_l2(m, m_ref):.2e} adstf {rel_l2(a, a_ref):.2e}")
    np.testing.assert_allclose(m, m_ref, rtol=1e-10, atol=1e-300)
    assert rel_l2(a, a_ref) < 1e-10


def test_adjoint_sources_host_entry_and_errors():
    from fd2d_noise_amd.measurement import make_adjoint_sources
    from oracle import measurement as M
    u, u0, _ = _traces(nr=4, seed=2)
    m, a = make_adjoint_sources(u, u0, None, T_AXIS, "dis", "cc_time_shift")
    m_ref, a_ref = M.make_adjoint_sources(u, u0, None, T_AXIS, "dis", "cc_time_shift")
    np.testing.assert_allclose(m, m_ref, rtol=1e-12)
    assert rel_l2(a, a_ref) < 1e-10
    with pytest.raises(abi.Fd2dError) as e:
        make_adjoint_sources(u, u0, u, T_AXIS, "dis", "cc_time_shift", deriv_order="2nd")
    assert e.value.code == abi.FD2D_EUNSUPPORTED
    t_short = np.arange(NT) * 0.1 - 30.0  # t(end) = 30 s < 100 s
    with pytest.raises(abi.Fd2dError) as e:
        make_adjoint_sources(u, u0, None, t_short, "dis", "amplitude_difference")
    assert e.value.code == abi.FD2D_EINVAL and "too short" in str(e.value)
    # data-independent (u_0 == 0) and time-shift threshold cases
    z = np.zeros_like(u)
    m, a = make_adjoint_sources(u, z, None, T_AXIS, "dis", "cc_time_shift")
    m_ref, a_ref = M.make_adjoint_sources(u, z, None, T_AXIS, "dis", "cc_time_shift")
    np.testing.assert_allclose(m, m_ref, rtol=1e-14)
    assert rel_l2(a, a_ref) < 1e-10


def test_dot_compensated_and_axpby():
    import torch
    from fd2d_noise_amd import inversion as V
    x = torch.tensor([1e16, 1.0, -1e16, 3.0], dtype=torch.float64, device="cuda")
    y = torch.ones(4, dtype=torch.float64, device="cuda")
    assert V.dot(x, y) == 4.0  # naive summation loses the 1
    rng = np.random.default_rng(7)
    for n in (0, 1, 1000, 3_000_001):
        a = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
        b = rng.standard_normal(n)
        p = a * b  # fsum of the rounded products: within eps/2 * sum|p| of the exact dot
        exact = math.fsum(p.tolist())
        got = V.dot(torch.as_tensor(a, device="cuda"), torch.as_tensor(b, device="cuda"))
        scale = float(np.sum(np.abs(p))) or 1.0
        assert abs(got - exact) <= 2e-15 * scale
    a = torch.as_tensor(rng.standard_normal(12345), device="cuda")
    b = torch.as_tensor(rng.standard_normal(12345), device="cuda")
    out = V.axpby(0.75, a, -2.5, b)
    np.testing.assert_allclose(out.cpu().numpy(), 0.75 * a.cpu().numpy() - 2.5 * b.cpu().numpy(), rtol=1e-15,
                               atol=1e-15)
    a_h, b_h = a.cpu().numpy().copy(), b.cpu().numpy()
    V.axpby(1.0, a, 3.0, b, out=a)  # out aliases x
    np.testing.assert_allclose(a.cpu().numpy(), a_h + 3.0 * b_h, rtol=1e-15, atol=1e-15)


def test_gaussblur2d_device():
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    rng = np.random.default_rng(8)
    n1, n2 = 37, 23
    Y = rng.standard_normal((n1, n2))
    x1 = np.arange(n1) * 1.5 + 10.0
    x2 = np.arange(n2) * 2.25
    ref = OI.gaussblur2d(Y, x1, x2, (2.7, 4.1))
    got = V.gaussblur2d(torch.as_tensor(np.ascontiguousarray(Y.T), device="cuda"), x1, x2, (2.7, 4.1))
    assert rel_l2(got.cpu().numpy().T, ref) < 1e-13


@pytest.mark.parametrize("n_maps,sig_src,sig_str", [(2, (3000.0, 3000.0), (2500.0, 4000.0)), (3, (1.0, 1.0), (5000.0,
                                                                                                              1500.0))])
def test_parameter_maps_device(n_maps, sig_src, sig_str):
    import torch
    from fd2d_noise_amd import Config
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    c = Config(nx=40, nz=34, absorb_right=0)
    usr = types.SimpleNamespace(config=c, sigma_source=sig_src, sigma_structure=sig_str, ref_source=0.2,
                                ref_structure=1.0, mu_0=4.8e10)
    rng = np.random.default_rng(9)
    m = rng.standard_normal(c.nx * c.nz * n_maps)
    ref = OI.map_m_to_parameters(m, c, sig_src, sig_str, 0.2, 1.0, 4.8e10)
    got = V.map_m_to_parameters(torch.as_tensor(m, device="cuda"), usr)
    got = got.cpu().numpy().reshape(-1).reshape((c.nx, c.nz, n_maps), order="F")
    assert rel_l2(got[:, :, :-1], ref[:, :, :-1]) < 1e-12
    assert rel_l2(got[:, :, -1], ref[:, :, -1]) < 1e-12
    g = rng.standard_normal((c.nx, c.nz, n_maps))
    gref = OI.map_gradparameters_to_gradm(g, c, sig_src, sig_str, 4.8e10)
    gd = torch.as_tensor(np.ascontiguousarray(g.transpose(2, 1, 0)), device="cuda")
    gm = V.map_gradparameters_to_gradm(gd, usr).cpu().numpy()
    assert rel_l2(gm, gref) < 1e-12


@pytest.mark.parametrize("weighting", ["scalar", "field"])
def test_regularization_device(weighting):
    import torch
    from fd2d_noise_amd import Config
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    c = Config(nx=30, nz=20)
    rng = np.random.default_rng(10)
    n = c.nx * c.nz * 3
    m, m0 = rng.standard_normal(n), rng.standard_normal(n)
    w = 2.0 if weighting == "scalar" else rng.uniform(1, 4, c.nx * c.nz)
    g = rng.standard_normal(n)
    j_ref, g_ref = OI.regularization(0.5, g, m, m0, c.nx, c.nz, 0.3, 1.7, w)
    usr = types.SimpleNamespace(config=c, reg_alpha=0.3, reg_beta=1.7, reg_weighting=w,
                                m0=torch.as_tensor(m0, device="cuda"))
    gd = torch.as_tensor(g, device="cuda")
    j, gd = V.regularization(0.5, gd, torch.as_tensor(m, device="cuda"), usr)
    assert j == pytest.approx(j_ref, rel=1e-13)
    assert rel_l2(gd.cpu().numpy(), g_ref) < 1e-14
    j2, _ = V.regularization(0.5, 0, torch.as_tensor(m, device="cuda"), usr)
    assert j2 == pytest.approx(j_ref, rel=1e-13)


def test_lbfgs_device_matches_oracle_on_quadratic():
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    rng = np.random.default_rng(11)
    n = 64
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    A = Q @ np.diag(np.geomspace(1, 20, n)) @ Q.T
    b = rng.standard_normal(n)
    Ad = torch.as_tensor(A, device="cuda")
    bd = torch.as_tensor(b, device="cuda")

    class Quad:
        def eval_objective_and_gradient(self, m, name=""):
            g = Ad @ m - bd
            return float(0.5 * (m @ (Ad @ m)) - bd @ m), g.contiguous(), None

        def eval_objective(self, m, name=""):
            j, _, c = self.eval_objective_and_gradient(m)
            return j, c

        def eval_grad_objective(self, m, name=""):
            _, g, c = self.eval_objective_and_gradient(m)
            return g, c

    opts = dict(tolerance=1e-6, successive_change=0.0, max_iterations=200)  # converges in ~35 iterations
    flag_ref, m_ref, hist_ref = OI.lbfgs(np.zeros(n), OI.Options(**opts),
                                         OI.Problem(lambda m: (0.5 * m @ A @ m - b @ m, A @ m - b, None)))
    usr = types.SimpleNamespace(model=[])
    flag, m, usr = V.optlib_lbfgs(torch.zeros(n, dtype=torch.float64, device="cuda"), V.Options(**opts), usr, Quad())
    print(f"lbfgs: flag {flag} its {len(usr.model) - 1} (oracle {flag_ref}, {len(hist_ref) - 1})")
    assert flag == flag_ref == 0
    assert len(usr.model) == len(hist_ref)
    for a, r in zip(usr.model, hist_ref):
        assert a["objective"] == pytest.approx(r["objective"], rel=1e-9)
    assert rel_l2(m.cpu().numpy(), m_ref) < 1e-9


@pytest.mark.parametrize("typ", ["source", "structure", "joint"])
def test_noise_objective_and_gradient_device(orc, typ):
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    from test_oracle_inversion import _noise_usr
    usr, m_true, _ = _noise_usr(typ)
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    usr.reg_alpha, usr.reg_beta = 1e-6, 2e-6
    usr.m0 = m_true
    j_ref, g_ref, c_ref = OI.eval_objective_and_gradient(m, usr, orc)
    prob = V.NoiseProblem(usr)
    j, g, c_all = prob.eval_objective_and_gradient(torch.as_tensor(m, device="cuda"))
    print(f"{typ}: j {abs(j - j_ref) / j_ref:.2e} g {rel_l2(g.cpu().numpy(), g_ref):.2e}")
    assert j == pytest.approx(j_ref, rel=1e-4)
    assert rel_l2(g.cpu().numpy(), g_ref) < 1e-3
    assert rel_l2(c_all.cpu().numpy().T, c_ref) < 1e-5
    jo, _ = prob.eval_objective(torch.as_tensor(m, device="cuda"))
    assert jo == pytest.approx(j, rel=1e-12)


@pytest.mark.parametrize("typ", ["source", "structure", "joint"])
def test_noise_objective_batched_equals_sequential(orc, typ):
    """All stations through one batch engine (NoiseProblem(batched=True)) give the sequential
    station loop's objective, gradient and traces bit for bit."""
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    from test_oracle_inversion import _noise_usr
    usr, m_true, _ = _noise_usr(typ)
    # three reference stations, each recorded at the other three points of the array
    usr.array = np.array([[2.0e4, 3.0e4], [4.0e4, 3.5e4], [3.5e4, 2.0e4], [2.6e4, 4.0e4]])
    usr.ref_stat = usr.array[:3].copy()
    c = usr.config
    P = c.nx * c.nz
    usr.c_data = np.zeros((9, c.n_corr))
    _, _, cd = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    usr.c_data = cd
    assert usr.c_data.shape[0] == 9
    m = torch.as_tensor(np.concatenate([np.ones(P), np.zeros(P)]), device="cuda")
    j1, g1, c1 = V.NoiseProblem(usr).eval_objective_and_gradient(m)
    j2, g2, c2 = V.NoiseProblem(usr, batched=True).eval_objective_and_gradient(m)
    print(f"{typ}: batched j {j2!r} sequential {j1!r}")
    assert j1 == j2 and torch.equal(g1, g2) and torch.equal(c1, c2)
    jo, _ = V.NoiseProblem(usr, batched=True).eval_objective(m)
    assert jo == V.NoiseProblem(usr).eval_objective(m)[0]


def test_noise_lbfgs_decreases_misfit(orc):
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    from test_oracle_inversion import _noise_usr
    usr, m_true, _ = _noise_usr("structure")
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m0 = torch.as_tensor(np.concatenate([np.ones(P), np.zeros(P)]), device="cuda")
    flag, m, usr = V.optlib_lbfgs(m0, V.Options(max_iterations=3, init_step_length=1.0, successive_change=0.0),
                                  usr, rng=np.random.default_rng(0))
    objs = [mod["objective"] for mod in usr.model]
    print("noise lbfgs objectives", objs)
    assert flag == 1 and len(objs) == 4
    assert all(b < a for a, b in zip(objs, objs[1:]))


@pytest.mark.parametrize("K,nb,n_src", [(121, 4, 1), (10, 3, 2), (91, 91, 1)])
def test_general_source_device(K, nb, n_src):
    from fd2d_noise_amd import inputs
    from oracle import inversion as OI
    rng = np.random.default_rng(K + nb)
    spec = rng.random((K, n_src))
    dist = rng.random((33, 17, n_src))
    got = inputs.general_source(spec, dist, nb)
    ref = OI.general_source(spec, dist, nb)
    assert got.shape == (33, 17, nb)
    assert rel_l2(got, ref) < 1e-14


def test_make_noise_source_basis_functions():
    from fd2d_noise_amd import Config, inputs
    from oracle import inversion as OI
    c = Config(Lx=2.0e5, Lz=2.0e5, nx=41, nz=41, n_basis_fct=4)
    dist, spec = inputs.make_noise_source(c, "gaussian")
    d0, s0 = inputs.make_noise_source(c, "gaussian", 0)
    assert np.all(spec == 1.0) and spec.shape == (c.n_sample, 1)
    assert rel_l2(dist, OI.general_source(s0, d0, 4)) < 1e-14


@pytest.mark.parametrize("typ", ["source", "structure", "joint"])
def test_noise_hessian_vector_product_device(orc, typ):
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    from test_oracle_inversion import _noise_usr
    usr, m_true, blob = _noise_usr(typ)
    usr.measurement_source = "waveform_difference"
    usr.measurement_structure = "amplitude_difference"
    usr.window = (2.0, 8.0)
    usr.reg_alpha, usr.reg_beta = 1e-7, 2e-7
    c = usr.config
    P = c.nx * c.nz
    _, _, usr.c_data = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = np.concatenate([np.ones(P), np.zeros(P)])
    rng = np.random.default_rng(2)
    dmp = np.zeros(2 * P)
    if typ != "structure":
        dmp[:P] = (blob * (0.5 + rng.random((c.nx, c.nz)))).ravel(order="F")
    if typ != "source":
        dmp[P:] = 0.03 * usr.mu_0 * (blob * (0.5 + rng.random((c.nx, c.nz)))).ravel(order="F")
    H_ref = OI.eval_hessian_vector_product(m, dmp, usr, orc)
    prob = V.NoiseProblem(usr)
    H = prob.eval_hessian_vector_product(torch.as_tensor(m, device="cuda"), dmp).cpu().numpy()
    for part, sl in (("source", slice(0, P)), ("structure", slice(P, 2 * P))):
        err = rel_l2(H[sl], H_ref[sl])
        print(f"hvp {typ} {part}: {err:.2e}")
        assert err < 1e-4


@pytest.mark.parametrize("typ", ["source", "structure", "joint"])
def test_noise_hessian_vector_products_batched_equal_sequential(orc, typ):
    """Several perturbations at once (NoiseProblem.eval_hessian_vector_products: every station's
    model Green and correlation runs shared by the perturbations, all solves through one batch
    engine) give eval_hessian_vector_product of each perturbation bit for bit, also when the
    perturbations of a station are split into groups; the first perturbation is a Rademacher
    (+-1) probe scaled to the parameters."""
    import torch
    from fd2d_noise_amd import inversion as V
    from oracle import inversion as OI
    from test_oracle_inversion import _noise_usr
    usr, m_true, blob = _noise_usr(typ)
    usr.array = np.array([[2.0e4, 3.0e4], [4.0e4, 3.5e4], [3.5e4, 2.0e4], [2.6e4, 4.0e4]])
    usr.ref_stat = usr.array[:2].copy()
    usr.measurement_source = "waveform_difference"
    usr.reg_alpha, usr.reg_beta = 1e-7, 2e-7
    c = usr.config
    P = c.nx * c.nz
    usr.c_data = np.zeros((6, c.n_corr))
    _, _, usr.c_data = OI.eval_objective_and_gradient(m_true, usr, orc, only_objective=True)
    m = torch.as_tensor(np.concatenate([np.ones(P), np.zeros(P)]), device="cuda")
    rng = np.random.default_rng(7)
    scale = np.concatenate([np.full(P, 0.1), np.full(P, 0.01 * usr.mu_0)])
    dms = [rng.choice([-1.0, 1.0], 2 * P) * scale]
    b = blob.ravel(order="F")
    dms.append(np.concatenate([0.5 * b, np.zeros(P)]))                # source maps only
    dms.append(np.concatenate([np.zeros(P), 0.02 * usr.mu_0 * b]))    # structure map only
    prob = V.NoiseProblem(usr)
    seq = [prob.eval_hessian_vector_product(m, dm) for dm in dms]
    got = prob.eval_hessian_vector_products(m, dms)
    print(f"{typ}: plan {prob.last_hvp_plan}")
    split = prob.eval_hessian_vector_products(m, dms, memory_fraction=1e-9)
    assert prob.last_hvp_plan["perturbations_per_group"] == 1
    for q, (a, b_, s) in enumerate(zip(got, split, seq)):
        print(f"{typ} probe {q}: |H| {float(s.norm()):.3e}")
        assert torch.equal(a, s), f"probe {q}"
        assert torch.equal(b_, s), f"probe {q} split"
    assert prob.eval_hessian_vector_products(m, []) == []

"""Batched solves against the single-solve entry points, through the C ABI.

Each solve of a batch must give exactly (bit for bit) what its own *_device call gives: the engine
runs the same tile kernel body over the same data, only with the tiles of several solves in one
launch. The mix covers the three solvers, two grid sizes under one FD order (heterogeneous nx/nz
and nt in one grouped launch), a second FD order, Born solves, point and distributed adjoint
sources, snapshots, DFT outputs, seismograms and both kernels (the structure-kernel product fused
into the grouped step). The single-solve path itself is held to the oracle by test_gpu_parity.py.
"""
import numpy as np
import pytest

from fd2d_noise_amd import Config, _lib, abi

from helpers import blob_distribution, hetero_model, interior_points

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    _lib.load()
    return torch


def _cases(torch):
    """Inputs of a mix of solves, prepared with the single-solve path. Returns a list of
    (kind, args, kwargs-without-outputs, output-dims-and-dtypes)."""
    from fd2d_noise_amd import device as D
    f32, f64, cd = torch.float32, torch.float64, torch.complex128
    c1 = Config(order=4, nx=70, nz=52, nt=160, fw_nth=2, freq_samp=3)
    c2 = Config(order=4, nx=50, nz=90, nt=120, fw_nth=3)
    c3 = Config(order=8, nx=64, nz=60, nt=140, fw_nth=4, Lx=63e3, Lz=59e3)
    # 25 x 35 = 875 tiles: large enough to keep its own launches inside the engine
    c4 = Config(order=4, nx=1600, nz=1100, nt=30, fw_nth=3, Lx=1.6e6, Lz=1.1e6, dt=0.1)
    rng = np.random.default_rng(901)
    cases = []
    for c in (c1, c2, c3, c4):
        mu, rho = hetero_model(rng, c)
        mu_d, rho_d = D.to_device(mu), D.to_device(rho)
        srcs = [interior_points(rng, c, 1, 0.4) for _ in range(2)]
        rec = interior_points(rng, c, 3)
        rec_d = D.to_device(rec)
        nr = rec.shape[0]
        spectrum = D.to_device(np.exp(-((np.asarray(c.f_sample) - 0.1) ** 2) / 0.05 ** 2)[:, None])
        sd = D.to_device(blob_distribution(rng, c))
        for q, src in enumerate(srcs):
            src_d = D.to_device(src)
            G = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
            Gout = D.empty_like_dims((c.nx, c.nz, c.n_fw_green()), f32)
            D.run_forward1_green(mu_d, rho_d, src_d, rec_d, 1, None, None, config=c, G_fft=G, G_out=Gout)
            gouts = [((c.nx, c.nz, c.n_sample), cd), ((c.nx, c.nz, c.n_fw_green()), f32), ((nr, c.nt), f64)]
            cases.append(("green", c, (mu_d, rho_d, src_d, rec_d, 1, None, None), gouts))
            corr = D.empty_like_dims((nr, c.n_corr), f64)
            C = D.empty_like_dims((c.nx, c.nz, c.n_fw_corr()), f32)
            D.run_forward2_correlation(mu_d, rho_d, G, spectrum, sd, rec_d, 1, None, None, config=c,
                                       seismograms=corr, C_out=C)
            couts = [((nr, c.n_corr), f64), ((c.nx, c.nz, c.n_fw_corr()), f32)]
            cases.append(("correlation", c, (mu_d, rho_d, G, spectrum, sd, rec_d, 1, None, None), couts))
            a = D.to_host(corr)
            scale = np.abs(a).max()  # zero on the large grid: no wave reaches its receivers in 30 steps
            adj = a[:, ::-1] / scale if scale > 0 else rng.standard_normal(a.shape)
            stf = D.to_device(np.ascontiguousarray(adj).astype(np.complex128))
            nb1 = max(c.n_basis_fct, 1) + 1
            aouts = [((c.nx, c.nz, nb1), f64), ((c.nx, c.nz, c.n_sample), cd), ((c.nx, c.nz, c.n_fw_corr()), f32),
                     ((nr, c.n_corr), f64)]
            # mode 1: point sources at the receivers, source and structure kernels, stf_fft
            cases.append(("adjoint", c, (mu_d, rho_d, C, stf, rec_d, rec_d, spectrum, None, G, 1, None, None),
                          aouts[:2] + [None, aouts[3]]))
            if q == 0:
                # mode 0 with the distributed source of a mode-1 run and the Green snapshots
                K1 = D.empty_like_dims((c.nx, c.nz, nb1), f64)
                a1 = D.empty_like_dims((c.nx, c.nz, c.n_sample), cd)
                D.run_noise_adjoint(mu_d, rho_d, C, stf, rec_d, rec_d, None, None, None, 1, None, None, config=c,
                                    K=K1, stf_fft=a1)
                cases.append(("adjoint", c, (mu_d, rho_d, Gout, a1, rec_d, rec_d, spectrum, sd, None, 0, None, None),
                              [aouts[0], None, None, aouts[3]]))
                # mode 13: saved snapshots too
                cases.append(("adjoint", c, (mu_d, rho_d, C, stf, rec_d, rec_d, spectrum, None, G, 13, None, None),
                              aouts))
                if c is c1:
                    dmu = D.to_device(0.05 * mu)
                    Gin = D.empty_like_dims((c.nx, c.nz, c.nt), f32)
                    cg = c.replace(fw_nth=1)
                    D.run_forward1_green(mu_d, rho_d, src_d, rec_d, 1, None, None, config=cg,
                                         G_fft=D.empty_like_dims((c.nx, c.nz, c.n_sample), cd), G_out=Gin)
                    cases.append(("green", cg, (mu_d, rho_d, src_d, rec_d, 0, dmu, Gin), gouts[:1] + [None, gouts[2]]))
                    Cin = D.empty_like_dims((c.nx, c.nz, cg.n_fw_corr()), f32)
                    D.run_forward2_correlation(mu_d, rho_d, G, spectrum, sd, rec_d, 1, None, None, config=cg,
                                               seismograms=D.empty_like_dims((nr, c.n_corr), f64), C_out=Cin)
                    cases.append(("correlation", cg, (mu_d, rho_d, G, spectrum, sd, rec_d, 0, dmu, Cin),
                                  [couts[0], None]))
    return cases


def _outs(torch, dims):
    from fd2d_noise_amd import device as D
    return [None if d is None else D.empty_like_dims(d[0], d[1]) for d in dims]


_NAMES = {"green": ("G_fft", "G_out", "seismograms"), "correlation": ("seismograms", "C_out"),
          "adjoint": ("K", "stf_fft", "u_out", "seismograms")}


def _alone(kind, c, args, outs):
    from fd2d_noise_amd import device as D
    fn = {"green": D.run_forward1_green, "correlation": D.run_forward2_correlation,
          "adjoint": D.run_noise_adjoint}[kind]
    fn(*args, config=c, **dict(zip(_NAMES[kind], outs)))


def _spec(kind, c, args, outs):
    from fd2d_noise_amd import batch as B
    fn = {"green": B.green, "correlation": B.correlation, "adjoint": B.adjoint}[kind]
    return fn(*args, config=c, **dict(zip(_NAMES[kind], outs)))


def _same(torch, tag, ref, got):
    for q, (x, y) in enumerate(zip(ref, got)):
        if x is None:
            continue
        assert torch.isfinite(torch.view_as_real(x) if x.is_complex() else x).all(), f"{tag}[{q}] not finite"
        same = torch.equal(x, y)
        print(f"{tag}[{q}] {tuple(x.shape)}: {'identical' if same else 'DIFFERENT'}", flush=True)
        assert same, f"{tag}[{q}]: batched result differs from the single solve"


@pytest.fixture(scope="module")
def mix(torch_mod):
    torch = torch_mod
    cases = _cases(torch)
    ref = []
    for kind, c, args, dims in cases:
        o = _outs(torch, dims)
        _alone(kind, c, args, o)
        ref.append(o)
    torch.cuda.synchronize()
    return cases, ref


def test_run_batch_matches_single_solves(torch_mod, mix):
    torch = torch_mod
    from fd2d_noise_amd import batch as B
    cases, ref = mix
    got = [_outs(torch, d) for _, _, _, d in cases]
    specs = [_spec(k, c, a, o) for (k, c, a, _), o in zip(cases, got)]
    print(f"{len(specs)} solves in one batch", flush=True)
    B.run_batch(specs)
    torch.cuda.synchronize()
    for q, (kind, c, args, _) in enumerate(cases):
        _same(torch, f"{kind} {c.nx}x{c.nz} order {c.order} mode {args[-3 if kind != 'green' else 4]}",
              ref[q], got[q])


def test_engines_one_after_another(torch_mod, mix):
    """Engines created and destroyed one after another: each destroys its streams while the
    library's caches keep the device blocks and phase tables that its solves released, and the next
    engine takes them again. Every batch must match the single solves, and torch's own allocations
    and launches between the batches must see no error left behind by the library."""
    torch = torch_mod
    from fd2d_noise_amd import batch as B
    cases, ref = mix
    for rep in range(4):
        got = [_outs(torch, d) for _, _, _, d in cases]
        B.run_batch([_spec(k, c, a, o) for (k, c, a, _), o in zip(cases, got)])
        torch.cuda.synchronize()
        for q, (kind, c, args, _) in enumerate(cases):
            _same(torch, f"engine {rep}: {kind} {c.nx}x{c.nz}", ref[q], got[q])
        del got
        x = torch.zeros((1 << 26) + rep, dtype=torch.float64, device="cuda")
        assert float(x.add_(1.0).sum()) == float(x.numel())
        torch.cuda.synchronize()


def test_engine_continuous_batching(torch_mod, mix):
    """Solves joining while others run (three waves, advanced by uneven round counts)."""
    torch = torch_mod
    from fd2d_noise_amd import batch as B
    cases, ref = mix
    got = [_outs(torch, d) for _, _, _, d in cases]
    specs = [_spec(k, c, a, o) for (k, c, a, _), o in zip(cases, got)]
    waves = [specs[0::3], specs[1::3], specs[2::3]]
    with B.Engine() as eng:
        tickets = [eng.submit(s) for s in waves[0]]
        eng.advance(37)
        tickets += [eng.submit(s) for s in waves[1]]
        eng.advance(101)
        # Which solves are still running is decided on the host: every round issues one or two
        # steps of each running solve, so after 101 rounds a solve of more than 202 steps is running
        # with 101..202 steps issued, whatever the device has completed. (Whether a finished solve's
        # last work has completed on the device, done(), depends on timing and is not asserted.)
        running = 0
        for t, sp in zip(tickets[len(waves[0]):], waves[1]):
            state, issued, total = eng.progress(t)
            n = sp.time_steps()
            if n > 2 * 101:
                assert state == abi.FD2D_SOLVE_RUNNING and total == n and 101 <= issued <= 202, (n, state, issued)
                running += 1
            else:
                assert state in (abi.FD2D_SOLVE_RUNNING, abi.FD2D_SOLVE_FINISHING, abi.FD2D_SOLVE_COMPLETE)
                if state == abi.FD2D_SOLVE_RUNNING:
                    assert 101 <= issued < n
        assert running >= 4  # the third wave joins while these run
        tickets += [eng.submit(s) for s in waves[2]]
        eng.run_until_done()
        assert all(eng.done(t) for t in tickets) and eng.active == 0
        assert all(eng.progress(t) == (abi.FD2D_SOLVE_COMPLETE, -1, -1) for t in tickets)
    order = list(range(0, len(specs), 3)) + list(range(1, len(specs), 3)) + list(range(2, len(specs), 3))
    torch.cuda.synchronize()
    for q in order:
        kind, c, _, _ = cases[q]
        _same(torch, f"wave case {q} {kind}", ref[q], got[q])


def test_engine_rejects_bad_solves(torch_mod, mix):
    torch = torch_mod
    from fd2d_noise_amd import batch as B, device as D
    cases, _ = mix
    kind, c, args, dims = cases[0]
    bad = list(args)
    bad[0] = D.to_device(np.ones((c.nx - 1, c.nz)))  # mu of the wrong size
    with B.Engine() as eng:
        with pytest.raises(abi.Fd2dError) as e:
            eng.submit(_spec(kind, c, tuple(bad), _outs(torch, dims)))
        assert e.value.code == abi.FD2D_EINVAL
        assert eng.advance(5) == 0
        with pytest.raises(abi.Fd2dError):
            eng.done(12345)


def test_calculate_data_matches_station_loop(torch_mod, tmp_path):
    """formats.calculate_data (code/calculate_data.m:180-223, stations batched) equals the
    per-station loop of single solves, and its files load back as the inversion reads them."""
    torch = torch_mod
    from fd2d_noise_amd import device as D, formats as F
    c = Config(order=4, nx=64, nz=56, nt=110, n_basis_fct=2)
    rng = np.random.default_rng(77)
    mu, rho = hetero_model(rng, c)
    array = interior_points(rng, c, 5)
    ref = array[[0, 2, 4]]
    spectrum = np.ones((c.n_sample, 1))
    sd = blob_distribution(rng, c, nb=2)
    out = F.calculate_data(c, mu, rho, spectrum, sd, array, ref, out_dir=str(tmp_path))
    rows = []
    for i in range(ref.shape[0]):
        rec = array[~np.all(array == ref[i], axis=1)]
        G = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
        D.run_forward1_green(D.to_device(mu), D.to_device(rho), D.to_device(ref[i:i + 1]), None, 0, None, None,
                             config=c, G_fft=G)
        tr = D.empty_like_dims((rec.shape[0], c.n_corr), torch.float64)
        D.run_forward2_correlation(D.to_device(mu), D.to_device(rho), G, D.to_device(spectrum), D.to_device(sd),
                                   D.to_device(rec), 0, None, None, config=c, seismograms=tr)
        rows.append(D.to_host(tr))
    ref_data = np.concatenate(rows, axis=0)
    assert out["c_data"].shape == (3 * 4, c.n_corr) and np.array_equal(out["c_data"], ref_data)
    dat = F.load_data(str(tmp_path / "data_3_ref_2.mat"))
    net = F.load_array(str(tmp_path / "array_3_ref.mat"))
    assert np.array_equal(dat["c_data"], ref_data) and dat["t"].shape == (1, c.n_corr)
    assert np.array_equal(net["ref_stat"], ref) and np.array_equal(net["array"], array)


def _unstable(c):
    """The same grid with dt 2.5 times the stability limit of its stencil (CFL): the fields grow by
    about 20x per step and overflow to Inf, then NaN, within a few dozen steps."""
    coef = {2: [1.0], 4: [9 / 8, 1 / 24], 6: [1.171875, 0.0651042, 0.0046875],
            8: [1.1962891, 0.0797526, 0.0095703, 0.0006975]}[c.order]
    vmax = np.sqrt(4.8e10 * 1.1 / 3000.0)
    dt_lim = min(c.Lx / (c.nx - 1), c.Lz / (c.nz - 1)) / (np.sqrt(2) * sum(coef) * vmax)
    return c.replace(dt=2.5 * dt_lim)


def test_nonfinite_solve_fails_alone(torch_mod, mix):
    """A solve whose time step exceeds the stability limit becomes NaN / Inf: the engine reports it
    FAILED and stops it, the solves batched with it are complete and bit-identical to their single
    runs, and the single-solve entry point returns FD2D_ENONFINITE with its outputs written."""
    torch = torch_mod
    from fd2d_noise_amd import batch as B, device as D
    cases, ref = mix
    good = [q for q, (k, c, _, _) in enumerate(cases) if k in ("green", "correlation") and c.nx < 100][:4]
    kind, c, args, dims = cases[good[0]]
    cu = _unstable(c).replace(nt=2000)
    bad_outs = _outs(torch, [((cu.nx, cu.nz, cu.n_sample), torch.complex128), None, ((3, cu.nt), torch.float64)])
    bad = B.green(*args[:4], 0, None, None, config=cu, G_fft=bad_outs[0], seismograms=bad_outs[2])
    got = [_outs(torch, cases[q][3]) for q in good]
    specs = [_spec(cases[q][0], cases[q][1], cases[q][2], o) for q, o in zip(good, got)]
    check_every = 512  # steps between the engine's periodic finiteness checks
    with B.Engine() as eng:
        tb = eng.submit(bad)
        tg = [eng.submit(s) for s in specs]
        while eng.progress(tb)[1] < check_every + 2:  # the first periodic check is queued
            eng.advance(1)
        torch.cuda.synchronize()  # ... and complete, so the next round's poll sees its flag
        eng.advance(1)
        state, issued, total = eng.progress(tb)
        print(f"unstable solve: state {state}, stopped after {issued} of {total} steps", flush=True)
        assert state == abi.FD2D_SOLVE_FAILED and total == cu.nt and issued <= check_every + 8
        eng.run_until_done()
        assert eng.progress(tb)[:2] == (abi.FD2D_SOLVE_FAILED, issued)
        assert all(eng.progress(t)[0] == abi.FD2D_SOLVE_COMPLETE for t in tg)
    torch.cuda.synchronize()
    for q, o in zip(good, got):
        _same(torch, f"beside the unstable solve: case {q}", ref[q], o)
    # the batch call names the failed solve and completes the others
    got2 = [_outs(torch, cases[q][3]) for q in good]
    with pytest.raises(abi.Fd2dError) as e:
        B.run_batch([_spec(cases[q][0], cases[q][1], cases[q][2], o) for q, o in zip(good, got2)] + [bad])
    assert e.value.code == abi.FD2D_ENONFINITE and f"solve(s) {len(good)} " in str(e.value)
    torch.cuda.synchronize()
    for q, o in zip(good, got2):
        _same(torch, f"run_batch beside the unstable solve: case {q}", ref[q], o)
    # the single-solve entry point: error code, outputs written
    seis = D.empty_like_dims((3, cu.nt), torch.float64)
    with pytest.raises(abi.Fd2dError) as e:
        D.run_forward1_green(*args[:4], 0, None, None, config=cu, G_fft=bad_outs[0], seismograms=seis)
    assert e.value.code == abi.FD2D_ENONFINITE and "stability" in str(e.value)
    assert not torch.isfinite(seis).all()


def test_engine_cancel(torch_mod, mix):
    """A cancelled solve issues no further steps; the others complete bit-identically."""
    torch = torch_mod
    from fd2d_noise_amd import batch as B
    cases, ref = mix
    sel = [q for q, (k, c, _, _) in enumerate(cases) if k == "correlation" and c.nx < 100][:3]
    got = [_outs(torch, cases[q][3]) for q in sel]
    specs = [_spec(cases[q][0], cases[q][1], cases[q][2], o) for q, o in zip(sel, got)]
    with B.Engine() as eng:
        ts = [eng.submit(s) for s in specs]
        eng.advance(10)
        assert eng.cancel(ts[1])
        state, issued, total = eng.progress(ts[1])
        assert state == abi.FD2D_SOLVE_CANCELLED and 10 <= issued <= 20 and total == specs[1].time_steps()
        eng.advance(5)
        assert eng.progress(ts[1])[1] == issued
        eng.run_until_done()
        assert eng.done(ts[1]) and not eng.cancel(ts[0])
        assert eng.progress(ts[0])[0] == abi.FD2D_SOLVE_COMPLETE
    torch.cuda.synchronize()
    for k in (0, 2):
        _same(torch, f"beside a cancelled solve: case {sel[k]}", ref[sel[k]], got[k])


def test_large_class_shared_with_partner_lane(torch_mod):
    """40 Green runs of one FD order in one batch, 20 tiles each: once 20 of them (400 tiles) run on
    their class's lane, the other 20 go to the partner lane (the lane of the other FD orders, idle
    here), whose launches run side by side with the first lane's. Every G_fft and seismogram equals
    its single solve's, bit for bit."""
    torch = torch_mod
    from fd2d_noise_amd import batch as B, device as D
    c = Config(order=4, nx=200, nz=150, nt=60, fw_nth=1, freq_samp=3, Lx=199e3, Lz=149e3)
    rng = np.random.default_rng(907)
    mu, rho = hetero_model(rng, c)
    mu_d, rho_d = D.to_device(mu), D.to_device(rho)
    rec = D.to_device(interior_points(rng, c, 2))
    n = 40
    srcs = [D.to_device(interior_points(rng, c, 1)) for _ in range(n)]
    G_b = [D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128) for _ in range(n)]
    s_b = [D.empty_like_dims((2, c.nt), torch.float64) for _ in range(n)]
    B.run_batch([B.green(mu_d, rho_d, srcs[i], rec, 0, None, None, config=c, G_fft=G_b[i], seismograms=s_b[i])
                 for i in range(n)])
    torch.cuda.synchronize()
    for i in range(n):
        G = D.empty_like_dims((c.nx, c.nz, c.n_sample), torch.complex128)
        s = D.empty_like_dims((2, c.nt), torch.float64)
        D.run_forward1_green(mu_d, rho_d, srcs[i], rec, 0, None, None, config=c, G_fft=G, seismograms=s)
        torch.cuda.synchronize()
        assert torch.equal(G, G_b[i]) and torch.equal(s, s_b[i]), f"solve {i}"
        assert float(G.abs().max()) > 0.0

"""NumPy FP64 restatement of fd2d_noise's three time-stepping solvers (TEST INFRASTRUCTURE ONLY).

A second restatement of code/run_forward1_green.m, code/run_forward2_correlation.m and
code/run_noise_adjoint.m, written from the .m files with whole-array numpy operations, independent
of oracle/fd2d_oracle.c (row loops in C): the two agree to 1e-13 on the reference's tiny set-up
(input/input_parameters.m:8-16), which catches a transcription error in either, and this one writes
the committed golden fixtures (tests/golden/). Never imported by the product path.

Arguments and outputs are those of the MATLAB functions (numpy arrays in MATLAB shape, [] = None),
with the configuration of input_parameters / input_interferometry / absorb_specs passed as a
fd2d_noise_amd.Config.
"""
from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------------------------------------
# code/propagation/
# ---------------------------------------------------------------------------------------------
def staggered_coefficients(order: int) -> np.ndarray:
    """c_m of sum_m c_m (f(x + (m - 1/2) h) - f(x - (m - 1/2) h)) / h = f'(x) + O(h^order): the
    Taylor conditions sum_m c_m (2m - 1)^(2q - 1) / 2^(2q - 2) = [q == 1], q = 1 .. order/2, solved
    here (orders 2 and 4 of the reference are 1 and 9/8, -1/24: dx_v.m:23,29)."""
    R = order // 2
    A = np.array([[(2 * m - 1) ** (2 * q - 1) / 2.0 ** (2 * q - 2) for m in range(1, R + 1)] for q in range(1, R + 1)])
    b = np.zeros(R)
    b[0] = 1.0
    return np.linalg.solve(A, b)


def dx_v(v, dx, order):
    """code/propagation/dx_v.m: (nx-1) x nz, zero outside the stencil's reach."""
    nx, nz = v.shape
    out = np.zeros((nx - 1, nz))
    if order == 2:
        out[:, :] = (v[1:, :] - v[:-1, :]) / dx
    elif order == 4:  # rows 2 .. nx-2 (1-based), dx_v.m:28-30
        out[1:nx - 2, :] = 9 * (v[2:nx - 1, :] - v[1:nx - 2, :]) / (8 * dx) - (v[3:nx, :] - v[0:nx - 3, :]) / (24 * dx)
    elif order in (6, 8):  # textbook extension: rows R .. nx-R (1-based)
        R = order // 2
        c = staggered_coefficients(order)
        i = np.arange(R - 1, nx - R)  # 0-based
        acc = np.zeros((i.size, nz))
        for m in range(1, R + 1):
            acc += c[m - 1] * (v[i + m, :] - v[i + 1 - m, :])
        out[i, :] = acc / dx
    return out


def dz_v(v, dz, order):
    """code/propagation/dz_v.m: nx x (nz-1)."""
    return dx_v(v.T, dz, order).T


def div_s(sxy, szy, dx, dz, order):
    """code/propagation/div_s.m: nx x nz."""
    nx, nz = szy.shape[0], sxy.shape[1]
    out = np.zeros((nx, nz))

    def part(s, h, n):
        o = np.zeros((n,) + s.shape[1:])
        if order == 2:  # 2 .. n-1
            o[1:n - 1] = (s[1:n - 1] - s[0:n - 2]) / h
        elif order == 4:  # 3 .. n-2
            o[2:n - 2] = 9 * (s[2:n - 2] - s[1:n - 3]) / (8 * h) - (s[3:n - 1] - s[0:n - 4]) / (24 * h)
        elif order in (6, 8):  # R+1 .. n-R (1-based)
            R = order // 2
            c = staggered_coefficients(order)
            i = np.arange(R, n - R)
            acc = np.zeros((i.size,) + s.shape[1:])
            for m in range(1, R + 1):
                acc += c[m - 1] * (s[i + m - 1] - s[i - m])
            o[i] = acc / h
        return o

    out += part(sxy, dx, nx)
    out += part(szy.T, dz, nz).T
    return out


def coordinates(c):
    """define_computational_domain.m: x = 0:dx:Lx, z = 0:dz:Lz (MATLAB colon counts)."""
    dx, dz = c.Lx / (c.nx - 1), c.Lz / (c.nz - 1)
    x = dx * np.arange(int(np.floor(c.Lx / dx + 1e-10)) + 1)
    z = dz * np.arange(int(np.floor(c.Lz / dz + 1e-10)) + 1)
    return x, z, dx, dz


def absbound(c):
    """code/propagation/init_absbound.m:16-30 (X' and Z' as nx x nz grids)."""
    x, z, _, _ = coordinates(c)
    X, Z = np.meshgrid(x, z)
    X, Z = X.T, Z.T
    w = c.absorb_width
    ab = np.ones((c.nx, c.nz))
    if c.absorb_left:
        ab = ab * ((X > w).astype(float) + np.exp(-(X - w) ** 2 / (2 * w) ** 2) * (X <= w))
    if c.absorb_right:
        ab = ab * ((X < c.Lx - w).astype(float) + np.exp(-(X - (c.Lx - w)) ** 2 / (2 * w) ** 2) * (X >= c.Lx - w))
    if c.absorb_bottom:
        ab = ab * ((Z > w).astype(float) + np.exp(-(Z - w) ** 2 / (2 * w) ** 2) * (Z <= w))
    if c.absorb_top:
        ab = ab * ((Z < c.Lz - w).astype(float) + np.exp(-(Z - (c.Lz - w)) ** 2 / (2 * w) ** 2) * (Z >= c.Lz - w))
    return ab


def node_ids(pos, c):
    """Nearest grid node per axis, the lowest index on ties (run_forward1_green.m:45-60), 0-based."""
    if pos is None or np.size(pos) == 0:
        return np.zeros((0, 2), dtype=int)
    x, z, _, _ = coordinates(c)
    pos = np.atleast_2d(pos)
    return np.array([[int(np.argmin(np.abs(x - p[0]))), int(np.argmin(np.abs(z - p[1])))] for p in pos])


def interpol_matrix(K, nb):
    """The band interpolation of run_forward2_correlation.m:121-165 (integration_limits.m:2-14)."""
    if nb == 0:
        return np.ones((K, 1))
    step = K / nb
    lim = np.array([[np.floor(i * step + 1), np.floor((i + 1) * step)] for i in range(nb)], dtype=int)
    Im = np.zeros((K, nb))
    for k in range(1, K + 1):
        ib = int(np.nonzero((k >= lim[:, 0]) & (k <= lim[:, 1]))[0][0])  # 0-based band
        idx = np.arange(lim[ib, 0], lim[ib, 1] + 1)
        ni = idx.size
        i = int(np.nonzero(idx == k)[0][0]) + 1  # 1-based position in the band
        below = ib - 1 if ib > 0 else None
        above = ib + 1 if ib < nb - 1 else None
        if i < ni / 2:
            Im[k - 1, ib] = 1 - (ni / 2 - i) / ni
            if below is not None:
                Im[k - 1, below] = (ni / 2 - i) / ni
        elif i == ni / 2:
            Im[k - 1, ib] = 1
        else:
            Im[k - 1, ib] = 1 - (i - ni / 2) / ni
            if above is not None:
                Im[k - 1, above] = (i - ni / 2) / ni
    return Im


def _empty(a):
    return a is None or np.size(a) == 0


def _colon(a, step, b):
    n = int(np.floor((b - a) / step + 1e-10)) + 1
    return a + step * np.arange(n)


# ---------------------------------------------------------------------------------------------
# code/run_forward1_green.m
# ---------------------------------------------------------------------------------------------
def run_forward1_green(mu, rho, src, rec, mode, dmu, G_in, *, config):
    c = config
    nx, nz, nt, dt, f, s, K = c.nx, c.nz, c.nt, c.dt, c.fw_nth, c.freq_samp, c.n_sample
    _, _, dx, dz = coordinates(c)
    mu, rho = np.reshape(mu, (nx, nz), order="F"), np.reshape(rho, (nx, nz), order="F")
    w = c.w_sample
    t = _colon(0.0, dt, (nt - 1) * dt)
    sid, rid = node_ids(src, c), node_ids(rec, c)
    stf = 1.0e9 * np.ones(nt)
    n_idx = [n for n in range(nt, 2 * nt) if n % s == 0]
    fft_coeff = np.array([1 / np.sqrt(2 * np.pi) * np.exp(-1j * w * t[n - nt]) * dt for n in n_idx]).reshape(-1, K)
    n_fw = sum(1 for n in range(nt, 2 * nt) if n % f == 0)
    v, sxy, szy = np.zeros((nx, nz)), np.zeros((nx - 1, nz)), np.zeros((nx, nz - 1))
    G_fft = np.zeros((nx, nz, K), dtype=complex)
    G_out = np.zeros((nx, nz, n_fw), dtype=np.float32) if mode != 0 else np.float32(0.0)
    seis = np.zeros((len(rid), nt))
    ab = absbound(c)
    i_ftc = i_out = i_in = 0
    born = not _empty(dmu)
    if born:
        dmu = np.reshape(dmu, (nx, nz), order="F")
    for n in range(1, nt + 1):
        if mode != 0 and (n + nt - 1) % f == 0:
            G_out[:, :, i_out] = v.astype(np.float32)
            i_out += 1
        if born:
            g = G_in[:, :, i_in]  # single: its stencil is evaluated in single precision
            sxy = sxy + dt * dmu[:nx - 1, :] * _dv_single(g, dx, c.order, 0)
            szy = szy + dt * dmu[:, :nz - 1] * _dv_single(g, dz, c.order, 1)
            i_in += 1
        DS = div_s(sxy, szy, dx, dz, c.order)
        if not born:
            for q in range(len(sid)):
                DS[sid[q, 0], sid[q, 1]] += stf[n - 1]
        v = v + dt * DS / rho
        v = v * ab
        sxy = sxy + dt * mu[:nx - 1, :] * dx_v(v, dx, c.order)
        szy = szy + dt * mu[:, :nz - 1] * dz_v(v, dz, c.order)
        if (n + nt - 1) % s == 0:
            G_fft += v[:, :, None] * fft_coeff[i_ftc][None, None, :]
            i_ftc += 1
        for r in range(len(rid)):
            seis[r, n - 1] = v[rid[r, 0], rid[r, 1]]
    return np.conj(G_fft), G_out, seis


# ---------------------------------------------------------------------------------------------
# code/run_forward2_correlation.m
# ---------------------------------------------------------------------------------------------
def run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in, *, config):
    c = config
    nx, nz, dt, f, s, K, nb = c.nx, c.nz, c.dt, c.fw_nth, c.freq_samp, c.n_sample, c.n_basis_fct
    _, _, dx, dz = coordinates(c)
    mu, rho = np.reshape(mu, (nx, nz), order="F"), np.reshape(rho, (nx, nz), order="F")
    w, dw = c.w_sample, c.dw
    spectrum = np.atleast_2d(np.asarray(spectrum, dtype=float).reshape(K, -1, order="F"))
    n_src = spectrum.shape[1]
    sd = np.reshape(source_dist, (nx, nz, nb if nb else n_src), order="F")
    rid = node_ids(rec, c)
    t = _colon(-(c.nt - 1) * dt, dt, (c.nt - 1) * dt)
    N = t.size
    ifft_coeff = np.array([1 / np.sqrt(2 * np.pi) * np.exp(1j * w * t[n - 1]) * dw for n in range(1, N + 1) if n % s == 0])
    v, sxy, szy, u = np.zeros((nx, nz)), np.zeros((nx - 1, nz)), np.zeros((nx, nz - 1)), np.zeros((nx, nz))
    C_out = np.zeros((nx, nz, N // f), dtype=np.float32) if mode != 0 else np.float32(0.0)
    seis = np.zeros((len(rid), N))
    ab = absbound(c)
    Im = interpol_matrix(K, nb)
    dist = np.einsum("kb,xzb->xzk", Im, sd)  # distribution(ix,iz,:) = interpol_matrix * sd(ix,iz,:)
    born = not _empty(dmu)
    if born:
        dmu = np.reshape(dmu, (nx, nz), order="F")
    i_ftc = i_out = 0
    for n in range(1, N + 1):
        if mode != 0 and n % f == 0:
            C_out[:, :, i_out] = u.astype(np.float32)
            i_out += 1
        DS = div_s(sxy, szy, dx, dz, c.order)
        if n % s == 0:
            for q in range(n_src):
                S = np.einsum("k,xzk,xzk,k->xz", spectrum[:, q], dist, G_fft, ifft_coeff[i_ftc])
                DS = DS + S.real
            i_ftc += 1
        v = v + dt * DS / rho
        v = v * ab
        ex, ez = dx_v(v, dx, c.order), dz_v(v, dz, c.order)
        if n == N or not born:
            sxy = sxy + dt * mu[:nx - 1, :] * ex
            szy = szy + dt * mu[:, :nz - 1] * ez
        else:
            d = C_in[:, :, n - 1] - C_in[:, :, n]  # the difference of single arrays, and its stencil, in single
            sxy = sxy + dt * mu[:nx - 1, :] * ex - dmu[:nx - 1, :] * _dv_single(d, dx, c.order, 0)
            szy = szy + dt * mu[:, :nz - 1] * ez - dmu[:, :nz - 1] * _dv_single(d, dz, c.order, 1)
        u = u + v * dt
        for r in range(len(rid)):
            seis[r, n - 1] = u[rid[r, 0], rid[r, 1]]
    return seis, C_out


# ---------------------------------------------------------------------------------------------
# code/run_noise_adjoint.m
# ---------------------------------------------------------------------------------------------
def run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode, dmu, u_in, *, config):
    c = config
    nx, nz, dt, f, s, K, nb = c.nx, c.nz, c.dt, c.fw_nth, c.freq_samp, c.n_sample, c.n_basis_fct
    _, _, dx, dz = coordinates(c)
    mu, rho = np.reshape(mu, (nx, nz), order="F"), np.reshape(rho, (nx, nz), order="F")
    w, dw = c.w_sample, c.dw
    t = _colon(-(c.nt - 1) * dt, dt, (c.nt - 1) * dt)
    N = t.size
    sid, rid = node_ids(src, c), node_ids(rec, c)
    stf = None if _empty(stf) else np.asarray(stf)
    stf_dist = stf is not None and stf.ndim == 3 and stf.shape[2] != 1
    coeff_n = [n for n in range(1, N + 1) if n % s == 0]
    fft_coeff = np.array([1 / np.sqrt(2 * np.pi) * np.exp(-1j * w * t[n - 1]) * dt for n in coeff_n])
    ifft_coeff = np.array([1 / np.sqrt(2 * np.pi) * np.exp(1j * w * t[n - 1]) * dw for n in coeff_n])
    ifft_coeff2 = 1 / np.sqrt(2 * np.pi) * np.exp(1j * np.outer(t, w)) * dw
    Im = interpol_matrix(K, nb)
    if not _empty(source_dist):
        spectrum = np.asarray(spectrum, dtype=float).reshape(K, -1, order="F")
        n_src = spectrum.shape[1]
        sd = np.reshape(source_dist, (nx, nz, nb if nb else n_src), order="F")
        dist = np.einsum("kb,xzb->xzk", Im, sd)
    else:
        n_src, dist = 0, None
    v, sxy, szy, u = np.zeros((nx, nz)), np.zeros((nx - 1, nz)), np.zeros((nx, nz - 1)), np.zeros((nx, nz))
    stf_fft = np.zeros((nx, nz, K), dtype=complex)
    u_out = np.zeros((nx, nz, N // f), dtype=np.float32) if mode in (11, 13) else np.float32(0.0)
    seis = np.zeros((len(rid), N))
    K_mu = np.zeros((nx, nz))
    K_s = np.zeros((nx, nz, max(nb, 1)))
    ab = absbound(c)
    nsteps = c.nt if (mode in (0, 1) and stf_dist) else N
    n_ufwd = 0 if _empty(u_fwd) else u_fwd.shape[2] if np.ndim(u_fwd) == 3 else 1
    born = not _empty(dmu) and not _empty(u_in) and np.ndim(u_in) == 3 and u_in.shape[2] == N
    if not _empty(dmu):
        dmu = np.reshape(dmu, (nx, nz), order="F")
    src_kernel = not _empty(G_fft) and np.ndim(G_fft) == 3 and G_fft.shape[2] == K
    if src_kernel:
        spec_k = np.asarray(spectrum, dtype=float).reshape(K)
    i_ftc = i_in = i_out = 0
    for n in range(1, nsteps + 1):
        if mode in (11, 13) and n % f == 0:
            u_out[:, :, i_out] = u.astype(np.float32)
            i_out += 1
        DS = div_s(sxy, szy, dx, dz, c.order)
        if stf is not None and not stf_dist:
            for q in range(len(sid)):
                DS[sid[q, 0], sid[q, 1]] += np.real(stf[q, n - 1])
        elif stf_dist and n % s == 0:
            T = np.zeros((nx, nz), dtype=complex)
            for q in range(n_src):
                T += np.einsum("k,xzk,xzk,k->xz", spectrum[:, q], dist, np.conj(stf), ifft_coeff[i_ftc])
            i_ftc += 1
            DS = DS + T.real
        v = v + dt * DS / rho
        v = v * ab
        ex, ez = dx_v(v, dx, c.order), dz_v(v, dz, c.order)
        if n == N or not born:
            sxy = sxy + dt * mu[:nx - 1, :] * ex
            szy = szy + dt * mu[:, :nz - 1] * ez
        else:
            d = u_in[:, :, n - 1] - u_in[:, :, n]
            sxy = sxy + dt * mu[:nx - 1, :] * ex - dmu[:nx - 1, :] * _dv_single(d, dx, c.order, 0)
            szy = szy + dt * mu[:, :nz - 1] * ez - dmu[:, :nz - 1] * _dv_single(d, dz, c.order, 1)
        u = u + v * dt
        exu, ezu = dx_v(u, dx, c.order), dz_v(u, dz, c.order)
        if src_kernel and n % s == 0:
            # M_tn(ix,iz,:) = interpol_matrix' * (spectrum .* conj(G_fft(ix,iz,:)) .* ifft_coeff2(n,:).')
            M = np.einsum("kb,k,xzk,k->xzb", Im, spec_k, np.conj(G_fft), ifft_coeff2[n - 1])
            K_s += np.real(M * u[:, :, None])
        if n_ufwd and n_ufwd >= i_in + 1 and n % f == 0:
            g = u_fwd[:, :, n_ufwd - 1 - i_in]
            # dx_v of a single array: single arithmetic, as MATLAB evaluates it
            K_mu[:nx - 1, :] -= exu * _dv_single(g, dx, c.order, 0) * f
            K_mu[:, :nz - 1] -= ezu * _dv_single(g, dz, c.order, 1) * f
            i_in += 1
        if mode in (1, 12, 13) and n % s == 0:
            stf_fft += u[:, :, None] * fft_coeff[i_ftc][None, None, :]
            i_ftc += 1
        for r in range(len(rid)):
            seis[r, n - 1] = u[rid[r, 0], rid[r, 1]]
    Kout = np.zeros((nx, nz, max(nb, 1) + 1))
    Kout[:, :, :max(nb, 1)] = K_s
    Kout[:, :, -1] = K_mu
    return Kout, stf_fft, u_out, seis


def _dv_single(g, h, order, axis):
    """dx_v / dz_v evaluated in single precision (MATLAB keeps a single array single): the stencil
    with float32 operands and the float32 spacing, then widened."""
    g = np.asarray(g, dtype=np.float32)
    if axis == 1:
        return _dv_single(g.T, h, order, 0).T
    nx, nz = g.shape
    out = np.zeros((nx - 1, nz), dtype=np.float32)
    hf = np.float32(h)
    if order == 2:
        out[:, :] = (g[1:, :] - g[:-1, :]) / hf
    elif order == 4:
        out[1:nx - 2, :] = (np.float32(9) * (g[2:nx - 1, :] - g[1:nx - 2, :]) / np.float32(8 * h)
                            - (g[3:nx, :] - g[0:nx - 3, :]) / np.float32(24 * h))
    else:
        R = order // 2
        cf = staggered_coefficients(order).astype(np.float32)
        i = np.arange(R - 1, nx - R)
        acc = np.zeros((i.size, nz), dtype=np.float32)
        for m in range(1, R + 1):
            acc += cf[m - 1] * (g[i + m, :] - g[i + 1 - m, :])
        out[i, :] = acc / hf
    return out.astype(np.float64)

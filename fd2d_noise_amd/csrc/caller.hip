// caller.hip -- FP64 device kernels of the solvers' callers: misfits and adjoint sources
// (tools/measurement/), compensated inner products and vector updates for the L-BFGS driver
// (inversion/optlib/), and the Gaussian parameter map (tools/addons/gaussblur2d.m).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "caller.h"

namespace fd2d {

// ---------------------------------------------------------------------------------------------
// Compensated (Neumaier) FP64 sums and a 256-thread block reduction
// ---------------------------------------------------------------------------------------------
struct CSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double x)
    {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

// Sum of one value per thread over a 256-thread block (compensated); every thread gets the result.
__device__ double block_sum(double v, double* lds_s, double* lds_c)
{
    CSum a;
    a.add(v);
    lds_s[threadIdx.x] = a.s;
    lds_c[threadIdx.x] = a.c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            CSum b;
            b.s = lds_s[threadIdx.x];
            b.c = lds_c[threadIdx.x];
            b.add(lds_s[threadIdx.x + off]);
            b.c += lds_c[threadIdx.x + off];
            lds_s[threadIdx.x] = b.s;
            lds_c[threadIdx.x] = b.c;
        }
        __syncthreads();
    }
    const double r = lds_s[0] + lds_c[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------------------------
// Adjoint sources (make_adjoint_sources.m:31-175)
// ---------------------------------------------------------------------------------------------
struct Series {
    const double* p;
    int64_t stride;  // receivers: element (i) of the row at p[i * stride]
    __device__ double operator[](int64_t i) const { return p[i * stride]; }
};

// get_window.m:26-34 ('cos_taper') at time t.
__device__ double cos_taper(double t, double t_min, double t_max)
{
    if (t_min > t_max) {
        const double q = t_min;
        t_min = t_max;
        t_max = q;
    }
    const double width = fabs(t_max - t_min) * 0.1;
    double w = (double)(t > t_min) * (double)(t < t_max);
    w += (0.5 + 0.5 * cos(M_PI * (t_max - t) / width)) * (double)(t >= t_max) * (double)(t < t_max + width);
    w += (0.5 + 0.5 * cos(M_PI * (t_min - t) / width)) * (double)(t > t_min - width) * (double)(t <= t_min);
    return w;
}

struct AdjCtx {
    int64_t nt;
    double dt;
};

// Signal value (displacement or, for 'vel', its forward difference, make_adjoint_sources.m:40-52).
__device__ double sig(const Series& s, int64_t i, int vel, const AdjCtx& c)
{
    if (!vel) return s[i];
    return i < c.nt - 1 ? (s[i + 1] - s[i]) / c.dt : 0.0;
}

__device__ double block_max_first(double v, int64_t idx, double* lds_v, int64_t* lds_i, int64_t* out_idx)
{
    lds_v[threadIdx.x] = v;
    lds_i[threadIdx.x] = idx;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            const double a = lds_v[threadIdx.x], b = lds_v[threadIdx.x + off];
            const int64_t ia = lds_i[threadIdx.x], ib = lds_i[threadIdx.x + off];
            if (b > a || (b == a && ib < ia)) {
                lds_v[threadIdx.x] = b;
                lds_i[threadIdx.x] = ib;
            }
        }
        __syncthreads();
    }
    const double r = lds_v[0];
    *out_idx = lds_i[0];
    __syncthreads();
    return r;
}

// One workgroup per receiver.
__global__ __launch_bounds__(256) void k_adjoint_sources(AdjArgs a)
{
    __shared__ double lds_s[256], lds_c[256];
    __shared__ int64_t lds_i[256];
    __shared__ int sh_flag;
    const int64_t n = blockIdx.x;
    const int64_t nt = a.nt;
    const Series U{a.u + n, a.nr}, U0{a.u0 + n, a.nr}, DU{a.du ? a.du + n : nullptr, a.nr};
    double* out = a.adstf + n;  // adstf(n, i) at out[i * nr]
    AdjCtx c{nt, a.dt};
    const double dt = a.dt;
    const int vel = a.veldis;
    auto T = [&](int64_t i) { return a.t[i]; };
    auto U_ = [&](int64_t i) { return sig(U, i, vel, c); };
    auto U0_ = [&](int64_t i) { return sig(U0, i, vel, c); };
    auto DU_ = [&](int64_t i) { return DU.p ? DU[i] : 0.0; };

    // sum(u_0 == 0) == length(t): data-independent measurement
    if (threadIdx.x == 0) sh_flag = 1;
    __syncthreads();
    for (int64_t i = threadIdx.x; i < nt; i += 256)
        if (U0_(i) != 0.0) sh_flag = 0;
    __syncthreads();
    const bool u0_zero = sh_flag != 0;

    for (int64_t i = threadIdx.x; i < nt; i += 256) out[i * a.nr] = 0.0;
    __syncthreads();

    double misfit = 0.0;
    const double left = a.left, right = a.right;
    if (a.measurement == FD2D_WAVEFORM_DIFFERENCE) {
        // waveform_difference.m:23-48 with the window of make_adjoint_sources.m:70-73
        const double tl = T(0), tr = T(nt - 1);
        double part = 0.0;
        CSum acc;
        for (int64_t i = threadIdx.x; i < nt; i += 256) {
            const double w = cos_taper(T(i), tl, tr);
            const double d = U_(i) - U0_(i);
            acc.add(w * w * d * d);
            const double ad = a.deriv_order == 1 ? w * w * d * dt : w * w * DU_(i) * dt;
            out[(nt - 1 - i) * a.nr] = ad;
        }
        part = acc.value();
        const double s = block_sum(part, lds_s, lds_c);
        misfit = a.deriv_order == 1 ? 0.5 * s * dt : 0.0;
    } else if (a.measurement == FD2D_CC_TIME_SHIFT || a.measurement == FD2D_AMPLITUDE_DIFFERENCE) {
        // causal window [left, right], then acausal [-right, -left] (make_adjoint_sources.m:126-140)
        for (int branch = 0; branch < 2; ++branch) {
            const double wl = branch == 0 ? left : -right, wr = branch == 0 ? right : -left;
            if (a.measurement == FD2D_CC_TIME_SHIFT) {
                // cc_time_shift.m:21-59
                double Tshift = 1.0;
                if (!u0_zero) {
                    // support of the window
                    int64_t lo = nt, hi = -1;
                    for (int64_t i = threadIdx.x; i < nt; i += 256)
                        if (cos_taper(T(i), wl, wr) != 0.0) {
                            lo = lo < i ? lo : i;
                            hi = hi > i ? hi : i;
                        }
                    int64_t tmp;
                    const double lo_d = -block_max_first(-(double)lo, 0, lds_s, lds_i, &tmp);
                    const double hi_d = block_max_first((double)hi, 0, lds_s, lds_i, &tmp);
                    const int64_t a0 = (int64_t)lo_d, b0 = (int64_t)hi_d;
                    int64_t best_i = 0;
                    double best = 0.0;
                    if (b0 >= a0) {
                        const int64_t L = b0 - a0 + 1;
                        double* xs = a.scratch + 3 * n * nt;  // win .* u_0 over the support
                        double* ys = xs + nt;                 // win .* u
                        for (int64_t i = a0 + threadIdx.x; i <= b0; i += 256) {
                            const double wv = cos_taper(T(i), wl, wr);
                            xs[i - a0] = wv * U0_(i);
                            ys[i - a0] = wv * U_(i);
                        }
                        __syncthreads();
                        // cc(m) = sum_i x(i+m) y(i) (xcorr(win.*u_0, win.*u)), first maximum
                        double bv = -DBL_MAX;
                        int64_t bi = 0;
                        for (int64_t m = -(L - 1) + (int64_t)threadIdx.x; m <= L - 1; m += 256) {
                            CSum s;
                            const int64_t i_lo = m < 0 ? -m : 0, i_hi = m < 0 ? L - 1 : L - 1 - m;
                            for (int64_t i = i_lo; i <= i_hi; ++i) s.add(xs[i + m] * ys[i]);
                            const double v = s.value();
                            if (v > bv) {
                                bv = v;
                                bi = m + (nt - 1);
                            }
                        }
                        best = block_max_first(bv, bi, lds_s, lds_i, &best_i);
                        // lags outside the support are exactly zero; the first of them is index 0
                        if (L < nt && best <= 0.0) {
                            best_i = 0;
                            best = 0.0;
                        }
                    }
                    Tshift = (double)(best_i - (nt - 1)) * dt;
                    if (fabs(Tshift) > 3.5) Tshift = 0.0;
                }
                misfit += Tshift * Tshift / 2.0;
                // v = diff(u)/dt; adstf = T * win.^2 .* v / (sum((win.*v).^2) * dt)
                CSum e;
                for (int64_t i = threadIdx.x; i < nt; i += 256) {
                    const double v = i < nt - 1 ? (U_(i + 1) - U_(i)) / dt : 0.0;
                    const double w = cos_taper(T(i), wl, wr);
                    e.add((w * v) * (w * v));
                }
                const double den = block_sum(e.value(), lds_s, lds_c) * dt;
                for (int64_t i = threadIdx.x; i < nt; i += 256) {
                    const double v = i < nt - 1 ? (U_(i + 1) - U_(i)) / dt : 0.0;
                    const double w = cos_taper(T(i), wl, wr);
                    out[(nt - 1 - i) * a.nr] += Tshift * (w * w * v) / den;
                }
            } else {
                // amp_diff.m:23-67
                CSum e0, e1, e2, e3;
                for (int64_t i = threadIdx.x; i < nt; i += 256) {
                    const double w = cos_taper(T(i), wl, wr);
                    const double u = U_(i), u0 = U0_(i);
                    e0.add((w * u0) * (w * u0));
                    e1.add((w * u) * (w * u));
                    e2.add(u * u);
                    e3.add(w * w * u * DU_(i));
                }
                const double s0 = block_sum(e0.value(), lds_s, lds_c);
                const double s1 = block_sum(e1.value(), lds_s, lds_c);
                const double s2 = block_sum(e2.value(), lds_s, lds_c);
                const double s3 = block_sum(e3.value(), lds_s, lds_c);
                const double energy_0 = s0 * dt;
                const double meas = (s1 * dt - energy_0) / (energy_0 + DBL_EPSILON);
                for (int64_t i = threadIdx.x; i < nt; i += 256) {
                    const double w = cos_taper(T(i), wl, wr);
                    double ad;
                    if (a.deriv_order == 1) {
                        // data-independent branch flips u once inside and once at the end
                        ad = u0_zero ? 2.0 * U_(nt - 1 - i) / (s2 * dt) : 2.0 * meas * w * w * U_(i) / s0;
                    } else {
                        ad = 2.0 * meas * w * w * DU_(i) / energy_0 * dt +
                             4.0 * w * w * U_(i) / (energy_0 * energy_0) * s3 * dt * dt;
                    }
                    out[(nt - 1 - i) * a.nr] += ad;
                }
                if (a.deriv_order == 1) misfit += u0_zero ? 1.0 : meas * meas / 2.0;
            }
            __syncthreads();
        }
    } else {
        // log_amp_ratio.m:23-93 with the Hann window of get_window.m:19-23 on [left, right]
        // support of the box (t > t_min) .* (t < t_max)
        const double tmin = left < right ? left : right, tmax = left < right ? right : left;
        int64_t lo = nt, hi = -1;
        for (int64_t i = threadIdx.x; i < nt; i += 256)
            if (T(i) > tmin && T(i) < tmax) {
                lo = lo < i ? lo : i;
                hi = hi > i ? hi : i;
            }
        int64_t tmp;
        const int64_t i1 = (int64_t)(-block_max_first(-(double)lo, 0, lds_s, lds_i, &tmp));
        const int64_t i2 = (int64_t)block_max_first((double)hi, 0, lds_s, lds_i, &tmp);
        const int64_t Lh = i2 - i1 + 1;
        auto hann = [&](int64_t i) -> double {
            if (i < i1 || i > i2) return 0.0;
            if (Lh == 1) return 1.0;
            return 0.5 * (1.0 - cos(2.0 * M_PI * (double)(i - i1) / (double)(Lh - 1)));
        };
        // causal / acausal by the sign of t at the window's first maximum
        double wmax = -1.0;
        int64_t imax = 0;
        for (int64_t i = threadIdx.x; i < nt; i += 256) {
            const double w = hann(i);
            if (w > wmax) {
                wmax = w;
                imax = i;
            }
        }
        block_max_first(wmax, imax, lds_s, lds_i, &imax);
        const bool flip = T(imax) < 0.0;
        auto wc = [&](int64_t i) { return flip ? hann(nt - 1 - i) : hann(i); };
        auto wa = [&](int64_t i) { return flip ? hann(i) : hann(nt - 1 - i); };
        // trapz over unit spacing: sum minus half the end points
        CSum ec, ea, e0c, e0a, dc, da;
        for (int64_t i = threadIdx.x; i < nt; i += 256) {
            const double h = (i == 0 || i == nt - 1) ? 0.5 : 1.0;
            const double u = U_(i), u0 = U0_(i), du = DU_(i);
            ec.add(h * (wc(i) * u) * (wc(i) * u));
            ea.add(h * (wa(i) * u) * (wa(i) * u));
            e0c.add(h * (wc(i) * u0) * (wc(i) * u0));
            e0a.add(h * (wa(i) * u0) * (wa(i) * u0));
            dc.add(h * 2.0 * wc(i) * wc(i) * u * du);
            da.add(h * 2.0 * wa(i) * wa(i) * u * du);
        }
        const double e_c = block_sum(ec.value(), lds_s, lds_c) * dt;
        const double e_a = block_sum(ea.value(), lds_s, lds_c) * dt + DBL_EPSILON;
        const double e0_c = block_sum(e0c.value(), lds_s, lds_c) * dt;
        const double e0_a = block_sum(e0a.value(), lds_s, lds_c) * dt + DBL_EPSILON;
        const double de_c2 = block_sum(dc.value(), lds_s, lds_c) * dt;
        const double de_a2 = block_sum(da.value(), lds_s, lds_c) * dt;
        const double A = log(e_c / e_a), A0 = log(e0_c / e0_a);
        if (a.deriv_order == 1) misfit = 0.5 * (A - A0) * (A - A0);
        for (int64_t i = threadIdx.x; i < nt; i += 256) {
            const double u = U_(i), du = DU_(i), w_c = wc(i), w_a = wa(i);
            double ad;
            if (a.deriv_order == 1) {
                const double dec = 2.0 * w_c * w_c * u * dt, dea = 2.0 * w_a * w_a * u * dt;
                ad = u0_zero ? dec / e_c - dea / e_a : (A - A0) * (dec / e_c - dea / e_a);
            } else {
                ad = (A - A0) * (2.0 / e_c * w_c * w_c - 2.0 / e_a * w_a * w_a) * du * dt +
                     (A - A0) * (-2.0 / (e_c * e_c) * w_c * w_c * u * de_c2 + 2.0 / (e_a * e_a) * w_a * w_a * u * de_a2) * dt +
                     (2.0 / e_c * w_c * w_c * u - 2.0 / e_a * w_a * w_a * u) * (de_c2 / e_c - de_a2 / e_a) * dt;
            }
            out[(nt - 1 - i) * a.nr] = ad;
        }
    }
    __syncthreads();
    // velocity measurements: adstf(1:nt-1) = -diff(adstf)/dt (make_adjoint_sources.m:166-168)
    if (vel) {
        double* tmp = a.scratch + 3 * n * nt + 2 * nt;
        for (int64_t i = threadIdx.x; i < nt - 1; i += 256) tmp[i] = -(out[(i + 1) * a.nr] - out[i * a.nr]) / dt;
        __syncthreads();
        for (int64_t i = threadIdx.x; i < nt - 1; i += 256) out[i * a.nr] = tmp[i];
    }
    if (threadIdx.x == 0) a.misfit[n] = misfit;
}

void launch_adjoint_sources(const AdjArgs& a, hipStream_t s)
{
    hipLaunchKernelGGL(k_adjoint_sources, dim3((unsigned)a.nr), dim3(256), 0, s, a);
}

// ---------------------------------------------------------------------------------------------
// Compensated dot product and vector updates (optlib_lbfgs.m, optlib_wolfe.m)
// ---------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_dot_partial(const double* __restrict__ x, const double* __restrict__ y,
                                                     int64_t n, double* __restrict__ part)
{
    __shared__ double ls[256], lc[256];
    CSum acc;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        // exact product split: p + e = x*y (FMA), both terms summed compensated
        const double p = x[i] * y[i];
        const double e = fma(x[i], y[i], -p);
        acc.add(p);
        acc.add(e);
    }
    ls[threadIdx.x] = acc.s;
    lc[threadIdx.x] = acc.c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            CSum b;
            b.s = ls[threadIdx.x];
            b.c = lc[threadIdx.x];
            b.add(ls[threadIdx.x + off]);
            b.c += lc[threadIdx.x + off];
            ls[threadIdx.x] = b.s;
            lc[threadIdx.x] = b.c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ls[0];
        part[2 * blockIdx.x + 1] = lc[0];
    }
}

__global__ __launch_bounds__(256) void k_dot_final(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    __shared__ double ls[256], lc[256];
    CSum acc;
    for (int i = threadIdx.x; i < nb; i += 256) {
        acc.add(part[2 * i]);
        acc.c += part[2 * i + 1];
    }
    ls[threadIdx.x] = acc.s;
    lc[threadIdx.x] = acc.c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            CSum b;
            b.s = ls[threadIdx.x];
            b.c = lc[threadIdx.x];
            b.add(ls[threadIdx.x + off]);
            b.c += lc[threadIdx.x + off];
            ls[threadIdx.x] = b.s;
            lc[threadIdx.x] = b.c;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = ls[0] + lc[0];
}

void launch_dot(const double* x, const double* y, int64_t n, double* part, int nb, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_dot_partial, dim3((unsigned)nb), dim3(256), 0, s, x, y, n, part);
    hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(256), 0, s, part, nb, out);
}

__global__ void k_axpby(double alpha, const double* __restrict__ x, double beta, const double* __restrict__ y,
                        double* __restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = alpha * x[i] + beta * y[i];
}

void launch_axpby(double alpha, const double* x, double beta, const double* y, double* out, int64_t n, hipStream_t s)
{
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_axpby, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, s, alpha, x, beta, y, out, n);
}

__global__ void k_axpby_f32(float alpha, const float* __restrict__ x, float beta, const float* __restrict__ y,
                            float* __restrict__ out, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = alpha * x[i] + beta * y[i];
}

void launch_axpby_f32(float alpha, const float* x, float beta, const float* y, float* out, int64_t n, hipStream_t s)
{
    const int64_t nb = (n + 255) / 256;
    hipLaunchKernelGGL(k_axpby_f32, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, s, alpha, x, beta, y, out,
                       n);
}

// ---------------------------------------------------------------------------------------------
// gaussblur2d.m:2-38: separable Gaussian blur truncated at sqrt(14) sigma, normalised by the
// outer product of the row sums of the two 1-D weight matrices.
// ---------------------------------------------------------------------------------------------
// pass 1: Mtmp(i,:) = sum_j w1(i,j) Y(j,:) over the first dimension (n1 x n2, column-major)
__global__ void k_blur_dim1(const double* __restrict__ Y, double* __restrict__ M, const double* __restrict__ x1,
                            int n1, int n2, int ldy, double sigma)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x, k = blockIdx.y;
    if (i >= n1 || k >= n2) return;
    const double cut = sqrt(14.0) * sigma;
    double acc = 0.0;
    for (int j = 0; j < n1; ++j) {
        const double d = fabs(x1[i] - x1[j]);
        if (d < cut) acc += exp(-(d / sigma) * (d / sigma) / 2.0) * Y[j + (int64_t)ldy * k];
    }
    M[i + (int64_t)n1 * k] = acc;
}

// pass 2: GY(:,i) = sum_j w2(i,j) Mtmp(:,j), divided by w1(.)' * w2(i)
__global__ void k_blur_dim2(const double* __restrict__ M, double* __restrict__ GY, const double* __restrict__ x1,
                            const double* __restrict__ x2, int n1, int n2, int ldg, double s1, double s2)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
    if (r >= n1 || i >= n2) return;
    const double c1 = sqrt(14.0) * s1, c2 = sqrt(14.0) * s2;
    double acc = 0.0, w2 = 0.0;
    for (int j = 0; j < n2; ++j) {
        const double d = fabs(x2[i] - x2[j]);
        if (d < c2) {
            const double w = exp(-(d / s2) * (d / s2) / 2.0);
            acc += w * M[r + (int64_t)n1 * j];
            w2 += w;
        }
    }
    double w1 = 0.0;
    for (int j = 0; j < n1; ++j) {
        const double d = fabs(x1[r] - x1[j]);
        if (d < c1) w1 += exp(-(d / s1) * (d / s1) / 2.0);
    }
    GY[r + (int64_t)ldg * i] = acc / (w1 * w2);
}

void launch_gaussblur(const double* Y, int ldy, double* GY, int ldg, double* tmp, const double* x1,
                      const double* x2, int n1, int n2, double s1, double s2, hipStream_t s)
{
    hipLaunchKernelGGL(k_blur_dim1, dim3((unsigned)((n1 + 127) / 128), (unsigned)n2), dim3(128), 0, s, Y, tmp, x1, n1,
                       n2, ldy, s1);
    hipLaunchKernelGGL(k_blur_dim2, dim3((unsigned)((n1 + 127) / 128), (unsigned)n2), dim3(128), 0, s, tmp, GY, x1, x2,
                       n1, n2, ldg, s1, s2);
}

// ---------------------------------------------------------------------------------------------
// Basis-function source maps (code/propagation/general_source.m:13-24), one thread per node and band
// ---------------------------------------------------------------------------------------------
__global__ void k_general_source(const double* __restrict__ spec, int K, int n_src, const double* __restrict__ dist,
                                 int64_t P, const int* __restrict__ k_lo, const int* __restrict__ k_hi,
                                 double* __restrict__ out)
{
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (x >= P) return;
    const double ni = (double)(k_hi[b] - k_lo[b] + 1);
    double acc = 0.0;
    for (int k = k_lo[b]; k <= k_hi[b]; ++k)
        for (int ns = 0; ns < n_src; ++ns) acc += spec[k + (int64_t)K * ns] * dist[x + P * ns] / ni;
    out[x + P * b] = acc;
}

void launch_general_source(const double* spec, int K, int n_src, const double* dist, int64_t P, const int* k_lo,
                           const int* k_hi, int nb, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_general_source, dim3((unsigned)((P + 255) / 256), (unsigned)nb), dim3(256), 0, s, spec, K,
                       n_src, dist, P, k_lo, k_hi, out);
}

// ---------------------------------------------------------------------------------------------
// Parameter map (map_m_to_parameters.m:48-49, map_gradparameters_to_gradm.m:101) and the Tikhonov
// terms of inversion/addons/regularization.m:1-29
// ---------------------------------------------------------------------------------------------
__global__ void k_affine(double* __restrict__ x, int64_t n, double a, double b)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = a * (b + x[i]);
}

void launch_affine(double* x, int64_t n, double a, double b, hipStream_t s)
{
    const int64_t nb = (n + 255) / 256;
    if (n) hipLaunchKernelGGL(k_affine, dim3((unsigned)(nb < 65536 ? nb : 65536)), dim3(256), 0, s, x, n, a, b);
}

// Reduces a (s, c) compensated pair per thread over the block into lds; thread 0 holds the result.
__device__ void block_pair(CSum& acc, double* ls, double* lc)
{
    ls[threadIdx.x] = acc.s;
    lc[threadIdx.x] = acc.c;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
            CSum b;
            b.s = ls[threadIdx.x];
            b.c = lc[threadIdx.x];
            b.add(ls[threadIdx.x + off]);
            b.c += lc[threadIdx.x + off];
            ls[threadIdx.x] = b.s;
            lc[threadIdx.x] = b.c;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void k_reg_partial(const double* __restrict__ m, const double* __restrict__ m0,
                                                     const double* __restrict__ w, int64_t n_w, int64_t n,
                                                     int64_t n_src, double alpha, double beta, double* __restrict__ g,
                                                     double* __restrict__ part)
{
    __shared__ double ls[256], lc[256];
    CSum a0, a1;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        const double d = m[i] - m0[i];
        const double wi = w[n_w == 1 ? 0 : i % n_w];
        const double t = wi * (d * d);
        if (i < n_src) {
            a0.add(t);
            if (g) g[i] = g[i] + alpha * wi * d;
        } else {
            a1.add(t);
            if (g) g[i] = g[i] + beta * wi * d;
        }
    }
    block_pair(a0, ls, lc);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x] = ls[0];
        part[4 * blockIdx.x + 1] = lc[0];
    }
    __syncthreads();
    block_pair(a1, ls, lc);
    if (threadIdx.x == 0) {
        part[4 * blockIdx.x + 2] = ls[0];
        part[4 * blockIdx.x + 3] = lc[0];
    }
}

__global__ __launch_bounds__(256) void k_reg_final(const double* __restrict__ part, int nb, double* __restrict__ sums)
{
    __shared__ double ls[256], lc[256];
    for (int q = 0; q < 2; ++q) {
        CSum acc;
        for (int i = threadIdx.x; i < nb; i += 256) {
            acc.add(part[4 * i + 2 * q]);
            acc.c += part[4 * i + 2 * q + 1];
        }
        block_pair(acc, ls, lc);
        if (threadIdx.x == 0) sums[q] = ls[0] + lc[0];
        __syncthreads();
    }
}

void launch_regularization(const double* m, const double* m0, const double* w, int64_t n_w, int64_t n,
                           int64_t n_src, double alpha, double beta, double* g, double* part, int nb,
                           double* sums, hipStream_t s)
{
    hipLaunchKernelGGL(k_reg_partial, dim3((unsigned)nb), dim3(256), 0, s, m, m0, w, n_w, n, n_src, alpha, beta, g,
                       part);
    hipLaunchKernelGGL(k_reg_final, dim3(1), dim3(256), 0, s, part, nb, sums);
}

}  // namespace fd2d

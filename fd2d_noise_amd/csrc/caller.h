// caller.h -- FP64 device kernels of the solvers' callers (measurement, L-BFGS, parameter map).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/fd2d_noise.h"  // enum fd2d_measurement

namespace fd2d {

struct AdjArgs {
    const double* u;     // nr x nt synthetics (MATLAB column-major)
    const double* u0;    // nr x nt observations
    const double* du;    // nr x nt perturbation (2nd order), or null
    const double* t;     // nt times
    int64_t nr, nt;
    double dt;
    int veldis;          // 1: 'vel'
    int measurement;     // FD2D_* above
    int deriv_order;     // 1 or 2
    double left, right;  // causal window [left, right]
    double* misfit;      // nr
    double* adstf;       // nr x nt
    double* scratch;     // 3 * nr * nt
};

void launch_adjoint_sources(const AdjArgs& a, hipStream_t s);
// *out = sum x.*y (compensated FP64); part holds 2*nb doubles
void launch_dot(const double* x, const double* y, int64_t n, double* part, int nb, double* out, hipStream_t s);
// out = alpha*x + beta*y (FP32)
void launch_axpby_f32(float alpha, const float* x, float beta, const float* y, float* out, int64_t n, hipStream_t s);
// out = alpha*x + beta*y
void launch_axpby(double alpha, const double* x, double beta, const double* y, double* out, int64_t n, hipStream_t s);
// GY = gaussblur2d(Y, x1, x2, [s1 s2]) on an n1 x n2 block (leading dimensions ldy, ldg); tmp n1*n2
void launch_gaussblur(const double* Y, int ldy, double* GY, int ldg, double* tmp, const double* x1,
                      const double* x2, int n1, int n2, double s1, double s2, hipStream_t s);

// general_source.m:13-24: out[b][x] = sum_{k in band b} sum_ns spec[k + K*ns] * dist[ns][x] / n_b
// (band b = k_lo[b]..k_hi[b], 0-based inclusive)
void launch_general_source(const double* spec, int K, int n_src, const double* dist, int64_t P, const int* k_lo,
                           const int* k_hi, int nb, double* out, hipStream_t s);
// x = a * (b + x)
void launch_affine(double* x, int64_t n, double a, double b, hipStream_t s);
// sums[0] = sum_{i < n_src} w(i) (m - m0)^2, sums[1] = the same over i >= n_src, with w(i) = w[i % n_w];
// g (if not null) += (i < n_src ? alpha : beta) * w(i) (m - m0). part holds 4*nb doubles.
void launch_regularization(const double* m, const double* m0, const double* w, int64_t n_w, int64_t n,
                           int64_t n_src, double alpha, double beta, double* g, double* part, int nb,
                           double* sums, hipStream_t s);

}  // namespace fd2d

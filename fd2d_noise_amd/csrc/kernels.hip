// kernels.hip -- gfx950 kernels of the 2-D SH staggered-grid engine.
//
// One time step is two launches over the nx x nz grid, 64 x 4 nodes per workgroup (one wavefront
// per z row, so every load and store of a row is 256 contiguous bytes):
//   velocity phase: DS = div(sigma) (zero frames of code/propagation/div_s.m), point sources,
//                   synthesised distributed source, v = (v + dt/rho*DS) * taper, u += v*dt,
//                   snapshot of the pre-update field, DFT sample store, receiver gather;
//   stress phase:   sigma += dt*mu * D(v) (zero frames of dx_v.m / dz_v.m), Born terms, and in
//                   the adjoint run the structure-kernel product of D(u) with D(u_fwd).
// The frequency-domain work is blocked over time: DFT samples and synthesised source slices are
// staged as FP32 slabs and contracted with FP64 phase tables in separate launches.
#include "engine.h"

namespace fd2d {

// ---------------------------------------------------------------------------------------------
// Velocity phase
// ---------------------------------------------------------------------------------------------
template <int R>
__global__ __launch_bounds__(kBX* kBZ) void k_velocity(StepArgs a)
{
    const int i0 = blockIdx.x * kBX, j0 = blockIdx.y * kBZ;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const long p = (long)j * a.pitch + i;
    const long c = (long)j * a.nx + i;

    float ds = 0.f;
    if (i >= R && i <= a.nx - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (a.sxy[p + m - 1] - a.sxy[p - m]);
        ds = s;
    }
    if (j >= R && j <= a.nz - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m)
            s += a.cz[m - 1] * (a.szy[p + (long)(m - 1) * a.pitch] - a.szy[p - (long)m * a.pitch]);
        ds += s;
    }
    // point sources: the tile test is uniform, the node test runs only in the owning tile
    for (int q = 0; q < a.nps; ++q) {
        const int2 ij = a.ps_ij[q];
        if (ij.x >= i0 && ij.x < i0 + kBX && ij.y >= j0 && ij.y < j0 + kBZ) {
            if (ij.x == i && ij.y == j) ds += a.ps_amp ? a.ps_amp[q] : a.ps_const;
        }
    }
    if (a.syn) ds += a.syn[c];

    const float vo = a.v[p];
    const float vn = (vo + a.bv[p] * ds) * (a.ax[i] * a.az[j]);
    a.v[p] = vn;
    float out = vn;  // recorded and sampled field: v in the Green run, u otherwise
    if (a.u) {
        const float uo = a.u[p];
        if (a.snap) a.snap[c] = uo;
        const float un = uo + vn * a.dt;
        a.u[p] = un;
        out = un;
    } else if (a.snap) {
        a.snap[c] = vo;
    }
    if (a.dsamp) a.dsamp[c] = out;
    for (int q = 0; q < a.nrec; ++q) {
        const int2 ij = a.rec_ij[q];
        if (ij.x >= i0 && ij.x < i0 + kBX && ij.y >= j0 && ij.y < j0 + kBZ) {
            if (ij.x == i && ij.y == j) a.seis_col[q] = (double)out;
        }
    }
}

// ---------------------------------------------------------------------------------------------
// Stress phase
// ---------------------------------------------------------------------------------------------
template <int R>
__device__ __forceinline__ float dx_pad(const float* f, long p, const float* cx)
{
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += cx[m - 1] * (f[p + m] - f[p - m + 1]);
    return s;
}

template <int R>
__device__ __forceinline__ float dz_pad(const float* f, long p, long pitch, const float* cz)
{
    float s = 0.f;
#pragma unroll
    for (int m = 1; m <= R; ++m) s += cz[m - 1] * (f[p + (long)m * pitch] - f[p - (long)(m - 1) * pitch]);
    return s;
}

template <int R>
__global__ __launch_bounds__(kBX* kBZ) void k_stress(StepArgs a)
{
    const int i = blockIdx.x * kBX + threadIdx.x, j = blockIdx.y * kBZ + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const long p = (long)j * a.pitch + i;
    const long c = (long)j * a.nx + i;
    const long nxl = a.nx;
    // dx_v is set for i = R-1 .. nx-R-1 (0-based), dz_v for j = R-1 .. nz-R-1
    const bool xin = i >= R - 1 && i <= a.nx - R - 1;
    const bool zin = j >= R - 1 && j <= a.nz - R - 1;
    const float dm = a.dtmu[p];
    if (xin) {
        float sn = a.sxy[p] + dm * dx_pad<R>(a.v, p, a.cx);
        if (a.born_dmu) {
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m)
                s += a.cx[m - 1] * ((a.born_in0[c + m] - a.born_in1[c + m]) -
                                    (a.born_in0[c - m + 1] - a.born_in1[c - m + 1]));
            sn -= a.born_dmu[c] * s;
        }
        a.sxy[p] = sn;
    }
    if (zin) {
        float sn = a.szy[p] + dm * dz_pad<R>(a.v, p, a.pitch, a.cz);
        if (a.born_dmu) {
            float s = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m)
                s += a.cz[m - 1] * ((a.born_in0[c + m * nxl] - a.born_in1[c + m * nxl]) -
                                    (a.born_in0[c - (m - 1) * nxl] - a.born_in1[c - (m - 1) * nxl]));
            sn -= a.born_dmu[c] * s;
        }
        a.szy[p] = sn;
    }
    if (a.kmu) {
        double acc = 0.0;
        if (xin) {
            float sf = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m) sf += a.cx[m - 1] * (a.kmu_fwd[c + m] - a.kmu_fwd[c - m + 1]);
            acc += (double)dx_pad<R>(a.u, p, a.cx) * (double)sf;
        }
        if (zin) {
            float sf = 0.f;
#pragma unroll
            for (int m = 1; m <= R; ++m)
                sf += a.cz[m - 1] * (a.kmu_fwd[c + m * nxl] - a.kmu_fwd[c - (m - 1) * nxl]);
            acc += (double)dz_pad<R>(a.u, p, a.pitch, a.cz) * (double)sf;
        }
        if (xin || zin) a.kmu[c] -= acc * a.kmu_scale;
    }
}

template <int R>
__global__ __launch_bounds__(kBX* kBZ) void k_born_green(StepArgs a, const float* dtdmu, const float* gin)
{
    const int i = blockIdx.x * kBX + threadIdx.x, j = blockIdx.y * kBZ + threadIdx.y;
    if (i >= a.nx || j >= a.nz) return;
    const long p = (long)j * a.pitch + i;
    const long c = (long)j * a.nx + i;
    const long nxl = a.nx;
    const float d = dtdmu[c];
    if (i >= R - 1 && i <= a.nx - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cx[m - 1] * (gin[c + m] - gin[c - m + 1]);
        a.sxy[p] += d * s;
    }
    if (j >= R - 1 && j <= a.nz - R - 1) {
        float s = 0.f;
#pragma unroll
        for (int m = 1; m <= R; ++m) s += a.cz[m - 1] * (gin[c + m * nxl] - gin[c - (m - 1) * nxl]);
        a.szy[p] += d * s;
    }
}

static dim3 step_grid(const StepArgs& a) { return dim3((a.nx + kBX - 1) / kBX, (a.nz + kBZ - 1) / kBZ); }

void launch_velocity(int R, const StepArgs& a, hipStream_t s)
{
    const dim3 g = step_grid(a), b(kBX, kBZ);
    switch (R) {
    case 1: hipLaunchKernelGGL(k_velocity<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_velocity<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_velocity<3>, g, b, 0, s, a); break;
    default: hipLaunchKernelGGL(k_velocity<4>, g, b, 0, s, a); break;
    }
}

void launch_stress(int R, const StepArgs& a, hipStream_t s)
{
    const dim3 g = step_grid(a), b(kBX, kBZ);
    switch (R) {
    case 1: hipLaunchKernelGGL(k_stress<1>, g, b, 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_stress<2>, g, b, 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_stress<3>, g, b, 0, s, a); break;
    default: hipLaunchKernelGGL(k_stress<4>, g, b, 0, s, a); break;
    }
}

void launch_born_green(int R, const StepArgs& a, const float* dtdmu, const float* gin, hipStream_t s)
{
    const dim3 g = step_grid(a), b(kBX, kBZ);
    switch (R) {
    case 1: hipLaunchKernelGGL(k_born_green<1>, g, b, 0, s, a, dtdmu, gin); break;
    case 2: hipLaunchKernelGGL(k_born_green<2>, g, b, 0, s, a, dtdmu, gin); break;
    case 3: hipLaunchKernelGGL(k_born_green<3>, g, b, 0, s, a, dtdmu, gin); break;
    default: hipLaunchKernelGGL(k_born_green<4>, g, b, 0, s, a, dtdmu, gin); break;
    }
}

// ---------------------------------------------------------------------------------------------
// Set-up and element-wise kernels
// ---------------------------------------------------------------------------------------------
__global__ void k_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                           int pitch, double dt)
{
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (long)nx * nz) return;
    const int i = (int)(c % nx), j = (int)(c / nx);
    const long p = (long)j * pitch + i;
    bv[p] = (float)(dt / rho[c]);
    dtmu[p] = (float)(dt * mu[c]);
}

void launch_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                     int pitch, double dt, hipStream_t s)
{
    const long n = (long)nx * nz;
    hipLaunchKernelGGL(k_material, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mu, rho, bv, dtmu,
                       nx, nz, pitch, dt);
}

__global__ void k_scale_to_float(const double* in, float* out, size_t n, double scale)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) out[c] = (float)(scale * in[c]);
}

void launch_scale_to_float(const double* in, float* out, size_t n, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_scale_to_float, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, scale);
}

__global__ void k_copy_f64(double* dst, const double* src, size_t n)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) dst[c] = src[c];
}

void launch_copy_f64(double* dst, const double* src, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, dst, src, n);
}

__global__ void k_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nq * nsteps) return;
    const int q = (int)(t % nq);
    const long n = t / nq;
    amp[t] = (float)stf[q + d0 * n].x;
}

void launch_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps, hipStream_t s)
{
    const long n = (long)nq * nsteps;
    hipLaunchKernelGGL(k_stf_real, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, amp, stf, d0, nq, nsteps);
}

// ---------------------------------------------------------------------------------------------
// Frequency domain
// ---------------------------------------------------------------------------------------------
constexpr int kDftKB = 8;      // complex frequencies per wavefront pass
constexpr int kMaxSlab = 128;  // time samples per flush / synthesis block

// One workgroup = 64 nodes x all K frequencies; the 4 wavefronts split the frequencies. The
// node's samples of the block are staged once in LDS; the phase table index is wave-uniform.
__global__ __launch_bounds__(256) void k_dft_flush(double2* __restrict__ acc, const float* __restrict__ samp,
                                                   const double2* __restrict__ coef, size_t P, int K, int cnt)
{
    __shared__ float tile[kMaxSlab * 64];
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t x0 = (size_t)blockIdx.x * 64;
    const size_t x = x0 + lane;
    for (int jj = w; jj < cnt; jj += 4) tile[jj * 64 + lane] = x < P ? samp[(size_t)jj * P + x] : 0.f;
    __syncthreads();
    for (int k0 = w * kDftKB; k0 < K; k0 += 4 * kDftKB) {
        double re[kDftKB], im[kDftKB];
#pragma unroll
        for (int kk = 0; kk < kDftKB; ++kk) re[kk] = im[kk] = 0.0;
        const int nk = K - k0 < kDftKB ? K - k0 : kDftKB;
        for (int jj = 0; jj < cnt; ++jj) {
            const double b = (double)tile[jj * 64 + lane];
            const double2* cr = coef + (size_t)jj * K + k0;
#pragma unroll
            for (int kk = 0; kk < kDftKB; ++kk) {
                if (kk < nk) {
                    const double2 cc = cr[kk];
                    re[kk] = fma(b, cc.x, re[kk]);
                    im[kk] = fma(b, cc.y, im[kk]);
                }
            }
        }
        if (x < P) {
#pragma unroll
            for (int kk = 0; kk < kDftKB; ++kk) {
                if (kk < nk) {
                    double2* d = acc + (size_t)(k0 + kk) * P + x;
                    double2 v = *d;
                    v.x += re[kk];
                    v.y += im[kk];
                    *d = v;
                }
            }
        }
    }
}

void launch_dft_flush(double2* acc, const float* samp, const double2* coef, size_t P, int K, int cnt,
                      hipStream_t s)
{
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_dft_flush, dim3((unsigned)((P + 63) / 64)), dim3(256), 0, s, acc, samp, coef, P, K, cnt);
}

constexpr int kSynJB = 32;  // time slices per pass of the synthesis kernel

__global__ __launch_bounds__(256) void k_synth(float* __restrict__ out, const double2* __restrict__ A,
                                               const double2* __restrict__ coef, size_t P, int K, int cnt)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int j0 = 0; j0 < cnt; j0 += kSynJB) {
        const int nj = cnt - j0 < kSynJB ? cnt - j0 : kSynJB;
        double s[kSynJB];
#pragma unroll
        for (int jj = 0; jj < kSynJB; ++jj) s[jj] = 0.0;
        for (int k = 0; k < K; ++k) {
            const double2 av = A[(size_t)k * P + x];
#pragma unroll
            for (int jj = 0; jj < kSynJB; ++jj) {
                if (jj < nj) {
                    const double2 cc = coef[(size_t)(j0 + jj) * K + k];
                    s[jj] = fma(av.x, cc.x, fma(-av.y, cc.y, s[jj]));
                }
            }
        }
#pragma unroll
        for (int jj = 0; jj < kSynJB; ++jj)
            if (jj < nj) out[(size_t)(j0 + jj) * P + x] = (float)s[jj];
    }
}

void launch_synth(float* out, const double2* A, const double2* coef, size_t P, int K, int cnt, hipStream_t s)
{
    if (cnt <= 0) return;
    hipLaunchKernelGGL(k_synth, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, out, A, coef, P, K, cnt);
}

__global__ void k_source_spectrum(double2* A, const double2* Z, int conj_in, const double* sd, int n_sd,
                                  const double* I, const double* spec_sum, size_t P, int K)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int k = 0; k < K; ++k) {
        double d = 0.0;
        for (int b = 0; b < n_sd; ++b) d += I[k + (size_t)K * b] * sd[x + P * b];
        const double a = spec_sum[k] * d;
        const double2 z = Z[(size_t)k * P + x];
        A[(size_t)k * P + x] = make_double2(a * z.x, conj_in ? -(a * z.y) : a * z.y);
    }
}

void launch_source_spectrum(double2* A, const double2* Z, int conj_in, const double* sd, int n_sd,
                            const double* I, const double* spec_sum, size_t P, int K, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_spectrum, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, A, Z, conj_in, sd,
                       n_sd, I, spec_sum, P, K);
}

__global__ void k_finalize_dft(double2* out, const double2* acc, size_t n, double scale, int conj)
{
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double2 a = acc[c];
    out[c] = make_double2(scale * a.x, conj ? -(scale * a.y) : scale * a.y);
}

void launch_finalize_dft(double2* out, const double2* acc, size_t n, double scale, int conj, hipStream_t s)
{
    hipLaunchKernelGGL(k_finalize_dft, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out, acc, n, scale, conj);
}

__global__ void k_source_kernel(double* Ks, const double2* G, const double2* D, const double* I,
                                const double* spec, size_t P, int K, int nb_cols, double scale)
{
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= P) return;
    for (int b = 0; b < nb_cols; ++b) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            const double2 g = G[(size_t)k * P + x], d = D[(size_t)k * P + x];
            s += I[k + (size_t)K * b] * spec[k] * (g.x * d.x - g.y * d.y);
        }
        Ks[(size_t)b * P + x] = scale * s;
    }
}

void launch_source_kernel(double* Ks, const double2* G, const double2* D, const double* I, const double* spec,
                          size_t P, int K, int nb_cols, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_source_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s, Ks, G, D, I, spec, P,
                       K, nb_cols, scale);
}

}  // namespace fd2d

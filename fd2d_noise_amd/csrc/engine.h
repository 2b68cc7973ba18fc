// engine.h -- device-side data layout and kernel launch interface of the gfx950 SH engine.
//
// Wavefield layout in HBM: every time-stepped field (v, sigma_xy, sigma_zy, u) and every per-node
// coefficient (dt/rho, dt*mu) is an FP32 nx x nz array, x fastest (the MATLAB column-major order,
// so a wavefront reads 64 consecutive x nodes of one z row: 256 contiguous bytes). Each array is
// padded with kGhost zero nodes on every side and a row pitch rounded up to 64 elements, so every
// stencil neighbour of an interior node is an in-bounds load. sigma_xy and sigma_zy share the node
// (i,j) of the MATLAB arrays sxy(i,j) ((nx-1) x nz) and szy(i,j) (nx x (nz-1)).
//
// Snapshots (G_out, C_out, u_out, u_fwd, G_in, C_in, u_in), DFT samples and synthesised source
// slices are compact nx x nz arrays (the MATLAB layout, so they are copied out unchanged).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fd2d {

constexpr int kGhost = 4;    // widest half-stencil (order 8)
constexpr int kBX = 64;      // step-kernel tile: one wavefront per z row
constexpr int kBZ = 4;
constexpr int kMaxPointSources = 1 << 16;

struct PaddedLayout {
    int nx = 0, nz = 0, pitch = 0, rows = 0;
    size_t elems() const { return (size_t)pitch * (size_t)rows; }
    size_t origin() const { return (size_t)kGhost * (size_t)pitch + kGhost; }
    static PaddedLayout make(int nx, int nz)
    {
        PaddedLayout l;
        l.nx = nx;
        l.nz = nz;
        l.pitch = ((nx + 2 * kGhost + 63) / 64) * 64;
        l.rows = nz + 2 * kGhost;
        return l;
    }
};

// Arguments of one time step (velocity phase and stress phase).
struct StepArgs {
    // state (origin pointers into padded arrays)
    float* v;
    float* sxy;
    float* szy;
    float* u;            // displacement (correlation / adjoint runs), else null
    const float* bv;     // dt / rho
    const float* dtmu;   // dt * mu
    const float* ax;     // separable taper, nx values
    const float* az;     // nz values
    int nx, nz, pitch;
    float cx[4], cz[4];  // staggered coefficients / dx, / dz
    float dt;

    // velocity phase
    int nps;                 // point sources
    const int2* ps_ij;
    const float* ps_amp;     // nps amplitudes of this step (null: ps_const for all)
    float ps_const;
    const float* syn;        // synthesised distributed source of this step (compact), or null
    // With u present (correlation / adjoint) the snapshot is u before the update and the DFT
    // sample and receivers take u after it; without u (Green) they take v.
    float* snap;             // snapshot destination (compact) or null
    float* dsamp;            // DFT sample destination (compact) or null
    int nrec;
    const int2* rec_ij;
    double* seis_col;        // seismograms(:, n)

    // stress phase
    const float* born_dmu;   // dmu as FP32 (compact) for C_in / u_in Born terms, or null
    const float* born_in0;   // C_in(:,:,n)
    const float* born_in1;   // C_in(:,:,n+1)
    const float* kmu_fwd;    // forward snapshot for the structure kernel (compact), or null
    double* kmu;             // K_mu accumulator (compact FP64)
    double kmu_scale;        // fw_nth
};

// Launchers (R = half stencil width: 1..4).
void launch_velocity(int R, const StepArgs& a, hipStream_t s);
void launch_stress(int R, const StepArgs& a, hipStream_t s);
// Green-function Born source applied at the start of a step: sigma += (dt*dmu) .* D(G_in slice).
void launch_born_green(int R, const StepArgs& a, const float* dtdmu, const float* gin, hipStream_t s);

// Material coefficients: bv = dt/rho, dtmu = dt*mu into padded arrays; optional dmu conversion.
void launch_material(const double* mu, const double* rho, float* bv, float* dtmu, int nx, int nz,
                     int pitch, double dt, hipStream_t s);
void launch_scale_to_float(const double* in, float* out, size_t n, double scale, hipStream_t s);

// acc[k][x] (+)= sum_j samp[j][x] * coef[j][k]   (FP64 complex accumulators, FP32 samples)
void launch_dft_flush(double2* acc, const float* samp, const double2* coef, size_t P, int K, int cnt,
                      hipStream_t s);
// out[j][x] = sum_k Re(A[k][x] * coef[j][k])      (FP64, rounded to FP32)
void launch_synth(float* out, const double2* A, const double2* coef, size_t P, int K, int cnt,
                  hipStream_t s);
// A[k][x] = spec_sum[k] * dist(x,k) * (conj_in ? conj(Z) : Z)[k][x], dist(x,k) = sum_b I[k,b] sd[b][x]
void launch_source_spectrum(double2* A, const double2* Z, int conj_in, const double* sd, int n_sd,
                            const double* I, const double* spec_sum, size_t P, int K, hipStream_t s);
// out[k][x] = scale * (conj ? conj(acc) : acc)
void launch_finalize_dft(double2* out, const double2* acc, size_t n, double scale, int conj,
                         hipStream_t s);
// Ks[b][x] = scale * sum_k I[k,b] spec[k] Re(conj(G[k][x] * D[k][x]))
void launch_source_kernel(double* Ks, const double2* G, const double2* D, const double* I,
                          const double* spec, size_t P, int K, int nb_cols, double scale,
                          hipStream_t s);
// point-source amplitudes: amp[n][q] = (float)Re(stf(q, n)) from complex nr x N interleaved
void launch_stf_real(float* amp, const double2* stf, int64_t d0, int nq, int nsteps, hipStream_t s);
// Ks/Kmu copy into the MATLAB K output
void launch_copy_f64(double* dst, const double* src, size_t n, hipStream_t s);

}  // namespace fd2d

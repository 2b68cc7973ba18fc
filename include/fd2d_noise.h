/*
 * fd2d_noise.h -- C ABI of the MI355X (gfx950) 2-D SH noise-correlation engine.
 *
 * The three solver entry points replace the three MATLAB-Coder MEX functions that fd2d_noise builds
 * from its solvers (code/mex_functions/compile_all.m:26-34):
 *
 *   run_forward1_green_mex        <- fd2d_run_forward1_green        (code/run_forward1_green.m:1)
 *   run_forward2_correlation_mex  <- fd2d_run_forward2_correlation  (code/run_forward2_correlation.m:1)
 *   run_noise_adjoint_mex         <- fd2d_run_noise_adjoint         (code/run_noise_adjoint.m:1)
 *
 * Positional arguments are the MATLAB ones, in the same order. Each array is a {data, dims[3]}
 * descriptor in MATLAB column-major layout (first index fastest); a NULL descriptor, a NULL data
 * pointer or a zero in dims is MATLAB's []. Element types are the ones the codegen type specs fix
 * (compile_all.m:5-21): real arrays are double, snapshot arrays (G_in, C_in, u_fwd, u_in and the
 * G_out / C_out / u_out outputs) are single, and complex arrays (G_fft, stf, stf_fft) are complex
 * double stored interleaved {re, im} (the MATLAB R2018a interleaved-complex API).
 *
 * The configuration that the MATLAB solvers read inside the call (input_parameters(),
 * input_interferometry() and absorb_specs(), e.g. code/run_forward1_green.m:30,37 and
 * code/propagation/init_absbound.m:9) is passed explicitly as an fd2d_config.
 *
 * Outputs are caller-allocated with the sizes fd2d_query_output_dims() reports. An output whose
 * descriptor is NULL is not returned (the MEX shim passes NULL for outputs beyond nargout).
 *
 * Host entry points take host memory and copy it to and from the device. The *_device entry
 * points take device pointers for every array (inputs and outputs resident in HBM) and run on the
 * given HIP stream (NULL = the null stream); they return after the work has been enqueued and the
 * stream has been synchronised.
 *
 * Every entry point returns FD2D_OK or an error code; fd2d_last_error() gives the message of the
 * calling thread's last failure. The library holds no global mutable state besides per-device
 * HIP resources, so it may be called from several host threads.
 */
#ifndef FD2D_NOISE_H
#define FD2D_NOISE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum fd2d_status {
    FD2D_OK = 0,
    FD2D_EINVAL = 1,        /* inconsistent sizes or arguments MATLAB would reject */
    FD2D_ENOMEM = 2,        /* device or host allocation failed */
    FD2D_EDEVICE = 3,       /* HIP runtime error or no usable gfx950 device */
    FD2D_EUNSUPPORTED = 4   /* FD order not in {2,4,6,8}, or a mode combination that is rejected */
};

/* Which solver fd2d_query_output_dims() describes. */
enum fd2d_solver {
    FD2D_GREEN = 1,
    FD2D_CORRELATION = 2,
    FD2D_ADJOINT = 3
};

/*
 * The configuration tuples of input/input_parameters.m:2, input/input_interferometry.m:2 and
 * input/absorb_specs.m:2 (model_type and source_type are not read by the solvers).
 */
typedef struct fd2d_config {
    /* input_parameters() */
    double  Lx, Lz;          /* domain extent [m]; dx = Lx/(nx-1) (define_computational_domain.m:19) */
    int32_t nx, nz;          /* grid points */
    double  dt;              /* time step [s] */
    int32_t nt;              /* number of Green-function steps; correlation/adjoint runs use 2nt-1 */
    int32_t order;           /* FD order: 2 or 4 as in dx_v.m:20-30, 6 or 8 as textbook extension */
    int32_t n_basis_fct;     /* number of frequency basis functions (0 = none) */
    int32_t fw_nth;          /* store every fw_nth step of the forward field */
    /* input_interferometry() */
    int32_t n_sample;        /* number of frequencies K */
    int32_t freq_samp;       /* DFT / synthesis every freq_samp steps */
    const double* w_sample;  /* K angular frequencies 2*pi*f_sample [rad/s]; borrowed for the call */
    double  dw;              /* w_sample(2) - w_sample(1) */
    /* absorb_specs() */
    double  absorb_width;    /* taper width [m] */
    int32_t absorb_left, absorb_right, absorb_top, absorb_bottom;
} fd2d_config;

/* Input array: MATLAB size(A,1), size(A,2), size(A,3). */
typedef struct fd2d_carray {
    const void* data;
    int64_t dims[3];
} fd2d_carray;

/* Output array, allocated by the caller. */
typedef struct fd2d_marray {
    void* data;
    int64_t dims[3];
} fd2d_marray;

/*
 * [G_fft, G_out, seismograms] = run_forward1_green(mu, rho, src, rec, mode, dmu, G_in)
 * code/run_forward1_green.m:1-201.
 *   mu, rho      nx x nz double (any shape with nx*nz elements, reshaped as at :32-33)
 *   src, rec     ns x 2 / nr x 2 double, metres (nearest grid node, lowest index on ties, :45-60)
 *   mode         0 = no snapshots, else store single(v) every fw_nth step (:133-137)
 *   dmu, G_in    Born mode when dmu is not [] (:140-144): G_in nx x nz x nt single
 *   G_fft        out nx x nz x n_sample complex double (conjugated DFT of v, :176-198)
 *   G_out        out nx x nz x n_fw single (mode != 0), n_fw from fd2d_query_output_dims
 *   seismograms  out nr x nt double (v at the receivers, :187-189)
 */
int fd2d_run_forward1_green(const fd2d_config* cfg,
                            const fd2d_carray* mu, const fd2d_carray* rho,
                            const fd2d_carray* src, const fd2d_carray* rec, double mode,
                            const fd2d_carray* dmu, const fd2d_carray* G_in,
                            fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seismograms);

/*
 * [seismograms, C_out] = run_forward2_correlation(mu, rho, G_fft, spectrum, source_dist, rec,
 *                                                 mode, dmu, C_in)
 * code/run_forward2_correlation.m:1-254.
 *   G_fft        nx x nz x n_sample complex double (output of the Green run)
 *   spectrum     n_sample x n_noise_sources double
 *   source_dist  nx x nz x (n_basis_fct or n_noise_sources) double
 *   dmu, C_in    Born mode (:232-238): C_in nx x nz x (2nt-1) single
 *   seismograms  out nr x (2nt-1) double (u at the receivers: the correlation traces)
 *   C_out        out nx x nz x floor((2nt-1)/fw_nth) single (mode != 0)
 */
int fd2d_run_forward2_correlation(const fd2d_config* cfg,
                                  const fd2d_carray* mu, const fd2d_carray* rho,
                                  const fd2d_carray* G_fft, const fd2d_carray* spectrum,
                                  const fd2d_carray* source_dist, const fd2d_carray* rec,
                                  double mode, const fd2d_carray* dmu, const fd2d_carray* C_in,
                                  fd2d_marray* seismograms, fd2d_marray* C_out);

/*
 * [K, stf_fft, u_out, seismograms] = run_noise_adjoint(mu, rho, u_fwd, stf, src, rec, spectrum,
 *                                                      source_dist, G_fft, mode, dmu, u_in)
 * code/run_noise_adjoint.m:1-368.
 *   u_fwd        nx x nz x n single forward snapshots (structure kernel, :324-331), or []
 *   stf          complex double: nr x (2nt-1) point adjoint sources at src (:249-253), or
 *                nx x nz x n_sample distributed source (:255-268)
 *   spectrum, source_dist, G_fft   as for the correlation run; G_fft enables the source kernel
 *   mode         0, 1 (gradient), 10, 11, 12, 13 (Hessian) as at :16-22
 *   K            out nx x nz x (max(n_basis_fct,1)+1) double: [K_s, K_mu]
 *   stf_fft      out nx x nz x n_sample complex double (modes 1, 12, 13)
 *   u_out        out nx x nz x floor((2nt-1)/fw_nth) single (modes 11, 13)
 *   seismograms  out nr x (2nt-1) double
 * A distributed stf together with mode 1, 12 or 13 makes the reference advance one DFT counter
 * twice per sample (:265, :343); that combination returns FD2D_EUNSUPPORTED.
 */
int fd2d_run_noise_adjoint(const fd2d_config* cfg,
                           const fd2d_carray* mu, const fd2d_carray* rho,
                           const fd2d_carray* u_fwd, const fd2d_carray* stf,
                           const fd2d_carray* src, const fd2d_carray* rec,
                           const fd2d_carray* spectrum, const fd2d_carray* source_dist,
                           const fd2d_carray* G_fft, double mode,
                           const fd2d_carray* dmu, const fd2d_carray* u_in,
                           fd2d_marray* K, fd2d_marray* stf_fft, fd2d_marray* u_out,
                           fd2d_marray* seismograms);

/* Device-resident forms: identical arguments, every data pointer is a device pointer, and the work
 * runs on `stream` (a hipStream_t, NULL for the null stream). No host<->device copies are made. */
int fd2d_run_forward1_green_device(const fd2d_config* cfg,
                                   const fd2d_carray* mu, const fd2d_carray* rho,
                                   const fd2d_carray* src, const fd2d_carray* rec, double mode,
                                   const fd2d_carray* dmu, const fd2d_carray* G_in,
                                   fd2d_marray* G_fft, fd2d_marray* G_out,
                                   fd2d_marray* seismograms, void* stream);

int fd2d_run_forward2_correlation_device(const fd2d_config* cfg,
                                         const fd2d_carray* mu, const fd2d_carray* rho,
                                         const fd2d_carray* G_fft, const fd2d_carray* spectrum,
                                         const fd2d_carray* source_dist, const fd2d_carray* rec,
                                         double mode, const fd2d_carray* dmu,
                                         const fd2d_carray* C_in,
                                         fd2d_marray* seismograms, fd2d_marray* C_out,
                                         void* stream);

int fd2d_run_noise_adjoint_device(const fd2d_config* cfg,
                                  const fd2d_carray* mu, const fd2d_carray* rho,
                                  const fd2d_carray* u_fwd, const fd2d_carray* stf,
                                  const fd2d_carray* src, const fd2d_carray* rec,
                                  const fd2d_carray* spectrum, const fd2d_carray* source_dist,
                                  const fd2d_carray* G_fft, double mode,
                                  const fd2d_carray* dmu, const fd2d_carray* u_in,
                                  fd2d_marray* K, fd2d_marray* stf_fft, fd2d_marray* u_out,
                                  fd2d_marray* seismograms, void* stream);

/*
 * Sizes of the outputs of one solver, in MATLAB output order:
 *   FD2D_GREEN:       G_fft, G_out, seismograms
 *   FD2D_CORRELATION: seismograms, C_out
 *   FD2D_ADJOINT:     K, stf_fft, u_out, seismograms
 * An output that the given mode does not produce gets dims {0,0,0} (the reference returns a
 * scalar single(0.0) there, e.g. code/run_forward1_green.m:109). n_rec is the number of receiver
 * rows. Returns the number of outputs, or a negative fd2d_status on error.
 */
int fd2d_query_output_dims(const fd2d_config* cfg, int32_t solver, double mode, int64_t n_rec,
                           int64_t dims[4][3]);

/*
 * Step-kernel timing (instrumentation; it has no counterpart in the reference). When enabled with
 * every > 0, every every-th time step of each solve brackets its fused step launch with HIP events
 * on the solve's stream; the durations are summed per solver until fd2d_timing_read(), which
 * returns and clears them. every = 0 disables it (the default).
 */
typedef struct fd2d_timing {
    int64_t launches[3][2];  /* [solver-1][0: fused step kernel, 1: reserved (0)] timed launches */
    double ms[3][2];         /* summed kernel durations [ms] */
    int64_t points[3];       /* nx*nz of the last timed solve of each solver */
} fd2d_timing;

int fd2d_timing_enable(int32_t every);
int fd2d_timing_read(fd2d_timing* out);

/* Message of the calling thread's last failure ("" if none). */
const char* fd2d_last_error(void);

/* Library version string. */
const char* fd2d_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FD2D_NOISE_H */

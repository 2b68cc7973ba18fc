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
    FD2D_EUNSUPPORTED = 4,  /* FD order not in {2,4,6,8}, or a mode combination that is rejected */
    FD2D_ENONFINITE = 5     /* the wavefield became NaN or Inf (e.g. dt beyond the stability limit of
                               the stencil); the outputs are still written, as the reference returns
                               its arrays whatever they hold */
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
 * runs on `stream` (a hipStream_t, NULL for the null stream). No host<->device copies of fields are
 * made. The small inputs src, rec and spectrum may also be in host memory: they are then read in
 * place, without the wait for queued work that reading them from device memory implies (the batch
 * engine and serving loops pass them this way). */
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

/* ---------------------------------------------------------------------------------------------
 * Callers of the solvers, FP64 on the device
 * ------------------------------------------------------------------------------------------- */

/* Measurements of tools/measurement/make_adjoint_sources.m:31 (the 'measurement' argument). */
enum fd2d_measurement {
    FD2D_WAVEFORM_DIFFERENCE = 0,   /* waveform_difference.m */
    FD2D_CC_TIME_SHIFT = 1,         /* cc_time_shift.m */
    FD2D_AMPLITUDE_DIFFERENCE = 2,  /* amp_diff.m */
    FD2D_LOG_AMPLITUDE_RATIO = 3    /* log_amp_ratio.m */
};

/*
 * [misfit_n, adstf] = make_adjoint_sources(u, u_0, du, t, veldis, measurement, src, rec, deriv_order)
 * tools/measurement/make_adjoint_sources.m:31-175, FP64 with compensated sums.
 *   u, u0, du    nr x nt double (du may be NULL for deriv_order 1); t: nt times
 *   veldis       0 = 'dis', 1 = 'vel';  deriv_order 1 = '1st', 2 = '2nd'
 *   win_left/right  the causal measurement window; the reference uses [100, 200] s
 *                (make_adjoint_sources.m:85-86), clipped to t(end) as it does, FD2D_EINVAL if
 *                win_left > t(end) ('time series too short')
 *   misfit       out nr (misfit_n);  adstf  out nr x nt (time-reversed adjoint sources)
 * cc_time_shift with deriv_order 2 returns FD2D_EUNSUPPORTED, as the reference errors there.
 */
int fd2d_make_adjoint_sources(const double* u, const double* u0, const double* du, const double* t, int64_t nr,
                              int64_t nt, int32_t veldis, int32_t measurement, int32_t deriv_order,
                              double win_left, double win_right, double* misfit, double* adstf);
int fd2d_make_adjoint_sources_device(const double* u, const double* u0, const double* du, const double* t,
                                     int64_t nr, int64_t nt, int32_t veldis, int32_t measurement,
                                     int32_t deriv_order, double win_left, double win_right, double* misfit,
                                     double* adstf, void* stream);

/* The same, enqueued on `stream` without waiting for it: t_host is the time axis t in host memory
 * (read for the window limits), t stays on the device for the kernels. For callers that queue
 * solves and caller work on one stream (the serving loop). */
int fd2d_make_adjoint_sources_queued(const double* u, const double* u0, const double* du, const double* t,
                                     const double* t_host, int64_t nr, int64_t nt, int32_t veldis,
                                     int32_t measurement, int32_t deriv_order, double win_left, double win_right,
                                     double* misfit, double* adstf, void* stream);

/*
 * fd2d_make_adjoint_sources_queued without any device allocation. It has no counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The kernels' scratch (3 * nr * nt doubles)
 * is a static device array of 2^21 doubles (16 MiB) that each GPU the library runs on holds until the
 * process ends, so nr * nt <= 699050; larger calls return FD2D_EINVAL. The array's address is looked
 * up on the first call on each device. Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (on other streams or from other host threads) race. The kernels write every
 * scratch element before reading it, so nothing is cleared between calls.
 */
int fd2d_make_adjoint_sources_noalloc(const double* u, const double* u0, const double* du, const double* t,
                                      const double* t_host, int64_t nr, int64_t nt, int32_t veldis,
                                      int32_t measurement, int32_t deriv_order, double win_left, double win_right,
                                      double* misfit, double* adstf, void* stream);

/* x'*y of two device FP64 vectors (n = 0: 0, pointers may be NULL) with compensated summation (the inner products of
 * inversion/optlib/optlib_lbfgs.m:51,92,102,110,175 and optlib_wolfe.m:59,85,105); result to host. */
int fd2d_dot_device(const double* x, const double* y, int64_t n, double* result, void* stream);
/* out = alpha*x + beta*y on the device (the model and direction updates of the same files). */
int fd2d_axpby_device(double alpha, const double* x, double beta, const double* y, double* out, int64_t n,
                      void* stream);
/* out = alpha*x + beta*y on FP32 device arrays (the single-precision wavefield sums of
 * inversion/optlib_callback_fct/eval_hessian_vector_product_2.m:91,143-144). */
int fd2d_axpby_f32_device(float alpha, const float* x, float beta, const float* y, float* out, int64_t n,
                          void* stream);
/* GY = gaussblur2d(Y, x1, x2, [sigma1 sigma2]) (tools/addons/gaussblur2d.m:2-38) on an n1 x n2
 * block of device arrays with leading dimensions ldy / ldg; x1, x2 device coordinate vectors. */
int fd2d_gaussblur2d_device(const double* Y, int32_t n1, int32_t n2, int32_t ldy, const double* x1,
                            const double* x2, double sigma1, double sigma2, double* GY, int32_t ldg,
                            void* stream);

/*
 * dist_basis_fct = general_source(noise_spectrum, noise_source_distribution, n_basis_fct)
 * (code/propagation/general_source.m:2-27, called by input/make_noise_source.m:224-227): band
 * averages of spectrum x distribution over the bands of integration_limits(n_sample, n_basis_fct).
 *   spectrum  n_sample x n_src FP64;  dist  nx x nz x n_src (P = nx*nz nodes);  dist_basis out
 *   nx x nz x n_basis_fct. All device arrays.
 */
int fd2d_general_source_device(const double* spectrum, int32_t n_sample, int32_t n_src, const double* dist,
                               int64_t P, int32_t n_basis_fct, double* dist_basis, void* stream);

/*
 * Parameter map of inversion/optlib_callback_fct/map_m_to_parameters.m:2-52 ('ring' switch 'no'):
 * m (nx x nz x n_maps, FP64 device) -> m_parameters. Inside the absorbing frame (rows ix..nx-ix+1,
 * columns iz..nz-iz+1, the first node where init_absbound == 1) every map is blurred with
 * gaussblur2d (sigma_source for maps 1..n_maps-1, sigma_structure for the last); then
 * maps 1..n_maps-1 get ref_source added and the last becomes mu_0 * (ref_structure + map).
 * Geometry and absorbing frame come from cfg (Lx, Lz, nx, nz, absorb_*). In place allowed.
 */
int fd2d_map_m_to_parameters_device(const fd2d_config* cfg, const double* m, int32_t n_maps,
                                    const double sigma_source[2], const double sigma_structure[2],
                                    double ref_source, double ref_structure, double mu_0, double* m_parameters,
                                    void* stream);
/* map_gradparameters_to_gradm.m:53-106: the blur of the same frame applied to grad_parameters,
 * zero outside the frame, the last map times mu_0. */
int fd2d_map_gradparameters_to_gradm_device(const fd2d_config* cfg, const double* grad_parameters, int32_t n_maps,
                                            const double sigma_source[2], const double sigma_structure[2],
                                            double mu_0, double* grad_m, void* stream);
/*
 * [j, g] = regularization(j, g, m, usr_par) (inversion/addons/regularization.m:1-29):
 *   j += alpha/2 sum w (m - m0)^2 over the source maps (the first n - n_struct entries),
 *   j += beta/2 sum w (m - m0)^2 over the structure map (the last n_struct = nx*nz entries),
 *   g += alpha w (m - m0) / beta w (m - m0) on the same parts (skipped when g is NULL, the
 *   reference's scalar g). w has n_w = 1 or nx*nz entries (repeated over the source maps).
 * *j is a host in/out scalar; m, m0, w, g are device arrays.
 */
int fd2d_regularization_device(double* j, double* g, const double* m, const double* m0, int64_t n, int64_t n_struct,
                               const double* w, int64_t n_w, double alpha, double beta, void* stream);

/* ---------------------------------------------------------------------------------------------
 * Batched solves (no counterpart in the reference, which runs one MEX call at a time; they serve
 * the loops that call the three solvers per station and per basis function, e.g.
 * inversion/optlib_callback_fct/eval_noise_objective_and_gradient.m:54-160)
 * ------------------------------------------------------------------------------------------- */

/*
 * One solve: the arguments of one of the three *_device entry points. Inputs a solver does not take
 * are NULL (see the per-solver lists). born_in is G_in, C_in or u_in. out[] are the outputs in
 * MATLAB order, as fd2d_query_output_dims() lists them (Green: G_fft, G_out, seismograms;
 * correlation: seismograms, C_out; adjoint: K, stf_fft, u_out, seismograms). The config, w_sample
 * and every descriptor are copied at submission; the device arrays they point to must stay valid,
 * and the inputs unchanged, until the solve is done.
 */
typedef struct fd2d_solve {
    int32_t solver;                    /* FD2D_GREEN, FD2D_CORRELATION or FD2D_ADJOINT */
    double mode;
    const fd2d_config* cfg;
    const fd2d_carray* mu;
    const fd2d_carray* rho;
    const fd2d_carray* src;            /* Green, adjoint */
    const fd2d_carray* rec;
    const fd2d_carray* G_fft;          /* correlation, adjoint */
    const fd2d_carray* spectrum;       /* correlation, adjoint */
    const fd2d_carray* source_dist;    /* correlation, adjoint */
    const fd2d_carray* stf;            /* adjoint */
    const fd2d_carray* u_fwd;          /* adjoint */
    const fd2d_carray* dmu;
    const fd2d_carray* born_in;
    fd2d_marray* out[4];
} fd2d_solve;

/*
 * An engine advances many solves together: each round advances every active solve by one time
 * step, or by two where its two-step launch applies, with one launch for up to 36 solves of the same
 * FD order and field set (Green runs carry no displacement field, correlation and adjoint runs do),
 * on streams of the engine's own that wait for the caller's stream at each submission (a large class
 * of small solves is shared between two of them, whose launches then run side by side). Each solve's
 * outputs equal those of its own entry point bit for bit. Solves may be submitted between rounds
 * (continuous batching); a solve's outputs are complete once fd2d_engine_done() reports it. An
 * engine is used by one host thread at a time.
 */
typedef struct fd2d_engine fd2d_engine;

/* The stream must outlive the engine. */
int fd2d_engine_create(void* stream, fd2d_engine** engine);
/*
 * Validates and sets up a solve (its device memory is allocated here); *ticket names it. Returns
 * FD2D_ENOMEM, with the engine unchanged, when the solve's buffers do not fit the device memory free
 * now (after the library's own cached blocks have been returned) or the engine holds 1024 solves: it
 * may be submitted again later, e.g. once running solves have completed.
 */
int fd2d_engine_submit(fd2d_engine* engine, const fd2d_solve* solve, int64_t* ticket);
/* Runs up to `rounds` rounds; *n_active (may be NULL) = solves not yet complete afterwards. When no
 * solve is left running it waits for the finished ones, so advancing until *n_active == 0 leaves
 * every output complete. */
int fd2d_engine_advance(fd2d_engine* engine, int32_t rounds, int32_t* n_active);
/* *done = 1 when the solve no longer runs and the work it queued has completed: its outputs are
 * complete, or it was cancelled or failed (fd2d_engine_progress tells which). */
int fd2d_engine_done(const fd2d_engine* engine, int64_t ticket, int32_t* done);

/* States of an engine's solve. */
enum fd2d_solve_state {
    FD2D_SOLVE_RUNNING = 0,    /* steps left to issue */
    FD2D_SOLVE_FINISHING = 1,  /* every step issued; its last work is still queued */
    FD2D_SOLVE_COMPLETE = 2,   /* outputs complete */
    FD2D_SOLVE_CANCELLED = 3,  /* removed by fd2d_engine_cancel; outputs undefined */
    FD2D_SOLVE_FAILED = 4      /* its wavefield became NaN or Inf (FD2D_ENONFINITE); the solves batched
                                  with it are not affected */
};
/*
 * *state of a solve and the time steps it has issued (*steps_issued, *steps_total; either pointer
 * may be NULL). The count is kept on the host: each fd2d_engine_advance round issues one or two
 * steps of every running solve, whatever the device has completed, so it is the same on every
 * machine. A complete solve reports -1 for both (the engine keeps no record of finished solves).
 * A running solve is checked for NaN / Inf every 512 steps and after its last step; a failed one
 * stops issuing steps. (No counterpart in the reference, whose solver calls run to the end.)
 */
int fd2d_engine_progress(const fd2d_engine* engine, int64_t ticket, int32_t* state, int64_t* steps_issued,
                         int64_t* steps_total);
/* Cancels a running solve: it issues no further steps and its outputs are left undefined.
 * *cancelled (may be NULL) = 1 if it was running, 0 if it had already finished. (No counterpart in
 * the reference; it serves callers that abandon a request, e.g. a serving loop.) */
int fd2d_engine_cancel(fd2d_engine* engine, int64_t ticket, int32_t* cancelled);
/* Waits for the stream and frees the engine; unfinished solves are abandoned. */
int fd2d_engine_destroy(fd2d_engine* engine);
/* Runs n solves to completion with one engine. FD2D_ENONFINITE names the solves that became NaN or
 * Inf; every other solve of the batch is complete. */
int fd2d_run_batch_device(const fd2d_solve* solves, int32_t n, void* stream);
/* Upper bound of the device bytes one solve allocates inside the library (the caller's arrays
 * are not counted): the figure an admission policy charges against an HBM budget. */
int fd2d_solve_device_bytes(const fd2d_solve* solve, int64_t* bytes);

/*
 * Device-resident arrays: the model and Green-spectrum handles of a cache or serving tier. A handle
 * owns device memory (on the device current at creation) holding one array in the C ABI's layout
 * (column-major, complex interleaved), so a caller without device pointers of its own -- a MEX
 * shim, a service in another language -- keeps mu and rho, Green spectra (G_fft, the input the
 * correlation and adjoint runs share) and snapshots on the GPU between calls, and passes them to
 * the *_device entry points and fd2d_solve through fd2d_resident_view. MATLAB moves every array
 * through the MEX boundary by value, so these have no counterpart in the reference; they replace
 * the PCIe round trip of a multi-GB G_fft between the Green and correlation runs.
 *   elem_bytes   4 (single), 8 (double) or 16 (complex double); any other value is FD2D_EINVAL
 */
typedef struct fd2d_resident fd2d_resident;
/* A resident copy of host array `host` (copied before the call returns). */
int fd2d_resident_upload(const fd2d_carray* host, int32_t elem_bytes, fd2d_resident** out);
/* A resident array of the given dims, zero-filled: an output of the *_device entry points. */
int fd2d_resident_alloc(const int64_t dims[3], int32_t elem_bytes, fd2d_resident** out);
/* Device descriptors of the array: `in` (may be NULL) for inputs, `out` (may be NULL) for outputs. */
int fd2d_resident_view(const fd2d_resident* r, fd2d_carray* in, fd2d_marray* out);
/* Copies the array into host memory after the work queued on `stream` (NULL: the null stream) has
 * finished; host->dims must have as many elements as the array. */
int fd2d_resident_download(const fd2d_resident* r, fd2d_marray* host, void* stream);
/* Frees the handle and its device memory (waits for the device first). NULL is a no-op. */
int fd2d_resident_free(fd2d_resident* r);

/*
 * Device memory of the solvers comes from a per-device cache: released buffers are kept (up to a
 * quarter of the device memory; when it is full, those released longest ago are freed to make room)
 * and handed out again to later solves, which otherwise pay a hipMalloc / hipFree of multi-GB
 * buffers per solve. The split phase tables of the frequency-domain
 * work are cached too (up to 256 MB per device), so solves of one configuration compute and upload
 * them once. The cached memory is not available to other allocators of the process (e.g.
 * PyTorch's); this call waits for the current device and frees every cached block and every phase
 * table no solve holds. Memory management; it has no counterpart in the reference.
 */
int fd2d_release_cached_memory(void);

/*
 * Steps per launch of the time loop (tuning; it has no counterpart in the reference). 1 launches
 * every step on its own; 2 advances two consecutive time steps per launch (temporal blocking)
 * wherever the two steps need at most one DFT sample, at most one synthesised source slice, no Born
 * term and no structure-kernel product after the first of them; 0 (the default) lets the first
 * solve of each shape (grid, FD order, field set) time both on its own first launches and keeps the
 * faster for that shape. Results are the same in every case, bit for bit.
 * Returns FD2D_EINVAL for other values. Applies to solves started afterwards.
 */
int fd2d_set_steps_per_launch(int32_t steps);

/*
 * Step-kernel timing (instrumentation; it has no counterpart in the reference). When enabled with
 * every > 0, each launch of the time loop that advances a step number divisible by `every` is
 * bracketed with HIP events on the solve's stream; durations and step counts are summed per solver
 * until fd2d_timing_read(), which returns and clears them. every = 0 disables it (the default).
 */
typedef struct fd2d_timing {
    int64_t launches[3][2];  /* [solver-1][0]: timed launches, [1]: time steps those launches advanced */
    double ms[3][2];         /* [solver-1][0]: summed launch durations [ms]; [1] reserved (0) */
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

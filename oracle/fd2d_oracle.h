/*
 * fd2d_oracle.h -- FP64 CPU restatement of fd2d_noise's three time-stepping solvers.
 *
 * TEST INFRASTRUCTURE. This library is the parity checker for the HIP engine and the timed CPU
 * baseline; it is never part of the product path. Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark load it.
 *
 * The solver entry points take exactly the arguments of the C ABI in include/fd2d_noise.h and
 * follow the MATLAB solvers step by step (file:line citations in fd2d_oracle.c). The primitives
 * are exported so that their closed-form properties can be tested on their own.
 */
#ifndef FD2D_ORACLE_H
#define FD2D_ORACLE_H

#include "../include/fd2d_noise.h"

#ifdef __cplusplus
extern "C" {
#endif

int oracle_run_forward1_green(const fd2d_config* cfg,
                              const fd2d_carray* mu, const fd2d_carray* rho,
                              const fd2d_carray* src, const fd2d_carray* rec, double mode,
                              const fd2d_carray* dmu, const fd2d_carray* G_in,
                              fd2d_marray* G_fft, fd2d_marray* G_out, fd2d_marray* seismograms);

int oracle_run_forward2_correlation(const fd2d_config* cfg,
                                    const fd2d_carray* mu, const fd2d_carray* rho,
                                    const fd2d_carray* G_fft, const fd2d_carray* spectrum,
                                    const fd2d_carray* source_dist, const fd2d_carray* rec,
                                    double mode, const fd2d_carray* dmu, const fd2d_carray* C_in,
                                    fd2d_marray* seismograms, fd2d_marray* C_out);

int oracle_run_noise_adjoint(const fd2d_config* cfg,
                             const fd2d_carray* mu, const fd2d_carray* rho,
                             const fd2d_carray* u_fwd, const fd2d_carray* stf,
                             const fd2d_carray* src, const fd2d_carray* rec,
                             const fd2d_carray* spectrum, const fd2d_carray* source_dist,
                             const fd2d_carray* G_fft, double mode,
                             const fd2d_carray* dmu, const fd2d_carray* u_in,
                             fd2d_marray* K, fd2d_marray* stf_fft, fd2d_marray* u_out,
                             fd2d_marray* seismograms);

/* Primitives (code/propagation/): out arrays are (nx-1) x nz, nx x (nz-1) and nx x nz. */
void oracle_dx_v(const double* v, double dx, int nx, int nz, int order, double* out);
void oracle_dz_v(const double* v, double dz, int nx, int nz, int order, double* out);
void oracle_div_s(const double* sxy, const double* szy, double dx, double dz, int nx, int nz,
                  int order, double* out);
void oracle_absbound(const fd2d_config* cfg, double* out);
int oracle_interpol_matrix(int n_sample, int n_basis_fct, double* out /* K x max(nb,1) */);

/* OpenMP threads used by the solvers (1 = serial, the reference's -singleCompThread). */
void oracle_set_threads(int n);
int oracle_get_threads(void);

const char* oracle_last_error(void);

#ifdef __cplusplus
}
#endif

#endif

/*
 * harness.c -- drives the MEX gateway fd2d_noise_amd/mex/fd2d_mex.c (built three times, as
 * mex_green / mex_corr / mex_adj) through the stub mex API of tests/mex/mex.h, the way MATLAB's
 * callers do (code/calculate_data.m:195-198, eval_objective_and_gradient.m:49-104), and compares
 * every output with a direct call of the FP64 oracle on the same arrays. TEST INFRASTRUCTURE.
 *
 *   harness exact   gateway bound to the oracle (CPU build): outputs must be identical
 *   harness gpu     gateway bound to libfd2d_noise_amd.so: relative L2 <= 1e-5 (fields, traces) and
 *                   1e-4 (kernels), as the parity tests
 *
 * Covered: the three solvers at FD orders 4 and 8, nargout 1..max (outputs beyond nargout are not
 * created; outputs a mode does not produce are single(0)), [] arguments (non-Born runs), the
 * distributed and point adjoint sources, class and size errors (fd2d:type, fd2d:args, fd2d:solver)
 * and that no array leaks. Exit status 0 when everything holds.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "fd2d_noise.h"
#include "../../oracle/fd2d_oracle.h"

void mex_green(int, mxArray**, int, const mxArray**);
void mex_corr(int, mxArray**, int, const mxArray**);
void mex_adj(int, mxArray**, int, const mxArray**);

static int g_exact = 1, g_fail = 0;

#define CHECK(cond, ...)                                                                           \
    do {                                                                                           \
        if (!(cond)) {                                                                             \
            printf("FAIL %s:%d: ", __FILE__, __LINE__);                                            \
            printf(__VA_ARGS__);                                                                   \
            printf("\n");                                                                          \
            ++g_fail;                                                                              \
        }                                                                                          \
    } while (0)

/* ---- arrays ---- */
static uint64_t g_rng = 88172645463325252ull;
static double urand(void)
{
    g_rng ^= g_rng << 13;
    g_rng ^= g_rng >> 7;
    g_rng ^= g_rng << 17;
    return (double)(g_rng >> 11) / 9007199254740992.0;
}
static mxArray* dbl(mwSize m, mwSize n, mwSize k)
{
    const mwSize d[3] = {m, n, k};
    return mxCreateNumericArray(3, d, mxDOUBLE_CLASS, mxREAL);
}
static mxArray* empty(mxClassID c) { return mxCreateNumericMatrix(0, 0, c, mxREAL); }
static fd2d_carray desc(const mxArray* a)
{
    fd2d_carray d = {NULL, {0, 0, 1}};
    if (!a || mxIsEmpty(a)) return d;
    const mwSize* s = mxGetDimensions(a);
    d.data = mxGetData(a);
    d.dims[0] = (int64_t)s[0];
    d.dims[1] = mxGetNumberOfDimensions(a) > 1 ? (int64_t)s[1] : 1;
    d.dims[2] = mxGetNumberOfDimensions(a) > 2 ? (int64_t)s[2] : 1;
    return d;
}
static const fd2d_carray* dp(const mxArray* a, fd2d_carray* d)
{
    *d = desc(a);
    return d->data ? d : NULL;
}
static fd2d_marray out_like(const mxArray* a)
{
    fd2d_marray m = {NULL, {0, 0, 1}};
    const size_t n = mxGetNumberOfElements(a);
    const size_t bytes = n * (mxIsSingle(a) ? 4 : 8) * (mxIsComplex(a) ? 2 : 1);
    m.data = calloc(bytes ? bytes : 1, 1);
    const mwSize* s = mxGetDimensions(a);
    m.dims[0] = (int64_t)s[0];
    m.dims[1] = mxGetNumberOfDimensions(a) > 1 ? (int64_t)s[1] : 1;
    m.dims[2] = mxGetNumberOfDimensions(a) > 2 ? (int64_t)s[2] : 1;
    return m;
}

/* relative L2 of the gateway's output against the oracle's (both of a's class) */
static double rel(const mxArray* a, const void* ref)
{
    const size_t n = mxGetNumberOfElements(a) * (mxIsComplex(a) ? 2 : 1);
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        const double x = mxIsSingle(a) ? ((const float*)mxGetData(a))[i] : ((const double*)mxGetData(a))[i];
        const double y = mxIsSingle(a) ? ((const float*)ref)[i] : ((const double*)ref)[i];
        num += (x - y) * (x - y);
        den += y * y;
    }
    return den > 0 ? sqrt(num / den) : sqrt(num);
}
static void same(const char* what, const mxArray* a, const fd2d_marray* ref, double tol)
{
    CHECK(a != NULL, "%s: output not created", what);
    if (!a) return;
    CHECK((int64_t)mxGetNumberOfElements(a) == ref->dims[0] * ref->dims[1] * ref->dims[2], "%s: size", what);
    const double e = rel(a, ref->data);
    printf("  %-28s rel-L2 %.2e\n", what, e);
    CHECK(g_exact ? e == 0.0 : e <= tol, "%s: relative L2 %.3e", what, e);
}
static int is_single_zero(const mxArray* a)
{
    return a && mxIsSingle(a) && mxGetNumberOfElements(a) == 1 && ((const float*)mxGetData(a))[0] == 0.0f;
}
static void destroy(mxArray** a, int n)
{
    for (int i = 0; i < n; ++i) {
        mxDestroyArray(a[i]);
        a[i] = NULL;
    }
}

/* ---- the configuration the gateway fetches ---- */
static fd2d_config g_cfg;
static double g_fs[121], g_ws[121];
static void set_config(int order, int nx, int nz, int nt, int fw_nth)
{
    const int absorb[4] = {1, 1, 1, 1};
    const double L = 1000.0 * (nx - 1), Lz = 1000.0 * (nz - 1);
    for (int k = 0; k < 121; ++k) {
        g_fs[k] = 0.02 + 0.0015 * k;
        g_ws[k] = 2.0 * 3.14159265358979323846 * g_fs[k];
    }
    stub_set_config(L, Lz, nx, nz, 0.09, nt, order, 0, fw_nth, g_fs, 121, 5, 5000.0, absorb);
    memset(&g_cfg, 0, sizeof g_cfg);
    g_cfg.Lx = L;
    g_cfg.Lz = Lz;
    g_cfg.nx = nx;
    g_cfg.nz = nz;
    g_cfg.dt = 0.09;
    g_cfg.nt = nt;
    g_cfg.order = order;
    g_cfg.fw_nth = fw_nth;
    g_cfg.n_sample = 121;
    g_cfg.freq_samp = 5;
    g_cfg.w_sample = g_ws;
    g_cfg.dw = g_ws[1] - g_ws[0];
    g_cfg.absorb_width = 5000.0;
    g_cfg.absorb_left = g_cfg.absorb_right = g_cfg.absorb_top = g_cfg.absorb_bottom = 1;
}

static void run_case(int order)
{
    const int nx = 52, nz = 46, nt = 140, fw = 2;
    set_config(order, nx, nz, nt, fw);
    const int N = 2 * nt - 1;
    printf("order %d, %d x %d, nt %d, fw_nth %d\n", order, nx, nz, nt, fw);
    const long live0 = stub_live_arrays();
    mxArray *mu = dbl(nx, nz, 1), *rho = dbl(nx, nz, 1), *src = dbl(1, 2, 1), *rec = dbl(3, 2, 1);
    for (int i = 0; i < nx * nz; ++i) {
        const double vs = 4000.0 * (1 + 0.05 * (urand() - 0.5));
        ((double*)mxGetData(rho))[i] = 3000.0;
        ((double*)mxGetData(mu))[i] = 3000.0 * vs * vs;
    }
    const double sx[1] = {23e3}, sz[1] = {21e3}, rx[3] = {30e3, 17e3, 28e3}, rz[3] = {25e3, 20e3, 14e3};
    ((double*)mxGetData(src))[0] = sx[0];
    ((double*)mxGetData(src))[1] = sz[0];
    for (int r = 0; r < 3; ++r) {
        ((double*)mxGetData(rec))[r] = rx[r];
        ((double*)mxGetData(rec))[3 + r] = rz[r];
    }
    mxArray *mode0 = mxCreateDoubleScalar(0), *mode1 = mxCreateDoubleScalar(1), *E = empty(mxDOUBLE_CLASS),
            *Es = empty(mxSINGLE_CLASS);
    fd2d_carray A[12];

    /* ---- Green: nargout 1 and 3, mode 1 ---- */
    mxArray* g[3] = {NULL, NULL, NULL};
    const mxArray* gi[7] = {mu, rho, src, rec, mode1, E, Es};
    CHECK(stub_call(mex_green, 1, g, 7, gi) == 0, "green nargout 1: %s %s", stub_error_id(), stub_error_msg());
    CHECK(g[0] && mxIsComplex(g[0]) && mxGetNumberOfDimensions(g[0]) == 3, "green: G_fft nx x nz x K complex");
    destroy(g, 3);
    CHECK(stub_call(mex_green, 3, g, 7, gi) == 0, "green nargout 3: %s %s", stub_error_id(), stub_error_msg());
    fd2d_marray G0 = out_like(g[0]), Go0 = out_like(g[1]), sg0 = out_like(g[2]);
    CHECK(oracle_run_forward1_green(&g_cfg, dp(mu, &A[0]), dp(rho, &A[1]), dp(src, &A[2]), dp(rec, &A[3]), 1.0, NULL,
                                    NULL, &G0, &Go0, &sg0) == 0, "oracle green: %s", oracle_last_error());
    same("green G_fft", g[0], &G0, 1e-5);
    same("green G_out", g[1], &Go0, 1e-5);
    same("green seismograms", g[2], &sg0, 1e-5);
    /* mode 0 with nargout 2: G_out is single(0) */
    mxArray* h[2] = {NULL, NULL};
    const mxArray* hi[7] = {mu, rho, src, rec, mode0, E, Es};
    CHECK(stub_call(mex_green, 2, h, 7, hi) == 0, "green mode 0: %s", stub_error_msg());
    CHECK(is_single_zero(h[1]), "green mode 0: G_out must be single(0)");
    destroy(h, 2);

    /* ---- correlation with the oracle's G_fft ---- */
    mxArray* Gm = mxCreateNumericArray(mxGetNumberOfDimensions(g[0]), mxGetDimensions(g[0]), mxDOUBLE_CLASS, mxCOMPLEX);
    memcpy(mxGetData(Gm), G0.data, mxGetNumberOfElements(g[0]) * 16);
    mxArray *spec = dbl(121, 1, 1), *sd = dbl(nx, nz, 1);
    for (int k = 0; k < 121; ++k) ((double*)mxGetData(spec))[k] = exp(-pow((g_fs[k] - 0.1) / 0.05, 2));
    for (int i = 0; i < nx * nz; ++i) ((double*)mxGetData(sd))[i] = 0.5 + urand();
    mxArray* c[2] = {NULL, NULL};
    const mxArray* ci[9] = {mu, rho, Gm, spec, sd, rec, mode1, E, Es};
    CHECK(stub_call(mex_corr, 2, c, 9, ci) == 0, "correlation: %s %s", stub_error_id(), stub_error_msg());
    fd2d_marray cc0 = out_like(c[0]), C0 = out_like(c[1]);
    CHECK(oracle_run_forward2_correlation(&g_cfg, dp(mu, &A[0]), dp(rho, &A[1]), dp(Gm, &A[2]), dp(spec, &A[3]),
                                          dp(sd, &A[4]), dp(rec, &A[5]), 1.0, NULL, NULL, &cc0, &C0) == 0,
          "oracle correlation: %s", oracle_last_error());
    same("correlation traces", c[0], &cc0, 1e-5);
    same("correlation C_out", c[1], &C0, 1e-5);

    /* ---- adjoint mode 1: point sources at the receivers (complex stf), K and stf_fft ---- */
    const mwSize sdims[2] = {3, (mwSize)N};
    mxArray* stf = mxCreateNumericArray(2, sdims, mxDOUBLE_CLASS, mxCOMPLEX);
    double amax = 1e-300;
    for (int i = 0; i < 3 * N; ++i) amax = fmax(amax, fabs(((double*)cc0.data)[i]));
    for (int r = 0; r < 3; ++r)
        for (int n = 0; n < N; ++n) ((double*)mxGetData(stf))[2 * (r + 3 * n)] = ((double*)cc0.data)[r + 3 * (N - 1 - n)] / amax;
    mxArray* Cm = mxCreateNumericArray(mxGetNumberOfDimensions(c[1]), mxGetDimensions(c[1]), mxSINGLE_CLASS, mxREAL);
    memcpy(mxGetData(Cm), C0.data, mxGetNumberOfElements(c[1]) * 4);
    mxArray* a[4] = {NULL, NULL, NULL, NULL};
    const mxArray* ai[12] = {mu, rho, Cm, stf, rec, rec, E, E, E, mode1, E, Es};
    CHECK(stub_call(mex_adj, 4, a, 12, ai) == 0, "adjoint mode 1: %s %s", stub_error_id(), stub_error_msg());
    fd2d_marray K0 = out_like(a[0]), sf0 = out_like(a[1]), sa0 = out_like(a[3]);
    CHECK(oracle_run_noise_adjoint(&g_cfg, dp(mu, &A[0]), dp(rho, &A[1]), dp(Cm, &A[2]), dp(stf, &A[3]), dp(rec, &A[4]),
                                   dp(rec, &A[5]), NULL, NULL, NULL, 1.0, NULL, NULL, &K0, &sf0, NULL, &sa0) == 0,
          "oracle adjoint: %s", oracle_last_error());
    same("adjoint K", a[0], &K0, 1e-4);
    same("adjoint stf_fft", a[1], &sf0, 1e-5);
    CHECK(is_single_zero(a[2]), "adjoint mode 1: u_out must be single(0)");
    same("adjoint seismograms", a[3], &sa0, 1e-5);

    /* ---- adjoint mode 0: distributed source (stf_fft), half length, against G_out, nargout 1 ---- */
    mxArray* sfm = mxCreateNumericArray(mxGetNumberOfDimensions(a[1]), mxGetDimensions(a[1]), mxDOUBLE_CLASS, mxCOMPLEX);
    memcpy(mxGetData(sfm), sf0.data, mxGetNumberOfElements(a[1]) * 16);
    mxArray* Gom = mxCreateNumericArray(mxGetNumberOfDimensions(g[1]), mxGetDimensions(g[1]), mxSINGLE_CLASS, mxREAL);
    memcpy(mxGetData(Gom), Go0.data, mxGetNumberOfElements(g[1]) * 4);
    mxArray* b[1] = {NULL};
    const mxArray* bi[12] = {mu, rho, Gom, sfm, rec, E, spec, sd, E, mode0, E, Es};
    CHECK(stub_call(mex_adj, 1, b, 12, bi) == 0, "adjoint mode 0: %s %s", stub_error_id(), stub_error_msg());
    fd2d_marray Kd0 = out_like(b[0]);
    CHECK(oracle_run_noise_adjoint(&g_cfg, dp(mu, &A[0]), dp(rho, &A[1]), dp(Gom, &A[2]), dp(sfm, &A[3]), dp(rec, &A[4]),
                                   NULL, dp(spec, &A[6]), dp(sd, &A[7]), NULL, 0.0, NULL, NULL, &Kd0, NULL, NULL, NULL) == 0,
          "oracle adjoint mode 0: %s", oracle_last_error());
    same("adjoint (mode 0) K", b[0], &Kd0, 1e-4);
    destroy(b, 1);

    /* ---- errors ---- */
    mxArray* e[4] = {NULL, NULL, NULL, NULL};
    mxArray* stf_real = dbl(3, N, 1);
    const mxArray* x1[12] = {mu, rho, Cm, stf_real, rec, rec, E, E, E, mode1, E, Es};
    CHECK(stub_call(mex_adj, 2, e, 12, x1) == 1 && !strcmp(stub_error_id(), "fd2d:type"),
          "real stf must raise fd2d:type (got '%s')", stub_error_id());
    mxArray* mu_s = mxCreateNumericMatrix(nx, nz, mxSINGLE_CLASS, mxREAL);
    const mxArray* x2[7] = {mu_s, rho, src, rec, mode0, E, Es};
    CHECK(stub_call(mex_green, 1, e, 7, x2) == 1 && !strcmp(stub_error_id(), "fd2d:type"), "single mu: fd2d:type");
    mxArray* Cd = dbl(nx, nz, 3);
    const mxArray* x3[12] = {mu, rho, Cd, stf, rec, rec, E, E, E, mode1, E, Es};
    CHECK(stub_call(mex_adj, 1, e, 12, x3) == 1 && !strcmp(stub_error_id(), "fd2d:type"), "double u_fwd: fd2d:type");
    mxArray* src3 = dbl(1, 3, 1);
    const mxArray* x4[7] = {mu, rho, src3, rec, mode0, E, Es};
    CHECK(stub_call(mex_green, 1, e, 7, x4) == 1 && !strcmp(stub_error_id(), "fd2d:type"), "1 x 3 src: fd2d:type");
    mxArray* mu_small = dbl(nx - 1, nz, 1);
    const mxArray* x5[7] = {mu_small, rho, src, rec, mode0, E, Es};
    CHECK(stub_call(mex_green, 1, e, 7, x5) == 1 && !strcmp(stub_error_id(), "fd2d:solver"),
          "mu of the wrong size: fd2d:solver (got '%s')", stub_error_id());
    CHECK(strstr(stub_error_msg(), "nx*nz") != NULL, "size error message: %s", stub_error_msg());
    CHECK(stub_call(mex_green, 1, e, 6, gi) == 1 && !strcmp(stub_error_id(), "fd2d:args"), "6 inputs: fd2d:args");
    CHECK(stub_call(mex_green, 4, e, 7, gi) == 1 && !strcmp(stub_error_id(), "fd2d:args"), "4 outputs: fd2d:args");
    mxArray* mode_v = dbl(1, 2, 1);
    const mxArray* x6[7] = {mu, rho, src, rec, mode_v, E, Es};
    CHECK(stub_call(mex_green, 1, e, 7, x6) == 1 && !strcmp(stub_error_id(), "fd2d:type"), "vector mode: fd2d:type");

    mxArray* all[] = {mu, rho, src, rec, mode0, mode1, E, Es, Gm, spec, sd, stf, Cm, sfm, Gom, stf_real, mu_s, Cd,
                      src3, mu_small, mode_v};
    destroy(g, 3);
    destroy(c, 2);
    destroy(a, 4);
    destroy(all, (int)(sizeof all / sizeof all[0]));
    CHECK(stub_live_arrays() == live0, "arrays leaked: %ld", stub_live_arrays() - live0);
    fd2d_marray* outs[] = {&G0, &Go0, &sg0, &cc0, &C0, &K0, &sf0, &sa0, &Kd0};
    for (size_t i = 0; i < sizeof outs / sizeof outs[0]; ++i) free(outs[i]->data);
}

int main(int argc, char** argv)
{
    setvbuf(stdout, NULL, _IONBF, 0); /* a sanitizer's exit report must not swallow the output */
    if (argc > 1 && !strcmp(argv[1], "init-only")) {
        /* the device and its runtime set up through the library, no solve: what a leak report of
           this run lists belongs to the runtime, not to the gateway or the library's solves */
        double r = 1.0;
        const int st = fd2d_dot_device(NULL, NULL, 0, &r, NULL);
        printf("init-only: status %d, result %g\n", st, r);
        return st == FD2D_OK && r == 0.0 ? 0 : 1;
    }
    g_exact = !(argc > 1 && !strcmp(argv[1], "gpu"));
    oracle_set_threads(1);
    run_case(4);
    run_case(8);
    printf(g_fail ? "HARNESS FAILED (%d)\n" : "HARNESS OK\n", g_fail);
    return g_fail ? 1 : 0;
}

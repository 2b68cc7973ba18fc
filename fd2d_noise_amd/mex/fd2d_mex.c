/*
 * fd2d_mex.c -- MATLAB MEX gateway of the gfx950 engine: the drop-in replacement of the three
 * MATLAB-Coder MEX functions that code/mex_functions/compile_all.m:26-34 builds from the solvers.
 *
 *   -DFD2D_SOLVER=1  run_forward1_green_mex(mu, rho, src, rec, mode, dmu, G_in)
 *   -DFD2D_SOLVER=2  run_forward2_correlation_mex(mu, rho, G_fft, spectrum, source_dist, rec, mode, dmu, C_in)
 *   -DFD2D_SOLVER=3  run_noise_adjoint_mex(mu, rho, u_fwd, stf, src, rec, spectrum, source_dist, G_fft, mode,
 *                                          dmu, u_in)
 *
 * Built with `mex -R2018a` (the interleaved complex API: complex arrays are {re, im} pairs, the layout
 * of the C ABI in include/fd2d_noise.h) against libfd2d_noise_amd.so. The MATLAB callers do not change:
 * positional arguments, nargout-dependent outputs and [] semantics are those of the .m solvers; the
 * configuration the .m solvers read internally (input_parameters, input_interferometry, absorb_specs)
 * is fetched with mexCallMATLAB.
 *
 * Argument classes follow the codegen type specs of compile_all.m:5-21 (i1 complex double, i2 and i5
 * double, i3 double n x 2, i4 double scalar, i6 single): a non-empty array of another class or
 * complexity raises fd2d:type instead of being read with the wrong element size. [] (empty of any
 * class) is MATLAB's empty argument. Solver failures raise fd2d:solver with the library's message; a
 * solve whose wavefield became non-finite returns its outputs (as the .m solvers return whatever
 * their arrays hold) with a warning.
 *
 * FD2D_GREEN_FN / FD2D_CORRELATION_FN / FD2D_ADJOINT_FN / FD2D_LAST_ERROR_FN name the solver entry
 * points and their error message (the C ABI's by default); a test harness may map them to another
 * implementation of the same signatures.
 */
#include "mex.h"
#include "fd2d_noise.h"

#include <string.h>

#ifndef FD2D_GREEN_FN
#define FD2D_GREEN_FN fd2d_run_forward1_green
#define FD2D_CORRELATION_FN fd2d_run_forward2_correlation
#define FD2D_ADJOINT_FN fd2d_run_noise_adjoint
#define FD2D_LAST_ERROR_FN fd2d_last_error
#endif

/* argument kinds of compile_all.m:5-21 */
enum { ARG_CDOUBLE = 1, ARG_DOUBLE = 2, ARG_NX2 = 3, ARG_SCALAR = 4, ARG_SINGLE = 6 };

static void check_class(const mxArray* a, int kind, int pos)
{
    if (!a) mexErrMsgIdAndTxt("fd2d:type", "argument %d is missing", pos);
    if (kind == ARG_SCALAR) {
        if (!mxIsDouble(a) || mxIsComplex(a) || mxGetNumberOfElements(a) != 1)
            mexErrMsgIdAndTxt("fd2d:type", "argument %d must be a real double scalar", pos);
        return;
    }
    if (mxIsEmpty(a)) return; /* [] of any class */
    if (kind == ARG_SINGLE) {
        if (!mxIsSingle(a) || mxIsComplex(a))
            mexErrMsgIdAndTxt("fd2d:type", "argument %d must be a real single array", pos);
    } else if (kind == ARG_CDOUBLE) {
        if (!mxIsDouble(a) || !mxIsComplex(a))
            mexErrMsgIdAndTxt("fd2d:type", "argument %d must be a complex double array (wrap it in complex())", pos);
    } else {
        if (!mxIsDouble(a) || mxIsComplex(a))
            mexErrMsgIdAndTxt("fd2d:type", "argument %d must be a real double array", pos);
        if (kind == ARG_NX2 && (mxGetNumberOfDimensions(a) != 2 || mxGetN(a) != 2))
            mexErrMsgIdAndTxt("fd2d:type", "argument %d must be an n x 2 array of positions", pos);
    }
}

/* MATLAB array -> C ABI descriptor after the class check; NULL for []. */
static const fd2d_carray* in_arr(const mxArray* a, int kind, int pos, fd2d_carray* d)
{
    check_class(a, kind, pos);
    d->data = NULL;
    d->dims[0] = d->dims[1] = 0;
    d->dims[2] = 1;
    if (mxIsEmpty(a)) return NULL;
    const mwSize* s = mxGetDimensions(a);
    const mwSize nd = mxGetNumberOfDimensions(a);
    mwSize tail = 1; /* trailing dimensions fold into the third, as reshape does */
    for (mwSize k = 3; k < nd; ++k) tail *= s[k];
    d->data = mxGetData(a);
    d->dims[0] = (int64_t)s[0];
    d->dims[1] = nd > 1 ? (int64_t)s[1] : 1;
    d->dims[2] = nd > 2 ? (int64_t)(s[2] * tail) : 1;
    return d;
}

/* fd2d_config from input_parameters(), input_interferometry(), absorb_specs(); *keep holds
 * w_sample, which the config borrows until the call ends. */
static fd2d_config get_config(mxArray** keep)
{
    fd2d_config c;
    mxArray *p[11], *q[5], *a[5];
    memset(&c, 0, sizeof c);
    mexCallMATLAB(11, p, 0, NULL, "input_parameters");
    mexCallMATLAB(5, q, 0, NULL, "input_interferometry");
    mexCallMATLAB(5, a, 0, NULL, "absorb_specs");
    c.Lx = mxGetScalar(p[0]);
    c.Lz = mxGetScalar(p[1]);
    c.nx = (int32_t)mxGetScalar(p[2]);
    c.nz = (int32_t)mxGetScalar(p[3]);
    c.dt = mxGetScalar(p[4]);
    c.nt = (int32_t)mxGetScalar(p[5]);
    c.order = (int32_t)mxGetScalar(p[6]);
    c.n_basis_fct = (int32_t)mxGetScalar(p[9]);
    c.fw_nth = (int32_t)mxGetScalar(p[10]);
    c.n_sample = (int32_t)mxGetScalar(q[1]);
    c.w_sample = mxGetPr(q[2]);
    c.dw = mxGetScalar(q[3]);
    c.freq_samp = (int32_t)mxGetScalar(q[4]);
    c.absorb_width = mxGetScalar(a[0]);
    c.absorb_left = (int32_t)mxGetScalar(a[1]);
    c.absorb_right = (int32_t)mxGetScalar(a[2]);
    c.absorb_top = (int32_t)mxGetScalar(a[3]);
    c.absorb_bottom = (int32_t)mxGetScalar(a[4]);
    *keep = q[2];
    for (int i = 0; i < 11; ++i) mxDestroyArray(p[i]);
    for (int i = 0; i < 5; ++i) {
        if (i != 2) mxDestroyArray(q[i]);
        mxDestroyArray(a[i]);
    }
    return c;
}

/* Output k: allocated with the dims fd2d_query_output_dims reports, or single(0) as the reference
 * returns when the mode does not produce it (e.g. code/run_forward1_green.m:109); not produced
 * beyond nargout (the first output always is, as MATLAB assigns ans). */
static fd2d_marray* out_arr(mxArray** slot, const int64_t d[3], mxClassID cls, mxComplexity cx, int wanted,
                            fd2d_marray* m)
{
    m->data = NULL;
    m->dims[0] = m->dims[1] = 0;
    m->dims[2] = 1;
    if (!wanted) return NULL;
    if (d[0] * d[1] * d[2] == 0) {
        *slot = mxCreateNumericMatrix(1, 1, mxSINGLE_CLASS, mxREAL);
        return NULL;
    }
    const mwSize s[3] = {(mwSize)d[0], (mwSize)d[1], (mwSize)d[2]};
    *slot = mxCreateNumericArray(3, s, cls, cx);
    m->data = mxGetData(*slot);
    memcpy(m->dims, d, sizeof m->dims);
    return m;
}

static void check(int st, mxArray* keep)
{
    if (st == FD2D_OK) return;
    if (st == FD2D_ENONFINITE) {
        mexWarnMsgIdAndTxt("fd2d:nonfinite", "%s", FD2D_LAST_ERROR_FN());
        return;
    }
    mxDestroyArray(keep);
    mexErrMsgIdAndTxt("fd2d:solver", "%s", FD2D_LAST_ERROR_FN());
}

static int64_t rows_of(const mxArray* a) { return mxIsEmpty(a) ? 0 : (int64_t)mxGetM(a); }

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[])
{
    int64_t d[4][3];
    fd2d_carray A[12];
    fd2d_marray M[4];
    mxArray* keep = NULL;
#if FD2D_SOLVER == 1
    if (nrhs != 7 || nlhs > 3) mexErrMsgIdAndTxt("fd2d:args", "run_forward1_green_mex: 7 inputs and up to 3 outputs");
    const fd2d_carray *mu = in_arr(prhs[0], ARG_DOUBLE, 1, &A[0]), *rho = in_arr(prhs[1], ARG_DOUBLE, 2, &A[1]),
                      *src = in_arr(prhs[2], ARG_NX2, 3, &A[2]), *rec = in_arr(prhs[3], ARG_NX2, 4, &A[3]);
    check_class(prhs[4], ARG_SCALAR, 5);
    const fd2d_carray *dmu = in_arr(prhs[5], ARG_DOUBLE, 6, &A[5]), *gin = in_arr(prhs[6], ARG_SINGLE, 7, &A[6]);
    const double mode = mxGetScalar(prhs[4]);
    fd2d_config c = get_config(&keep);
    if (fd2d_query_output_dims(&c, FD2D_GREEN, mode, rows_of(prhs[3]), d) < 0) check(FD2D_EINVAL, keep);
    fd2d_marray* o0 = out_arr(&plhs[0], d[0], mxDOUBLE_CLASS, mxCOMPLEX, 1, &M[0]);
    fd2d_marray* o1 = out_arr(&plhs[1], d[1], mxSINGLE_CLASS, mxREAL, nlhs > 1, &M[1]);
    fd2d_marray* o2 = out_arr(&plhs[2], d[2], mxDOUBLE_CLASS, mxREAL, nlhs > 2, &M[2]);
    check(FD2D_GREEN_FN(&c, mu, rho, src, rec, mode, dmu, gin, o0, o1, o2), keep);
#elif FD2D_SOLVER == 2
    if (nrhs != 9 || nlhs > 2)
        mexErrMsgIdAndTxt("fd2d:args", "run_forward2_correlation_mex: 9 inputs and up to 2 outputs");
    const fd2d_carray *mu = in_arr(prhs[0], ARG_DOUBLE, 1, &A[0]), *rho = in_arr(prhs[1], ARG_DOUBLE, 2, &A[1]),
                      *G = in_arr(prhs[2], ARG_CDOUBLE, 3, &A[2]), *spec = in_arr(prhs[3], ARG_DOUBLE, 4, &A[3]),
                      *sd = in_arr(prhs[4], ARG_DOUBLE, 5, &A[4]), *rec = in_arr(prhs[5], ARG_NX2, 6, &A[5]);
    check_class(prhs[6], ARG_SCALAR, 7);
    const fd2d_carray *dmu = in_arr(prhs[7], ARG_DOUBLE, 8, &A[7]), *cin = in_arr(prhs[8], ARG_SINGLE, 9, &A[8]);
    const double mode = mxGetScalar(prhs[6]);
    fd2d_config c = get_config(&keep);
    if (fd2d_query_output_dims(&c, FD2D_CORRELATION, mode, rows_of(prhs[5]), d) < 0) check(FD2D_EINVAL, keep);
    fd2d_marray* o0 = out_arr(&plhs[0], d[0], mxDOUBLE_CLASS, mxREAL, 1, &M[0]);
    fd2d_marray* o1 = out_arr(&plhs[1], d[1], mxSINGLE_CLASS, mxREAL, nlhs > 1, &M[1]);
    check(FD2D_CORRELATION_FN(&c, mu, rho, G, spec, sd, rec, mode, dmu, cin, o0, o1), keep);
#else
    if (nrhs != 12 || nlhs > 4) mexErrMsgIdAndTxt("fd2d:args", "run_noise_adjoint_mex: 12 inputs and up to 4 outputs");
    const fd2d_carray *mu = in_arr(prhs[0], ARG_DOUBLE, 1, &A[0]), *rho = in_arr(prhs[1], ARG_DOUBLE, 2, &A[1]),
                      *ufwd = in_arr(prhs[2], ARG_SINGLE, 3, &A[2]), *stf = in_arr(prhs[3], ARG_CDOUBLE, 4, &A[3]),
                      *src = in_arr(prhs[4], ARG_NX2, 5, &A[4]), *rec = in_arr(prhs[5], ARG_NX2, 6, &A[5]),
                      *spec = in_arr(prhs[6], ARG_DOUBLE, 7, &A[6]), *sd = in_arr(prhs[7], ARG_DOUBLE, 8, &A[7]),
                      *G = in_arr(prhs[8], ARG_CDOUBLE, 9, &A[8]);
    check_class(prhs[9], ARG_SCALAR, 10);
    const fd2d_carray *dmu = in_arr(prhs[10], ARG_DOUBLE, 11, &A[10]), *uin = in_arr(prhs[11], ARG_SINGLE, 12, &A[11]);
    const double mode = mxGetScalar(prhs[9]);
    fd2d_config c = get_config(&keep);
    if (fd2d_query_output_dims(&c, FD2D_ADJOINT, mode, rows_of(prhs[5]), d) < 0) check(FD2D_EINVAL, keep);
    fd2d_marray* o0 = out_arr(&plhs[0], d[0], mxDOUBLE_CLASS, mxREAL, 1, &M[0]);
    fd2d_marray* o1 = out_arr(&plhs[1], d[1], mxDOUBLE_CLASS, mxCOMPLEX, nlhs > 1, &M[1]);
    fd2d_marray* o2 = out_arr(&plhs[2], d[2], mxSINGLE_CLASS, mxREAL, nlhs > 2, &M[2]);
    fd2d_marray* o3 = out_arr(&plhs[3], d[3], mxDOUBLE_CLASS, mxREAL, nlhs > 3, &M[3]);
    check(FD2D_ADJOINT_FN(&c, mu, rho, ufwd, stf, src, rec, spec, sd, G, mode, dmu, uin, o0, o1, o2, o3), keep);
#endif
    mxDestroyArray(keep);
}

/* mex_stub.c -- implementation of tests/mex/mex.h (TEST INFRASTRUCTURE; see the header). */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    mxComplexity cx;
    mwSize nd;
    mwSize dims[8];
    void* data;
    int created_in_call; /* destroyed by the stub when the gateway raises */
    struct mxArray_tag* next;
};

static long g_live = 0;
static int g_in_call = 0;
static struct mxArray_tag* g_call_list = NULL;
static jmp_buf g_jmp;
static char g_err_id[128], g_err_msg[1024], g_warn_id[128];

static size_t elem_size(mxClassID c, mxComplexity x) { return (c == mxSINGLE_CLASS ? 4 : 8) * (x == mxCOMPLEX ? 2 : 1); }

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity cx)
{
    mxArray* a = (mxArray*)calloc(1, sizeof *a);
    mwSize n = 1;
    a->cls = cls;
    a->cx = cx;
    a->nd = ndim < 2 ? 2 : ndim;
    for (mwSize k = 0; k < a->nd; ++k) {
        a->dims[k] = k < ndim ? dims[k] : 1;
        n *= a->dims[k];
    }
    while (a->nd > 2 && a->dims[a->nd - 1] == 1) --a->nd; /* MATLAB drops trailing singletons */
    a->data = n ? calloc(n, elem_size(cls, cx)) : NULL;
    ++g_live;
    if (g_in_call) {
        a->created_in_call = 1;
        a->next = g_call_list;
        g_call_list = a;
    }
    return a;
}

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity cx)
{
    const mwSize d[2] = {m, n};
    return mxCreateNumericArray(2, d, cls, cx);
}

mxArray* mxCreateDoubleScalar(double v)
{
    mxArray* a = mxCreateNumericMatrix(1, 1, mxDOUBLE_CLASS, mxREAL);
    *(double*)a->data = v;
    return a;
}

static void unlink_call(mxArray* a)
{
    for (mxArray** p = &g_call_list; *p; p = &(*p)->next)
        if (*p == a) {
            *p = a->next;
            return;
        }
}

void mxDestroyArray(mxArray* a)
{
    if (!a) return;
    if (a->created_in_call) unlink_call(a);
    free(a->data);
    free(a);
    --g_live;
}

void* mxGetData(const mxArray* a) { return a->data; }
double* mxGetPr(const mxArray* a) { return a->cls == mxDOUBLE_CLASS ? (double*)a->data : NULL; }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims; }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->nd; }
size_t mxGetM(const mxArray* a) { return a->dims[0]; }
size_t mxGetN(const mxArray* a)
{
    size_t n = 1;
    for (mwSize k = 1; k < a->nd; ++k) n *= a->dims[k];
    return n;
}
size_t mxGetNumberOfElements(const mxArray* a) { return a->dims[0] * mxGetN(a); }
bool mxIsEmpty(const mxArray* a) { return mxGetNumberOfElements(a) == 0; }
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
bool mxIsComplex(const mxArray* a) { return a->cx == mxCOMPLEX; }
double mxGetScalar(const mxArray* a)
{
    if (mxIsEmpty(a)) return 0.0;
    return a->cls == mxSINGLE_CLASS ? (double)*(const float*)a->data : *(const double*)a->data;
}
mxClassID stub_class(const mxArray* a) { return a->cls; }

/* ---- configuration ---- */
static double g_p[11], g_q_scalar[3], g_abs[5];
static double* g_fs = NULL;
static int g_K = 0;

void stub_set_config(double Lx, double Lz, int nx, int nz, double dt, int nt, int order, int n_basis_fct,
                     int fw_nth, const double* f_sample, int n_sample, int freq_samp, double absorb_width,
                     const int absorb[4])
{
    const double p[11] = {Lx, Lz, nx, nz, dt, nt, order, 0.0, 0.0, n_basis_fct, fw_nth};
    memcpy(g_p, p, sizeof p);
    free(g_fs);
    g_fs = (double*)malloc(sizeof(double) * (size_t)(n_sample > 0 ? n_sample : 1));
    memcpy(g_fs, f_sample, sizeof(double) * (size_t)n_sample);
    g_K = n_sample;
    g_q_scalar[0] = n_sample;
    /* dw = w_sample(2) - w_sample(1) with w = 2 pi f (input/input_interferometry.m) */
    g_q_scalar[1] = n_sample > 1 ? 2.0 * 3.14159265358979323846 * f_sample[1] - 2.0 * 3.14159265358979323846 * f_sample[0]
                                 : 0.0;
    g_q_scalar[2] = freq_samp;
    g_abs[0] = absorb_width;
    for (int k = 0; k < 4; ++k) g_abs[k + 1] = absorb[k];
}

static mxArray* row(const double* v, int n)
{
    mxArray* a = mxCreateNumericMatrix(1, (mwSize)n, mxDOUBLE_CLASS, mxREAL);
    memcpy(a->data, v, sizeof(double) * (size_t)n);
    return a;
}

int mexCallMATLAB(int nlhs, mxArray* plhs[], int nrhs, mxArray* prhs[], const char* name)
{
    (void)nrhs;
    (void)prhs;
    if (!strcmp(name, "input_parameters") && nlhs <= 11) {
        for (int i = 0; i < nlhs; ++i) plhs[i] = mxCreateDoubleScalar(g_p[i]);
    } else if (!strcmp(name, "input_interferometry") && nlhs <= 5) {
        for (int i = 0; i < nlhs; ++i) {
            if (i == 0) {
                plhs[i] = row(g_fs, g_K);
            } else if (i == 2) {
                plhs[i] = row(g_fs, g_K);
                for (int k = 0; k < g_K; ++k) ((double*)plhs[i]->data)[k] *= 2.0 * 3.14159265358979323846;
            } else {
                plhs[i] = mxCreateDoubleScalar(g_q_scalar[i == 1 ? 0 : i == 3 ? 1 : 2]);
            }
        }
    } else if (!strcmp(name, "absorb_specs") && nlhs <= 5) {
        for (int i = 0; i < nlhs; ++i) plhs[i] = mxCreateDoubleScalar(g_abs[i]);
    } else {
        mexErrMsgIdAndTxt("MATLAB:UndefinedFunction", "undefined function %s", name);
    }
    return 0;
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    snprintf(g_err_id, sizeof g_err_id, "%s", id);
    vsnprintf(g_err_msg, sizeof g_err_msg, fmt, ap);
    va_end(ap);
    longjmp(g_jmp, 1);
}

void mexWarnMsgIdAndTxt(const char* id, const char* fmt, ...)
{
    (void)fmt;
    snprintf(g_warn_id, sizeof g_warn_id, "%s", id);
}

int stub_call(mex_gateway fn, int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[])
{
    g_err_id[0] = g_err_msg[0] = g_warn_id[0] = 0;
    for (int i = 0; i < nlhs; ++i) plhs[i] = NULL;
    g_in_call = 1;
    g_call_list = NULL;
    if (setjmp(g_jmp)) {
        /* MATLAB frees what the gateway allocated when it raises */
        g_in_call = 0;
        while (g_call_list) mxDestroyArray(g_call_list);
        for (int i = 0; i < nlhs; ++i) plhs[i] = NULL;
        return 1;
    }
    fn(nlhs, plhs, nrhs, prhs);
    g_in_call = 0;
    /* arrays that survive the call are the outputs: no longer the call's */
    while (g_call_list) {
        mxArray* a = g_call_list;
        g_call_list = a->next;
        a->created_in_call = 0;
        a->next = NULL;
    }
    return 0;
}

const char* stub_error_id(void) { return g_err_id; }
const char* stub_error_msg(void) { return g_err_msg; }
const char* stub_warning_id(void) { return g_warn_id; }
long stub_live_arrays(void) { return g_live; }
